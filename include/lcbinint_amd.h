/*
 * lcbinint_amd.h - C ABI of the MI355X (gfx950) binary-lens magnification library.
 *
 * Plain C: pointers, sizes and POD structs only.  Entry points whose names end
 * in _device take device pointers and a HIP stream (passed as void*, NULL = the
 * default stream) and return once the work is enqueued; the others take host
 * pointers and return when the results are in host memory.
 *
 * Status codes keep the values of the reference's lcbi_status
 * (include/lcbinint/lcbinint.h:8-13) and add one for device failures.
 */
#ifndef LCBINIINT_AMD_H
#define LCBINIINT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lcbg_status {
    LCBG_OK = 0,
    LCBG_INVALID_ARGUMENT = 1,
    LCBG_NUMERICAL_ERROR = 2,
    LCBG_UNSUPPORTED = 3,
    LCBG_DEVICE_ERROR = 4
} lcbg_status;

/* Finite-source method codes, the order of the reference's FiniteSourceMethod
 * (src/lcbinint/magnification/finite_source_magnifier.hpp:14-21). */
typedef enum lcbg_method {
    LCBG_METHOD_POINT_SOURCE = 0,
    LCBG_METHOD_HEXADECAPOLE = 1,
    LCBG_METHOD_INVERSE_RAY_CARTESIAN = 2,
    LCBG_METHOD_INVERSE_RAY_POLAR = 3
} lcbg_method;

/* One light curve's model, VBM coordinates (the Python LightCurve default).
 * Field meanings are those of lcbi_params (include/lcbinint/lcbinint.h:21-82):
 * t0, tE, u0 (= umin), alpha (= theta, rad), s (= sep), q (mass ratio as the
 * user passes it; the library applies the VBM flip q -> 1/q of
 * python/lcbinint_pybind.cpp:1125), rho, linear/sqrt limb-darkening c, d. */
typedef struct lcbg_curve {
    double t0;
    double tE;
    double u0;
    double alpha;
    double s;
    double q;
    double rho;
    double ld_c;
    double ld_d;
} lcbg_curve;

/* Finite-source options: the fields of lcbi_options
 * (include/lcbinint/lcbinint.h:85-104) that the finite-source path reads.
 * adaptive_source_bins != 0 enables the topology retry of
 * fixed_inverse_ray_binary (src/lcbinint/magnification/
 * finite_source_magnifier.cpp:1108-1121, 3166-3196: a Cartesian epoch whose
 * scan shows >= 16 seeds, >= 100 overlaps, >= 1000 gap repairs and a row-edge
 * jump of >= 1000 cells is scanned again at min(max_source_bins,
 * max(50, 2 source_bins)) bins, for source_bins < 50 and rho < 1e-2); with a
 * positive finite_source_tol or finite_source_reltol it is the adaptive mode:
 * a bounding-box margin of max(point_source_threshold, 60) rho, point-source
 * and hexadecapole thresholds derived from the tolerances (:4679-4790), and
 * the local refinement of the Cartesian area (:3071-3289).  The gradient
 * entry points and lcbg_light_curve_device_noalloc reject
 * adaptive_source_bins != 0 (LCBG_UNSUPPORTED). */
typedef struct lcbg_options {
    int source_bins;              /* nbin, cells per source radius (default 50) */
    int caustic_bins;             /* critical-curve phases (default 1400) */
    int mode;                     /* 1 cartesian, 2 polar, 4 auto (default 4) */
    int polar_source_bins;        /* 0 = source_bins */
    double grid_ratio;            /* default 4.0 */
    double polar_grid_ratio;      /* <= 0 = grid_ratio */
    double point_source_threshold;/* bbox margin in rho (default 20) */
    double hexadecapole_threshold;/* caustic-distance threshold in rho (default 3) */
    double adaptive_hex_threshold;/* hexadecapole tolerance (default 1e-3) */
    double finite_source_tol;     /* absolute tolerance (default 0): converged flag, adaptive mode */
    double finite_source_reltol;  /* relative tolerance (default 0): converged flag, adaptive mode */
    int adaptive_source_bins;     /* 0 fixed source_bins (default), else retry / adaptive refinement */
    int max_source_bins;          /* finest retry lattice and refinement depth bound (default 400) */
} lcbg_options;

/* Per-epoch result record, the fields of lcbi_result
 * (include/lcbinint/lcbinint.h:106-116) plus the finite-source method:
 * magnification, the point-source magnification (binary_mag0 on the static
 * path, the sum over binary_images on the LensModel path), the finite-source
 * magnification (= magnification), the source position the magnifier
 * received, the point-source image count, the lcbg_method, the method's error
 * estimate (FiniteSourceResult::error_estimate: 0 for point source and polar,
 * the derivative or hexadecapole relative error times max(|A|, 1), the
 * Cartesian area error indicator of finite_source_magnifier.cpp:1065-1106 with
 * the high-magnification floor of :2745-2756; in the adaptive mode the
 * refinement's estimate of :3284-3289), the refinement level (1 after a
 * topology retry, the deepest split of the local refinement; 0 without
 * adaptive_source_bins) and the converged flag (:3290-3294: a Cartesian epoch
 * converged iff its error estimate is at most 0.999 (tol + reltol max(|A|, 1))
 * and, in the adaptive mode, it was scanned at 32 or more bins; other methods
 * iff A is finite). */
typedef struct lcbg_result {
    double magnification;
    double point_source_magnification;
    double finite_source_magnification;
    double source_x;
    double source_y;
    int32_t image_count;
    int32_t method;
    double error_estimate;
    int32_t refinement_level;
    int32_t converged;
} lcbg_result;

/* Defaults of lcbi_default_options (src/lcbinint/lcbinint.cpp:36-54). */
lcbg_options lcbg_default_options(void);

const char* lcbg_status_string(lcbg_status status);

/* Point-source magnification of n independent (s, q, x, y) rows, device
 * pointers.  Replaces PointSourceMagnifier::binary_mag0
 * (src/lcbinint/magnification/point_source_magnifier.cpp:567) called in a loop;
 * q is passed straight through (VBM BinaryMag0 convention).  image_count may be
 * NULL. */
lcbg_status lcbg_binary_mag0_device(const double* s, const double* q, const double* x,
                                    const double* y, int64_t n, double* mag,
                                    int32_t* image_count, void* stream);

/* A batch of light-curve models prepared on the device of the calling thread.
 * Creation copies the per-curve records (with cos/sin(alpha) rounded by the
 * host libm, as the reference's std::cos/std::sin are).  The device workspace
 * is allocated by the first finite-source run on the batch (caustic caches,
 * angle tables, side streams, and per device the inverse-ray scratch kept
 * until the process ends), and again, after a device-wide synchronisation,
 * by a run with more epochs than any earlier one; other runs allocate
 * nothing.  lcbg_batch_capture_finite_source allocates before it records.
 * A batch is used by one host thread at a time.  `curves` is host memory. */
typedef struct lcbg_batch lcbg_batch;

lcbg_status lcbg_batch_create(const lcbg_curve* curves, int32_t n_curves,
                              const lcbg_options* options, lcbg_batch** out);
lcbg_status lcbg_batch_destroy(lcbg_batch* batch);

/* Replaces the models of every curve of the batch (host array of n_curves
 * records, as for lcbg_batch_create; per-curve source_bins are kept).  The
 * copy is ordered on `stream`, and the next finite-source run on the batch
 * rebuilds the per-curve caustic caches in-stream, as the reference's
 * magnifier rebuilds its cache when (s, q) change
 * (ensure_binary_caustic_cache, finite_source_magnifier.cpp:4014-4030). */
lcbg_status lcbg_batch_update_curves(lcbg_batch* batch, const lcbg_curve* curves, void* stream);

/* Per-curve inverse-ray resolution: source_bins[c] (host array of n_curves)
 * replaces options.source_bins for curve c; 0 keeps the batch option.  The
 * reference reads one FiniteSourceSettings::source_bins per magnifier
 * (src/lcbinint/magnification/finite_source_magnifier.hpp:24, Options.nbin at
 * python/lcbinint_pybind.cpp:2985-3058); a batch of curves with different
 * nbin runs here in one launch instead of one LightCurve per nbin. */
lcbg_status lcbg_batch_set_source_bins(lcbg_batch* batch, const int32_t* source_bins);

/* Point-source light curves of every curve of the batch in one launch.
 * Curve c owns epochs [offsets[c], offsets[c+1]) of `times` and `mag` (device
 * pointers, offsets has n_curves+1 entries).  Replaces the rho == 0 epoch loop
 * of evaluate_binary_light_curve_numpy (python/lcbinint_pybind.cpp:1120-1139);
 * rho is ignored. */
lcbg_status lcbg_batch_point_source_device(lcbg_batch* batch, const int64_t* offsets,
                                           const double* times, double* mag, void* stream);

/* Host-pointer point-source light curve of one curve (copies in, runs, copies
 * out): the drop-in path behind LightCurve(...)(times, params) with rho == 0. */
lcbg_status lcbg_light_curve_point_source(const lcbg_curve* curve, const double* times,
                                          int64_t n, double* mag);

/* Second-order effects of a curve: annual parallax (piEN, piEE, applied when
 * either is nonzero, as LightCurve(parallax=True) passes them,
 * python/lcbinint_pybind.cpp:2734-2735), the event coordinates (ra, dec in
 * degrees, reference time tfix, 0 = t0) of EventCoordinates
 * (python/lcbinint_pybind.cpp:95-101), and the orbital motion parameters of
 * binary_params_from_dict (:145-179): VBM (w1, w2, w3) = (g1, g2, g3) in
 * 1/day, and for the Keplerian orbit lom_szs, lom_ar (default 1).
 * The fields after lom_ar are the reference's source-trajectory extensions
 * (src/lcbinint/model/trajectory.cpp:287-401), zero when unused:
 * terrestrial parallax from an observatory at (obs_lat, obs_lon) degrees,
 * applied with annual parallax when either is nonzero (lcbi_params obs_lat,
 * obs_lon; LightCurve(terrestrial_parallax=True)); xallarap, the source's
 * own orbit, with xallarap_type = lcbi_xallarap_param_type (0 none, 1
 * angular velocity: xi_1, xi_2, omega_xa, inc_xa, phi_xa; 2 orbital
 * elements: piEN_xa, piEE_xa, period_xa, ecc_xa, peri_xa, inc_xa).
 * Amplitudes of the other xallarap form than xallarap_type make the curve
 * unsupported, as has_unsupported_dynamic_effects does
 * (src/lcbinint/model/lens_model.cpp:15-26). */
typedef struct lcbg_dynamics {
    double piEN;
    double piEE;
    double ra;
    double dec;
    double tfix;
    double g1;
    double g2;
    double g3;
    double lom_szs;
    double lom_ar;
    double obs_lat;
    double obs_lon;
    int32_t xallarap_type;
    int32_t reserved;
    double xi_1;
    double xi_2;
    double omega_xa;
    double inc_xa;
    double phi_xa;
    double piEN_xa;
    double piEE_xa;
    double period_xa;
    double ecc_xa;
    double peri_xa;
} lcbg_dynamics;

/* A triple-lens light curve in VBM coordinates: the trajectory of t0, tE, u0,
 * alpha and the geometry of make_triple_lens_geometry_vbm
 * (src/lcbinint/model/triple_lens_geometry.cpp:36-50): separation s and mass
 * ratio q of the first pair, the third lens at distance sep2 from the first at
 * angle ang, mass ratio q2 (> 0) to the first lens (lcbi_params q, sep, q2,
 * sep2, ang; LightCurve(lens="triple_lens")). */
typedef struct lcbg_triple {
    double t0;
    double tE;
    double u0;
    double alpha;
    double s;
    double q;
    double q2;
    double sep2;
    double ang;
} lcbg_triple;

/* Triple-lens point-source light curves on the device: PointSourceMagnifier::
 * triple_mag0 (src/lcbinint/magnification/point_source_magnifier.cpp:861-882,
 * 929-994) on the static trajectory of LensModel (src/lcbinint/model/
 * lens_model.cpp:90-99, 126-140), one thread per epoch: the degree-10 lens
 * polynomial, its Skowron-Gould roots, 2D Newton polish, physical selection by
 * residual.  curves is a host array of n_curves records; offsets (n_curves + 1),
 * times, mag and image_count (may be NULL) are device arrays; synchronous on
 * `stream`.  LCBG_UNSUPPORTED if a polynomial has fewer than three effective
 * degrees (its epochs are NaN). */
lcbg_status lcbg_triple_point_source_device(const lcbg_triple* curves, int32_t n_curves,
                                            const int64_t* offsets, const double* times, double* mag,
                                            int32_t* image_count, void* stream);

/* Host-pointer form for one curve (copies in, runs, copies out): the drop-in
 * path behind LightCurve(lens="triple_lens")(times, params) with rho == 0.
 * LCBG_NUMERICAL_ERROR if a magnification is not finite. */
lcbg_status lcbg_triple_light_curve_point_source(const lcbg_triple* curve, const double* times, int64_t n,
                                                 double* mag, int32_t* image_count);

/* Orbital motion models, the values of lcbi_orbital_motion_mode
 * (include/lcbinint/lcbinint.h:15-19). */
enum { LCBG_ORBIT_STATIC = 0, LCBG_ORBIT_CIRCULAR = 1, LCBG_ORBIT_KEPLER = 2 };

/* Earth ephemeris for annual parallax: the Horizons vector-table CSV that
 * load_earth_ephemeris reads (src/lcbinint/model/trajectory.cpp:147-183).
 * path NULL reads LCBININT_EARTH_EPHEMERIS, as the reference does
 * (trajectory.cpp:134-145); without a call the table is read from that
 * variable when first needed.  Process-wide.  LCBG_INVALID_ARGUMENT if fewer
 * than two rows were read (parallax epochs then evaluate to NaN). */
lcbg_status lcbg_load_earth_ephemeris(const char* path);

/* Host computation, no device: for each epoch the source position the
 * magnifier receives and the projected separation, xys[3i..3i+2] = (x, y, s),
 * as LensModel::magnification (src/lcbinint/model/lens_model.cpp:84-104,
 * 123-178) derives them: Trajectory::source_position with annual parallax
 * (trajectory.cpp:436-484), orbital_state (orbital_motion.cpp:183-199) and
 * the rotation into the instantaneous binary frame (VBM convention).  Curves
 * with terrestrial parallax or xallarap: LCBG_UNSUPPORTED (the device entry
 * points cover them). */
lcbg_status lcbg_dynamic_sources(const lcbg_curve* curve, const lcbg_dynamics* dyn,
                                 int32_t orbital_motion_mode, const double* times, int64_t n,
                                 double* xys);

/* lcbg_dynamic_sources with the forward-mode derivatives of (x, y, s) with
 * respect to (t0, tE, u0, alpha, s, piEN, piEE, g1, g2):
 * jac[27 i + 9 j + k] = d xys[3 i + j] / d param_k (no counterpart in the
 * reference).  Host computation. */
lcbg_status lcbg_dynamic_sources_jacobian(const lcbg_curve* curve, const lcbg_dynamics* dyn,
                                          int32_t orbital_motion_mode, const double* times, int64_t n,
                                          double* xys, double* jac);

/* lcbg_dynamic_sources for every curve of the batch, on the device: device
 * arrays offsets (n_curves + 1), times, and outputs x, y, s (n_epochs each)
 * and, if jac is not NULL, jac[27 i + 9 j + k] as lcbg_dynamic_sources_jacobian
 * lays it out; dyn is a host array of n_curves records.  Per-curve constants
 * come from the host (its libm, as lcbg_dynamic_sources); per epoch the
 * parallax and the rotation run on the device, which reproduces the host's
 * values bitwise without orbital motion and within a few ulp with it (the
 * device libm's cos, sin and atan2 of the orbit phase).  Enqueued on `stream`
 * and waited for; LCBG_NUMERICAL_ERROR if an orbit is not finite.  The device
 * form of Trajectory::source_position and orbital_state over a batch
 * (trajectory.cpp:436-484, orbital_motion.cpp:183-199). */
lcbg_status lcbg_batch_dynamic_sources_device(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                              int32_t orbital_motion_mode, const int64_t* offsets,
                                              const double* times, int64_t n_epochs, double* x, double* y,
                                              double* s, double* jac, void* stream);

/* The source trajectory of the geometry helper source_trajectory
 * (python/lcbinint_pybind.cpp:1819-1880, source_for_geometry): per epoch the
 * source position, with parallax as Trajectory::source_position and, with
 * orbital motion, the sky-frame position rotated into the binary frame of the
 * epoch (always through (x cos a + y sin a, -x sin a + y cos a), which differs
 * from LensModel's (time - t0) / tE without parallax by rounding only).  Host
 * arrays; evaluated on the device by the trajectory kernel on a one-curve
 * handle kept per host thread (allocated on its first call). */
lcbg_status lcbg_source_trajectory(const lcbg_curve* curve, const lcbg_dynamics* dyn,
                                   int32_t orbital_motion_mode, const double* times, int64_t n, double* x,
                                   double* y);

/* Critical curves (map_to_source_plane = 0) or caustics (1) of the geometry
 * helpers critical_curves / caustics (python/lcbinint_pybind.cpp:1882-1926):
 * the critical-curve quartic at the n_points phases 2 pi i / n_points (the
 * polynomial and solve of the caustic cache), its four roots or their images
 * through the lens equation, tracked into four branches as
 * py_append_tracked_points (:621-658) does (first phase ordered by atan2,
 * then the permutation of least total squared step).  vbm != 0 takes the
 * mass ratio as 1/q (:1911-1912).  x, y: host arrays of 4 * n_points, branch
 * b at [b * n_points, (b + 1) * n_points).  n_points >= 8, else
 * LCBG_INVALID_ARGUMENT (the reference raises invalid_argument).  The roots
 * are solved one thread per phase and tracked by one wave on the device; the
 * call allocates and frees its device buffers. */
lcbg_status lcbg_geometry_branches(double s, double q, int32_t n_points, int32_t map_to_source_plane,
                                   int32_t vbm, double* x, double* y);

/* Light curves of the batch's curves with annual parallax and/or orbital
 * motion (dyn: n_curves records), the LensModel path of
 * evaluate_binary_light_curve_numpy (python/lcbinint_pybind.cpp:1159-1173):
 * per epoch the source position and separation of lcbg_dynamic_sources, the
 * point-source magnification summed over binary_images
 * (point_source_magnifier.cpp:723-778), those images as seed hints, and
 * FiniteSourceMagnifier::binary_mag on the caustic of that epoch's
 * separation.  Host pointers (offsets: n_curves + 1, times/mag/method:
 * n_epochs); per-curve source_bins as set on the batch.  The trajectories are
 * evaluated on the device from per-curve constants the host computes (the
 * same values as lcbg_dynamic_sources: bitwise without orbital motion, within
 * a few ulp with it, whose per-epoch cos, sin and atan2 come from the device
 * libm).  Without orbital motion the separation is constant and every curve
 * keeps one caustic cache, reused by later calls while s and q are unchanged;
 * with it, every epoch has its own caustic cache, in chunks of up to 65536
 * epochs.  Device buffers are kept on the batch and grow only when a call
 * holds more epochs or curves than any before (then allocated once). */
lcbg_status lcbg_batch_light_curves_dynamic(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                            int32_t orbital_motion_mode, const int64_t* offsets,
                                            const double* times, int64_t n_epochs, double* mag,
                                            int32_t* method);

/* lcbg_batch_light_curves_dynamic writing one lcbg_result per epoch (host
 * array of n_epochs), the fields LensModel::magnification fills
 * (src/lcbinint/model/lens_model.cpp:186-234): the source position in the
 * instantaneous binary frame, the point-source magnification and image count
 * of binary_images, the finite-source method, error estimate and converged
 * flag.  Returns LCBG_NUMERICAL_ERROR if a magnification is not finite. */
lcbg_status lcbg_batch_light_curves_dynamic_results(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                                    int32_t orbital_motion_mode, const int64_t* offsets,
                                                    const double* times, int64_t n_epochs,
                                                    lcbg_result* results);

/* lcbg_batch_light_curves_dynamic with, for every epoch, the derivatives of
 * the magnification with respect to (t0, tE, u0, alpha, s, q, rho, u_LD, piEN,
 * piEE, g1, g2): dmag[12 i + k] (host, n_epochs x 12; no counterpart in the
 * reference).  The device trajectory is differentiated in forward mode (x, y
 * and s(t) with respect to t0, tE, u0, alpha, s, piEN, piEE, g1, g2; g3,
 * lom_szs, lom_ar, ra, dec and tfix held fixed), the magnification as in
 * lcbg_batch_finite_source_grad_device with respect to the source position,
 * s, q, rho and u_LD, and the two are chained per epoch.  mag_aa (may be
 * NULL) as there.  The rates in the config-5 parameterization follow by the
 * chain rule: d/d(ds/dt) = d/dg1 / (s 365.25), d/d(dalpha/dt) = d/dg2 / 365.25
 * (with g1 = (ds/dt) / s / 365.25 also moving with s). */
lcbg_status lcbg_batch_light_curves_dynamic_grad(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                                 int32_t orbital_motion_mode, const int64_t* offsets,
                                                 const double* times, int64_t n_epochs, double* mag,
                                                 int32_t* method, double* mag_aa, double* dmag);

/* lcbg_batch_light_curves_dynamic(_grad) on device arrays (offsets, times,
 * mag, method, mag_aa, dmag; dmag NULL: values only, mag_aa may be NULL),
 * enqueued on `stream`; dyn is a host array of n_curves records, from which
 * the host computes the per-curve trajectory constants.  The call waits for
 * the trajectory kernel (a non-finite orbit returns LCBG_NUMERICAL_ERROR
 * before any magnification is enqueued, as lens_model.cpp:117-120 raises);
 * the rest is asynchronous, and non-finite magnifications are left in mag
 * for the caller to test.  offsets are trusted as
 * lcbg_batch_finite_source_device trusts them (no counterpart in the
 * reference: the device form of the LensModel loop). */
lcbg_status lcbg_batch_light_curves_dynamic_device(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                                   int32_t orbital_motion_mode, const int64_t* offsets,
                                                   const double* times, int64_t n_epochs, double* mag,
                                                   int32_t* method, double* mag_aa, double* dmag,
                                                   void* stream);

/* Finite-source light curves of every curve of the batch: the plain method of
 * FiniteSourceMagnifier::binary_mag (finite_source_magnifier.cpp:4633-4843) as
 * the static VBM epoch loop of evaluate_binary_light_curve_numpy
 * (python/lcbinint_pybind.cpp:1141-1158) runs it: point source, hexadecapole,
 * or inverse-ray shooting on the Cartesian or polar lattice with limb
 * darkening.  Device pointers; n_epochs = offsets[n_curves].  method[i]
 * receives the lcbg_method of each epoch.  The first call on a batch, and the
 * first after lcbg_batch_update_curves, builds the per-curve caustic caches
 * in-stream; later calls reuse them and the workspace. */
lcbg_status lcbg_batch_finite_source_device(lcbg_batch* batch, const int64_t* offsets,
                                            const double* times, int64_t n_epochs, double* mag,
                                            int32_t* method, void* stream);

/* Value and gradient of the finite-source light curves (no counterpart in the
 * reference, which has no derivatives).  Runs lcbg_batch_finite_source_device
 * (mag, method as there, the reference's values) and then, for every epoch,
 * the derivatives of the magnification with respect to the curve parameters
 * (t0, tE, u0, alpha, s, q, rho, u_LD), dmag[8*i + k] (device, n_epochs x 8),
 * q and alpha as in lcbg_curve.  Point-source and hexadecapole epochs are
 * differentiated exactly (implicit differentiation of the lens equation per
 * image); inverse-ray epochs through an antialiased quadrature on a lattice
 * anchored at the lens-frame origin (cell side the power of two at or below
 * rho/nbin, the source-disk indicator replaced by the covered fraction of
 * each cell).  mag_aa (may be
 * NULL) receives the value of the differentiated function: equal to mag up to
 * rounding for point-source and hexadecapole epochs, the antialiased
 * quadrature for inverse-ray epochs. */
lcbg_status lcbg_batch_finite_source_grad_device(lcbg_batch* batch, const int64_t* offsets,
                                                 const double* times, int64_t n_epochs,
                                                 double* mag, int32_t* method, double* mag_aa,
                                                 double* dmag, void* stream);

/* chi^2 of each curve against observed fluxes, chi2[c] = sum over its epochs of
 * ((flux - (fs * mag + fb)) / sigma)^2 with mag the finite-source
 * magnification, and its gradient grad[8*c + k] with respect to (t0, tE, u0,
 * alpha, s, q, rho, u_LD) through the derivatives of
 * lcbg_batch_finite_source_grad_device (fs and fb held fixed).  flux and sigma
 * have n_epochs entries; mag, method (and mag_aa, dmag if not NULL) are filled
 * as there.  Device pointers. */
lcbg_status lcbg_batch_chi2_grad_device(lcbg_batch* batch, const int64_t* offsets,
                                        const double* times, int64_t n_epochs, const double* flux,
                                        const double* sigma, double fs, double fb, double* mag,
                                        int32_t* method, double* mag_aa, double* dmag,
                                        double* chi2, double* grad, void* stream);

/* Launch schedule of the batch's inverse-ray stage (no counterpart in the
 * reference, which runs one epoch at a time).  Zero fields take the defaults:
 * resident waves per CU for the seed search and the small-slot Cartesian and
 * polar scans (default: as many as the slots the device's occupancy gives),
 * the inside positions a first-crossing scan probes before the probe farm
 * takes it over (default 16), and the number of leading inverse-ray epochs (in
 * queue order, i.e. largest cost first) whose seeds are searched first and whose
 * area scans start on large slots while the other searches run (default 32).
 * Results do not depend on the schedule. */
typedef struct lcbg_schedule {
    int32_t seed_waves_per_cu;
    int32_t cart_waves_per_cu;
    int32_t polar_waves_per_cu;
    int32_t heavy_budget;
    int32_t lead_items;
} lcbg_schedule;

lcbg_status lcbg_batch_set_schedule(lcbg_batch* batch, const lcbg_schedule* schedule);

/* Work budget of one inverse-ray epoch, in wave steps (a step is one 64-cell
 * chunk of a row, one probe merge or one polar run; 0 = the default 2^22, also
 * the maximum).  An epoch that exceeds it is abandoned: its magnification is
 * NaN and lcbg_batch_last_run_status reports it (flag 4).  The reference has
 * no such limit; the budget bounds how long any one epoch can hold the
 * device. */
lcbg_status lcbg_batch_set_work_budget(lcbg_batch* batch, int64_t steps);
lcbg_status lcbg_batch_get_schedule(const lcbg_batch* batch, lcbg_schedule* schedule);

/* Autotuner: runs lcbg_batch_finite_source_device on the given workload with
 * candidate schedules (one pass of coordinate descent over the fields of
 * lcbg_schedule, each candidate timed over one run with HIP events, at most
 * max_runs runs), keeps the fastest on the batch and writes it to *best and
 * its run time to *best_ms.  mag and method receive the last run's results. */
lcbg_status lcbg_batch_autotune(lcbg_batch* batch, const int64_t* offsets, const double* times,
                                int64_t n_epochs, double* mag, int32_t* method, void* stream,
                                int32_t max_runs, lcbg_schedule* best, float* best_ms);

/* Records one run of lcbg_batch_finite_source_device with these arguments
 * (device pointers, the caustic caches always rebuilt) into a HIP graph kept by
 * the batch, replacing an earlier one; nothing runs.  The reference evaluates
 * epoch by epoch and has no counterpart.  All workspace is allocated here, so
 * a replay allocates nothing.  lcbg_batch_launch_graph replays it on `stream`
 * with the same buffers: a caller updates the curves (lcbg_batch_update_curves
 * on the same stream) and the times in place between replays.  A later run of
 * the batch with more epochs than any before reallocates the workspace and
 * drops the graph (lcbg_batch_launch_graph then returns LCBG_INVALID_ARGUMENT).
 *
 * Queue order: every run keeps each inverse-ray epoch's device time (seed
 * search + area); the next run over the same number of epochs starts the
 * epochs that took longest first, so that none of them starts last.  The
 * first run, and any run over a different epoch count, orders them by the
 * estimate |A_ps| nbin^2.  The order changes no result. */
lcbg_status lcbg_batch_capture_finite_source(lcbg_batch* batch, const int64_t* offsets,
                                             const double* times, int64_t n_epochs, double* mag,
                                             int32_t* method, void* stream);
lcbg_status lcbg_batch_launch_graph(lcbg_batch* batch, void* stream);

/* lcbg_batch_finite_source_device, then waits and reports the device time of
 * its four stages (stage_ms[4]: caustic-cache build (0 when the curves did not
 * change), point source + bounding box, caustic distance + hexadecapole,
 * inverse ray), the queue sizes (queue_counts[2]: epochs past
 * the bounding box, inverse-ray epochs) and the overflow flags (error_flags). */
lcbg_status lcbg_batch_finite_source_profile(lcbg_batch* batch, const int64_t* offsets,
                                             const double* times, int64_t n_epochs, double* mag,
                                             int32_t* method, void* stream, float* stage_ms,
                                             uint32_t* queue_counts, uint32_t* error_flags);

/* Diagnostics with no counterpart in the reference (its closest relative is the
 * LCBININT_AREA_DIAGNOSTICS switch of finite_source_magnifier.cpp): after this
 * call every finite-source run of the batch writes, for each inverse-ray epoch
 * e of an n-epoch run, cost[e] = device wall-clock ticks of its seed search,
 * cost[n + e] = ticks of its area scans (added to the value found there, so
 * the caller zeroes it), cost[2n + e] = quintic probe solves of the seed
 * search, cost[3n + e] = Cartesian rows or polar runs closed and cost[4n + e]
 * = wave steps of the area scans (both added likewise), and for each Cartesian
 * epoch the fields of the reference's LegacyAreaDiagnostics
 * (finite_source_magnifier.cpp:1006-1015): cost[5n + e] processed images,
 * cost[6n + e] fold-guarded images, cost[7n + e] counted rows (boundary_rows),
 * cost[8n + e] gap repairs, cost[9n + e] overlap hits and cost[10n + e] the
 * bits of the double max_jump_cells.  Gradient runs add, per inverse-ray
 * epoch, the ticks of the cell enumeration of its quadrature to cost[11n + e],
 * its row runs to cost[12n + e], its cells to cost[13n + e] and its wave steps
 * to cost[14n + e].  cost is a device pointer with 15 n_epochs entries; other
 * epochs are left untouched.  NULL turns it off.
 * lcbg_wall_clock_rate_khz() gives the tick rate of the current device. */
lcbg_status lcbg_batch_set_epoch_cost(lcbg_batch* batch, int64_t* cost);
int64_t lcbg_wall_clock_rate_khz(void);

/* Work of the inverse-ray stage of the batch's last finite-source run
 * (synchronous): counts[0] quintic probe solves, [1] quartic solves of the
 * critical-source refinement, [2] Cartesian lattice cells the row scan of
 * cartesian_image_area_impl (finite_source_magnifier.cpp:1892-2014) evaluates
 * (its guard count), [3] the same for Jacobian-guarded fold seeds, [4] polar
 * lattice cells (run cells, the two cells ending each run, neighbour-row and
 * start-cell tests), [5] device wall-clock ticks spent in seed searches and
 * [6] in the area scans (summed over epochs), [7] Cartesian rows closed plus
 * polar runs recorded.  Cells the device evaluates beyond a run's end are not
 * counted.  counts has 8 entries. */
lcbg_status lcbg_batch_work_counts(lcbg_batch* batch, uint64_t* counts);

/* Durations (ms, HIP events on the stream each kernel runs on) of the seven
 * kernel launches of the inverse-ray stage of the batch's last profiled run
 * (lcbg_batch_finite_source_profile): [0] seed search (k_ir_seeds), [1]
 * Cartesian and [2] polar area scans of the largest epochs on large slots, [3]
 * polar and [4] Cartesian area scans on small slots, [5] Cartesian and [6]
 * polar retries of epochs that outgrew a small slot.  A diagnostic with no
 * counterpart in the reference.  ms has 7 entries. */
lcbg_status lcbg_batch_ir_kernel_ms(lcbg_batch* batch, float* ms);

/* Stage accessor for one epoch of the batch's last finite-source run (host
 * pointers, synchronous, allocates a few KB; a diagnostic, not for hot loops).
 * With n_in < 0 it runs the device seed search of that epoch (the stage of
 * augmented_image_seeds + augment_seeds_from_caustic_branches,
 * finite_source_magnifier.cpp:1709-1871, 4845-4926) up to seed phase
 * last_phase (0 point-source images, 1 first caustic crossing, 2 branch scan,
 * 3 boundary probes, 4 interior probes, 5 caustic-segment augment; -1 = all)
 * and writes up to
 * `capacity` seeds as (x, y) pairs to `seeds`, *n_out = the seed count (-1 if
 * a capacity was exceeded).  With n_in >= 0 it uses the n_in seeds given.
 * Either way *mag receives the inverse-ray magnification computed from those
 * seeds with `method` (LCBG_METHOD_INVERSE_RAY_CARTESIAN: the row scan of
 * inverse_ray_cartesian_binary_mag, :2470-2768; _POLAR: the flood fill of
 * inverse_ray_polar_boundary_binary, :1156-1319). */
lcbg_status lcbg_batch_debug_epoch(lcbg_batch* batch, int64_t epoch, int32_t method,
                                   int32_t last_phase, double* seeds, int32_t capacity,
                                   int32_t n_in, int32_t* n_out, double* mag);

/* The per-curve critical sources of the 1400-phase seed scan (phase k, root r
 * at out[2*(4k+r)], out[2*(4k+r)+1]: 11200 doubles), which the reference
 * recomputes per inverse-ray epoch with critical_sources_at_phase
 * (finite_source_magnifier.cpp:1622-1641).  Builds the batch's caches if
 * needed; synchronous (diagnostics). */
lcbg_status lcbg_batch_debug_critical_sources(lcbg_batch* batch, int32_t curve, double* out);

/* Host-pointer light curve of one curve with the given options (NULL =
 * defaults): the drop-in path behind LightCurve(...)(times, params).  Returns
 * LCBG_NUMERICAL_ERROR if any magnification is not finite (the reference raises
 * RuntimeError("numerical error"), python/lcbinint_pybind.cpp:1153-1155),
 * including epochs this library abandons at a workspace capacity or work
 * budget; mag[i] of such an epoch is NaN. */
lcbg_status lcbg_light_curve(const lcbg_curve* curve, const lcbg_options* options,
                             const double* times, int64_t n, double* mag, int32_t* method);

/* lcbi_magnification_array (src/lcbinint/lcbinint.cpp:89-121) for the static
 * VBM binary lens: one lcbg_result per epoch (host arrays).  Every epoch is
 * evaluated; the status is that of the first epoch whose magnification is not
 * finite (LCBG_NUMERICAL_ERROR, or LCBG_UNSUPPORTED for a method outside this
 * library), where the reference's loop stops, and *first_failed_epoch (may be
 * NULL) receives that epoch's index, -1 if none. */
lcbg_status lcbg_light_curve_results(const lcbg_curve* curve, const lcbg_options* options,
                                     const double* times, int64_t n, lcbg_result* results,
                                     int64_t* first_failed_epoch);

/* Outcome of the batch's last finite-source run (any of the run entry points
 * or a graph replay): waits for that run, then writes the error flags of
 * epochs abandoned by this library's kernels (error_flags: bit 0 seed list
 * full, 1 Cartesian row records full, 2 polar run records full, 3 a method
 * outside this library, 4 work budget, 5 seed pool full, 6 a Cartesian team
 * wait ran out of time) and the index of the first epoch whose magnification
 * is not finite (-1 if none), where lcbi_magnification_array stops
 * (src/lcbinint/lcbinint.cpp:110-119).  Either pointer may be NULL.  Returns
 * LCBG_OK if every epoch is finite, else LCBG_UNSUPPORTED (flag 3),
 * LCBG_DEVICE_ERROR (flag 6) or LCBG_NUMERICAL_ERROR. */
lcbg_status lcbg_batch_last_run_status(lcbg_batch* batch, uint32_t* error_flags,
                                       int64_t* first_failed_epoch);

/* Per-epoch result records of the batch's last finite-source run (device
 * pointer to n_epochs records; n_epochs must be that run's epoch count),
 * enqueued on `stream`, which must be ordered after the run: the lcbg_result
 * fields, converged with the tolerances of the batch's options.  The run's
 * mag and method buffers are read, so they must still hold its results. */
lcbg_status lcbg_batch_epoch_results_device(lcbg_batch* batch, int64_t n_epochs, lcbg_result* results,
                                            void* stream);

/* No counterpart in the reference: an entry point that allocates nothing, for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  It runs what
 * lcbg_batch_finite_source_device runs on a one-curve batch, for one curve
 * (host record, as for lcbg_batch_create) with options (NULL = defaults):
 * times, mag and method are device pointers of n epochs, and the work is
 * enqueued on `stream` only.  Every piece of workspace lies in one static
 * device array of 128 MiB, which each device that runs this entry point
 * holds until the process ends; no device memory is allocated or freed.
 * Limits: n <= 8192 (else LCBG_INVALID_ARGUMENT), caustic_bins <= 2048
 * (else LCBG_INVALID_ARGUMENT), 2^17 Cartesian rows per image and 2^20 polar
 * runs per epoch (an epoch past them, like one past the work budget, gets a
 * NaN magnification; there is no batch to query, so callers test mag).
 * The first call on a device, and a call whose caustic_bins differs from the
 * previous call's on that device, upload a small angle table synchronously:
 * a caller that captures the call into a HIP graph makes one call with the
 * same caustic_bins first.  All calls on one device share the array, so
 * calls that overlap in time on one device (other streams, other threads)
 * race and give wrong results; serialise them.  Results equal those of the
 * batch entry point (the launch schedule changes no result). */
lcbg_status lcbg_light_curve_device_noalloc(const lcbg_curve* curve, const lcbg_options* options,
                                            const double* times, int64_t n, double* mag, int32_t* method,
                                            void* stream);

/* Device time in milliseconds of the last kernel launched by the calling host
 * thread through a *_device entry point, measured with HIP events recorded on
 * the launch stream around that kernel.  Returns a negative value if none. */
double lcbg_last_kernel_ms(void);

#ifdef __cplusplus
}
#endif

#endif
