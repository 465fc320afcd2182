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
 * (include/lcbinint/lcbinint.h:85-104) that the plain method reads. */
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
} lcbg_options;

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
 * host libm, as the reference's std::cos/std::sin are) and allocates the
 * batch's device workspace; the run entry points allocate nothing.  A batch
 * is used by one host thread at a time.  `curves` is host memory. */
typedef struct lcbg_batch lcbg_batch;

lcbg_status lcbg_batch_create(const lcbg_curve* curves, int32_t n_curves,
                              const lcbg_options* options, lcbg_batch** out);
lcbg_status lcbg_batch_destroy(lcbg_batch* batch);

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

/* Finite-source light curves of every curve of the batch: the plain method of
 * FiniteSourceMagnifier::binary_mag (finite_source_magnifier.cpp:4633-4843) as
 * the static VBM epoch loop of evaluate_binary_light_curve_numpy
 * (python/lcbinint_pybind.cpp:1141-1158) runs it: point source, hexadecapole,
 * or inverse-ray shooting on the Cartesian or polar lattice with limb
 * darkening.  Device pointers; n_epochs = offsets[n_curves].  method[i]
 * receives the lcbg_method of each epoch.  The first call on a batch builds the
 * per-curve caustic caches; later calls reuse them and the workspace. */
lcbg_status lcbg_batch_finite_source_device(lcbg_batch* batch, const int64_t* offsets,
                                            const double* times, int64_t n_epochs, double* mag,
                                            int32_t* method, void* stream);

/* lcbg_batch_finite_source_device, then waits and reports the device time of
 * its three stages (stage_ms[3]: point source + bounding box, caustic distance
 * + hexadecapole, inverse ray), the queue sizes (queue_counts[2]: epochs past
 * the bounding box, inverse-ray epochs) and the overflow flags (error_flags). */
lcbg_status lcbg_batch_finite_source_profile(lcbg_batch* batch, const int64_t* offsets,
                                             const double* times, int64_t n_epochs, double* mag,
                                             int32_t* method, void* stream, float* stage_ms,
                                             uint32_t* queue_counts, uint32_t* error_flags);

/* Host-pointer light curve of one curve with the given options (NULL =
 * defaults): the drop-in path behind LightCurve(...)(times, params).  Returns
 * LCBG_NUMERICAL_ERROR if any magnification is not finite (the reference raises
 * RuntimeError("numerical error"), python/lcbinint_pybind.cpp:1153-1155) and
 * LCBG_UNSUPPORTED if a workspace capacity was exceeded. */
lcbg_status lcbg_light_curve(const lcbg_curve* curve, const lcbg_options* options,
                             const double* times, int64_t n, double* mag, int32_t* method);

/* Device time in milliseconds of the last kernel launched by the calling host
 * thread through a *_device entry point, measured with HIP events recorded on
 * the launch stream around that kernel.  Returns a negative value if none. */
double lcbg_last_kernel_ms(void);

#ifdef __cplusplus
}
#endif

#endif
