// C-ABI shim over the reference lcbinint C++ core, compiled from the reference
// sources (never copied) by oracle/Makefile into oracle/_ref/liblcbinint_ref.so.
//
// Test infrastructure only: the parity tests, the golden-vector generator and
// bench.py's cpu_baseline leg load this library through ctypes as the checker.
// Every entry point mirrors one reference call path so that a test can compare
// the HIP product against the reference on identical inputs.
#include "lcbinint/lcbinint.h"
#include "lcbinint/magnification/finite_source_magnifier.hpp"
#include "lcbinint/magnification/point_source_magnifier.hpp"
#include "lcbinint/math/polynomial_roots.hpp"
#include "lcbinint/model/lens_model.hpp"
#include "lcbinint/model/lens_parameters.hpp"
#include "lcbinint/model/orbital_motion.hpp"
#include "lcbinint/model/trajectory.hpp"
#include "SkowronGould.h"

// The reference's finite-source translation unit, compiled here from where it
// lies (the include path points into the reference tree) so that its internal
// stage functions can be exposed below.
#include "lcbinint/magnification/finite_source_magnifier.cpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

using lcbinint::SourcePosition;
using lcbinint::magnification::FiniteSourceMagnifier;
using lcbinint::magnification::FiniteSourceSettings;
using lcbinint::magnification::PointSourceMagnifier;

namespace {

// Flat settings record shared with the Python side (ctypes Structure of the
// same layout): the fields of FiniteSourceSettings (finite_source_magnifier.hpp:23-39).
struct RefSettings {
    int source_bins;
    int caustic_bins;
    double grid_ratio;
    int polar_source_bins;
    double polar_grid_ratio;
    int finite_mode;
    double kinji_threshold;
    double hex_threshold;
    double adaptive_hex_threshold;
    double limb_darkening_c;
    double limb_darkening_d;
    int adaptive_source_bins;
    int max_source_bins;
};

FiniteSourceSettings to_settings(const RefSettings& s)
{
    FiniteSourceSettings out;
    out.source_bins = s.source_bins;
    out.caustic_bins = s.caustic_bins;
    out.grid_ratio = s.grid_ratio;
    out.polar_source_bins = s.polar_source_bins;
    out.polar_grid_ratio = s.polar_grid_ratio;
    out.finite_mode = s.finite_mode;
    out.kinji_threshold = s.kinji_threshold;
    out.hex_threshold = s.hex_threshold;
    out.adaptive_hex_threshold = s.adaptive_hex_threshold;
    out.limb_darkening_c = s.limb_darkening_c;
    out.limb_darkening_d = s.limb_darkening_d;
    out.adaptive_source_bins = s.adaptive_source_bins;
    out.max_source_bins = s.max_source_bins;
    return out;
}

} // namespace

extern "C" {

// PointSourceMagnifier::binary_mag0 (point_source_magnifier.cpp:567), uncached,
// VBM convention (s, q as passed).
double ref_binary_mag0(double s, double q, double x, double y, int* image_count)
{
    PointSourceMagnifier pm;
    const auto r = pm.binary_mag0(s, q, {x, y});
    if (image_count) *image_count = r.image_count;
    return r.magnification;
}

// Batch of binary_mag0 calls; one (s, q) per call row.
void ref_binary_mag0_many(const double* s, const double* q, const double* x, const double* y,
                          long n, double* out, int* image_count)
{
    PointSourceMagnifier pm;
    for (long i = 0; i < n; ++i) {
        const auto r = pm.binary_mag0(s[i], q[i], {x[i], y[i]});
        out[i] = r.magnification;
        if (image_count) image_count[i] = r.image_count;
    }
}

// binary_mag0_with_derivatives (point_source_magnifier.cpp:583).
double ref_binary_mag0_with_derivatives(double s, double q, double x, double y,
                                        int* image_count, double* indicator)
{
    PointSourceMagnifier pm;
    const auto r = pm.binary_mag0_with_derivatives(s, q, {x, y});
    if (image_count) *image_count = r.image_count;
    if (indicator) *indicator = r.derivative_error_indicator;
    return r.magnification;
}

// binary_image_candidates (point_source_magnifier.cpp:739): out rows of
// (x, y, jacobian, residual, physical).
int ref_binary_image_candidates(double s, double q, double x, double y, double* out)
{
    PointSourceMagnifier pm;
    const auto c = pm.binary_image_candidates(s, q, {x, y});
    for (std::size_t i = 0; i < c.size(); ++i) {
        out[5 * i + 0] = c[i].position.x;
        out[5 * i + 1] = c[i].position.y;
        out[5 * i + 2] = c[i].jacobian_determinant;
        out[5 * i + 3] = c[i].residual;
        out[5 * i + 4] = c[i].physical ? 1.0 : 0.0;
    }
    return static_cast<int>(c.size());
}

// binary_lens_equation (point_source_magnifier.cpp:780).
void ref_binary_lens_equation(double s, double q, double zx, double zy, double* out)
{
    PointSourceMagnifier pm;
    const auto w = pm.binary_lens_equation(s, q, {zx, zy});
    out[0] = w.x;
    out[1] = w.y;
}

// cmplx_roots_gen (SkowronGould.cpp:79) on interleaved (re, im) coefficients.
void ref_cmplx_roots_gen(const double* coeffs, int degree, int polish, int use_start,
                         double* roots_inout)
{
    std::vector<complex> poly(static_cast<std::size_t>(degree) + 1);
    std::vector<complex> roots(static_cast<std::size_t>(degree));
    for (int i = 0; i <= degree; ++i) poly[i] = complex(coeffs[2 * i], coeffs[2 * i + 1]);
    for (int i = 0; i < degree; ++i) roots[i] = complex(roots_inout[2 * i], roots_inout[2 * i + 1]);
    cmplx_roots_gen(roots.data(), poly.data(), degree, polish != 0, use_start != 0);
    for (int i = 0; i < degree; ++i) {
        roots_inout[2 * i] = roots[i].re;
        roots_inout[2 * i + 1] = roots[i].im;
    }
}

// PolynomialRootSolver::solve (polynomial_roots.cpp:35) with default options.
int ref_poly_solve(const double* coeffs, int ncoeff, double* roots)
{
    std::vector<lcbinint::Complex> c(static_cast<std::size_t>(ncoeff));
    for (int i = 0; i < ncoeff; ++i) c[i] = {coeffs[2 * i], coeffs[2 * i + 1]};
    lcbinint::math::PolynomialRootSolver solver;
    const auto r = solver.solve(c);
    if (r.status != lcbinint::math::RootSolverStatus::ok) return -1;
    for (std::size_t i = 0; i < r.roots.size(); ++i) {
        roots[2 * i] = r.roots[i].real();
        roots[2 * i + 1] = r.roots[i].imag();
    }
    return static_cast<int>(r.roots.size());
}

// FiniteSourceMagnifier::binary_mag (finite_source_magnifier.cpp:4633) with a
// magnifier kept per settings, as the pybind layer does
// (lcbinint_pybind.cpp:927-943).  Returns the magnification; method and error
// estimate are written through the pointers.
double ref_fs_binary_mag(const RefSettings* rs, double s, double q, double x, double y,
                         double rho, double a_ps, int* method, double* error, int* converged)
{
    thread_local std::unique_ptr<FiniteSourceMagnifier> cached;
    thread_local RefSettings cached_settings {};
    thread_local bool valid = false;
    if (!valid || std::memcmp(&cached_settings, rs, sizeof(RefSettings)) != 0) {
        cached = std::make_unique<FiniteSourceMagnifier>(to_settings(*rs));
        cached_settings = *rs;
        valid = true;
    }
    PointSourceMagnifier pm;
    const auto r = cached->binary_mag(s, q, {x, y}, rho, a_ps, nullptr, true);
    (void)pm;
    if (method) *method = static_cast<int>(r.decision.method);
    if (error) *error = r.error_estimate;
    if (converged) *converged = r.converged ? 1 : 0;
    return r.magnification;
}

// The static VBM light-curve loop of evaluate_binary_light_curve_numpy
// (lcbinint_pybind.cpp:1120-1158): source from (t0, tE, u0, alpha), the uncached
// point-source solve, then binary_mag with effective q = 1/q.  Per-epoch
// method codes go to `method` when non-null.  Returns 0, or 2 on a non-finite
// magnification (the reference raises RuntimeError("numerical error")).
int ref_light_curve_static(const RefSettings* rs, const double* times, long n, double t0,
                           double tE, double u0, double alpha, double s, double q, double rho,
                           double* out, int* method)
{
    PointSourceMagnifier pm;
    const double cos_alpha = std::cos(alpha);
    const double sin_alpha = std::sin(alpha);
    const double inverse_tE = 1.0 / tE;
    const double effective_q = q != 0.0 ? 1.0 / q : q;
    if (rho == 0.0) {
        for (long i = 0; i < n; ++i) {
            const double tau = (times[i] - t0) * inverse_tE;
            const SourcePosition src {tau * cos_alpha - u0 * sin_alpha,
                                      tau * sin_alpha + u0 * cos_alpha};
            out[i] = pm.binary_mag0(s, effective_q, src).magnification;
            if (method) method[i] = 0;
        }
        return 0;
    }
    FiniteSourceMagnifier fm(to_settings(*rs));
    for (long i = 0; i < n; ++i) {
        const double tau = (times[i] - t0) * inverse_tE;
        const SourcePosition src {tau * cos_alpha - u0 * sin_alpha,
                                  tau * sin_alpha + u0 * cos_alpha};
        const auto point = pm.binary_mag0(s, effective_q, src);
        const auto fs = fm.binary_mag(s, effective_q, src, std::abs(rho), point.magnification,
                                      nullptr, true);
        out[i] = fs.magnification;
        if (method) method[i] = static_cast<int>(fs.decision.method);
        if (!std::isfinite(fs.magnification)) return 2;
    }
    return 0;
}

// evaluate_binary_light_curve_info (lcbinint_pybind.cpp:1176-1253), static VBM
// branch: per epoch the row (magnification, point-source magnification,
// finite-source magnification, source x, source y, image count, method, error
// estimate, refinement level, converged) in rec[10 i .. 10 i + 9], with the
// finite-source tolerances tol / reltol of the settings (they decide
// `converged`).  Returns 0, or 2 at the first non-finite magnification (rows
// from there on are left unwritten), where the reference raises.
int ref_light_curve_static_info(const RefSettings* rs, double tol, double reltol, const double* times,
                                long n, double t0, double tE, double u0, double alpha, double s,
                                double q, double rho, double* rec)
{
    PointSourceMagnifier pm;
    FiniteSourceSettings settings = to_settings(*rs);
    settings.finite_source_tol = tol;
    settings.finite_source_reltol = reltol;
    FiniteSourceMagnifier fm(settings);
    const double cos_alpha = std::cos(alpha);
    const double sin_alpha = std::sin(alpha);
    const double inverse_tE = 1.0 / tE;
    const double effective_q = q != 0.0 ? 1.0 / q : q;
    for (long i = 0; i < n; ++i) {
        const double tau = (times[i] - t0) * inverse_tE;
        const SourcePosition src {tau * cos_alpha - u0 * sin_alpha, tau * sin_alpha + u0 * cos_alpha};
        const auto point = pm.binary_mag0(s, effective_q, src);
        double* r = rec + 10 * i;
        r[1] = point.magnification;
        r[3] = src.x;
        r[4] = src.y;
        r[5] = point.image_count;
        if (rho == 0.0) {
            r[0] = r[2] = point.magnification;
            r[6] = 0.0;
            r[7] = 0.0;
            r[8] = 0.0;
            r[9] = 1.0;
            continue;
        }
        const auto fs = fm.binary_mag(s, effective_q, src, std::abs(rho), point.magnification, nullptr, true);
        if (!std::isfinite(fs.magnification)) return 2;
        r[0] = r[2] = fs.magnification;
        r[6] = static_cast<double>(static_cast<int>(fs.decision.method));
        r[7] = fs.error_estimate;
        r[8] = fs.refinement_level;
        r[9] = fs.converged ? 1.0 : 0.0;
    }
    return 0;
}

// Forced inverse-ray, as lcbinint.binary_inverse_ray (lcbinint_pybind.cpp:3826-3847):
// bbox exit, derivative check and hexadecapole are switched off.
double ref_binary_inverse_ray(const RefSettings* rs, double s, double q, double x, double y,
                              double rho, int* method)
{
    FiniteSourceSettings settings = to_settings(*rs);
    settings.kinji_threshold = 1e18;
    settings.hex_threshold = 0.0;
    settings.adaptive_hex_threshold = 0.0;
    PointSourceMagnifier pm;
    FiniteSourceMagnifier fm(settings);
    const auto point = pm.binary_mag0(s, q, {x, y});
    const auto r = fm.binary_mag(s, q, {x, y}, std::abs(rho), point.magnification, nullptr, true);
    if (method) *method = static_cast<int>(r.decision.method);
    return r.magnification;
}

// Orbital state (orbital_motion.cpp:183): separation and angle at `time`.
void ref_orbital_state(const lcbi_params* params, double time, double* out)
{
    const auto p = lcbinint::model::from_c_params(*params);
    const auto st = lcbinint::model::orbital_state(p, time);
    out[0] = st.separation;
    out[1] = st.angle;
    out[2] = st.line_of_sight_separation;
}


// ---------------------------------------------------------------- stages
// Stage accessors over functions internal to finite_source_magnifier.cpp.
namespace fsm = lcbinint::magnification;

// caustic_points_at_phase (finite_source_magnifier.cpp:547-567): 4 source-plane
// points (x, y interleaved) of the critical curve at one phase.
int ref_caustic_points_at_phase(double s, double q, double phase, double* out)
{
    PointSourceMagnifier pm;
    const auto pts = fsm::caustic_points_at_phase(pm, s, q, phase);
    for (std::size_t i = 0; i < pts.size(); ++i) {
        out[2 * i] = pts[i].x;
        out[2 * i + 1] = pts[i].y;
    }
    return static_cast<int>(pts.size());
}

// The roots of the critical-curve quartic at one phase (x, y interleaved), as
// py_critical_points_at_phase (python/lcbinint_pybind.cpp:585-604) and the
// first half of caustic_points_at_phase compute them.
int ref_critical_points_at_phase(double s, double q, double phase, double* out)
{
    const lcbinint::Complex ph = std::polar(1.0, phase);
    lcbinint::math::PolynomialRootSolver solver;
    const auto roots = solver.solve(fsm::critical_curve_polynomial_coefficients(s, q, ph));
    if (roots.status != lcbinint::math::RootSolverStatus::ok) return 0;
    for (std::size_t i = 0; i < roots.roots.size(); ++i) {
        out[2 * i] = roots.roots[i].real();
        out[2 * i + 1] = roots.roots[i].imag();
    }
    return static_cast<int>(roots.roots.size());
}

// critical_sources_at_phase (finite_source_magnifier.cpp:1622-1641).
int ref_critical_sources_at_phase(double s, double q, double phase, double* out)
{
    const auto mapper = fsm::make_binary_lens_mapper(s, q);
    const auto pts = fsm::critical_sources_at_phase(mapper, s, q, phase);
    for (std::size_t i = 0; i < pts.size(); ++i) {
        out[2 * i] = pts[i].x;
        out[2 * i + 1] = pts[i].y;
    }
    return static_cast<int>(pts.size());
}

// augmented_image_seeds (finite_source_magnifier.cpp:1709-1871) followed by
// augment_seeds_from_caustic_branches (:4845-4926), as fixed_inverse_ray_binary
// runs them (:3055-3063).  Writes up to `cap` seeds; returns the seed count.
int ref_image_seeds(const RefSettings* rs, double s, double q, double x, double y, double rho,
                    double* out, int cap)
{
    PointSourceMagnifier pm;
    FiniteSourceMagnifier fm(to_settings(*rs));
    const auto mapper = fsm::make_binary_lens_mapper(s, q);
    auto seeds = fsm::augmented_image_seeds(pm, mapper, s, q, {x, y}, rho);
    fm.augment_seeds_from_caustic_branches(s, q, {x, y}, rho, seeds);
    for (std::size_t i = 0; i < seeds.size() && static_cast<int>(i) < cap; ++i) {
        out[2 * i] = seeds[i].x;
        out[2 * i + 1] = seeds[i].y;
    }
    return static_cast<int>(seeds.size());
}

// inverse_ray_cartesian_binary_mag (finite_source_magnifier.cpp:2470-2768) on a
// given seed list; diagnostics as (seed_count, processed, fold, rows, gaps,
// overlaps, max_jump, estimated_error).
double ref_cartesian_area(const RefSettings* rs, double s, double q, double x, double y,
                          double rho, const double* seeds_xy, int n_seeds, double* diag)
{
    PointSourceMagnifier pm;
    FiniteSourceMagnifier fm(to_settings(*rs));
    std::vector<SourcePosition> seeds(static_cast<std::size_t>(n_seeds));
    for (int i = 0; i < n_seeds; ++i) seeds[i] = {seeds_xy[2 * i], seeds_xy[2 * i + 1]};
    fsm::LegacyAreaDiagnostics d;
    const double a = fsm::inverse_ray_cartesian_binary_mag(pm, s, q, {x, y}, rho, fm.settings(),
                                                           &fm, &seeds, &d, nullptr);
    if (diag) {
        diag[0] = d.seed_count;
        diag[1] = d.processed_images;
        diag[2] = d.fold_seed_count;
        diag[3] = d.boundary_rows;
        diag[4] = d.gap_repairs;
        diag[5] = d.overlaps;
        diag[6] = d.max_jump_cells;
        diag[7] = d.estimated_error;
    }
    return a;
}

// inverse_ray_polar_boundary_binary (finite_source_magnifier.cpp:1156-1319) on a
// given seed list, with the settings as passed (callers raise polar_grid_ratio
// to 12 for the automatic mode as binary_mag does, :4813-4821).
double ref_polar_area(const RefSettings* rs, double s, double q, double x, double y, double rho,
                      const double* seeds_xy, int n_seeds)
{
    PointSourceMagnifier pm;
    FiniteSourceMagnifier fm(to_settings(*rs));
    std::vector<SourcePosition> seeds(static_cast<std::size_t>(n_seeds));
    for (int i = 0; i < n_seeds; ++i) seeds[i] = {seeds_xy[2 * i], seeds_xy[2 * i + 1]};
    return fsm::inverse_ray_polar_boundary_binary(pm, s, q, {x, y}, rho, fm.settings(), &fm,
                                                  nullptr, &seeds);
}

// hexadecapole_binary (finite_source_magnifier.cpp:792-847) with known A0.
double ref_hexadecapole(const RefSettings* rs, double s, double q, double x, double y,
                        double rho, double a0, double* rel_err)
{
    PointSourceMagnifier pm;
    const auto settings = to_settings(*rs);
    const auto h = fsm::hexadecapole_binary(pm, s, q, {x, y}, rho, settings, &a0);
    if (rel_err) *rel_err = h.relative_error;
    return h.magnification;
}

// A binary-lens curve with annual parallax and/or orbital motion, as
// LightCurve(parallax=..., orbital_motion_mode=...) evaluates it through
// LensModel::magnification (lens_model.cpp:83-236; binary lens, VBM convention,
// the path evaluate_binary_light_curve_numpy takes for such curves,
// python/lcbinint_pybind.cpp:1159-1173).  p: t0, tE, u0, alpha, s, q, rho,
// piEN, piEE, ra, dec, tfix, g1, g2, g3, lom_szs, lom_ar.  src (optional, 3n):
// the source position the magnifier receives and the separation s(t).
// Returns 0, 2 on a numerical error (first failing epoch left in out).
int ref_light_curve_dynamic(const RefSettings* rs, const double* times, long n, const double* p,
                            int orbit_mode, double* out, int* method, double* src)
{
    using namespace lcbinint::model;
    LensParameters lp;
    lp.t0 = p[0];
    lp.tE = p[1];
    lp.umin = p[2];
    lp.theta = p[3];
    lp.sep = p[4];
    lp.q = p[5];
    lp.rho = p[6];
    lp.piEN = p[7];
    lp.piEE = p[8];
    lp.ra = p[9];
    lp.dec = p[10];
    lp.tfix = p[11];
    lp.g1 = p[12];
    lp.g2 = p[13];
    lp.g3 = p[14];
    lp.lom_szs = p[15];
    lp.lom_ar = p[16];
    lp.orbital_motion_mode = static_cast<lcbi_orbital_motion_mode>(orbit_mode);
    lp.limb_darkening_c = rs->limb_darkening_c;
    lp.limb_darkening_d = rs->limb_darkening_d;
    const Trajectory trajectory(lp);
    const double cos_theta = std::cos(lp.theta);
    const double sin_theta = std::sin(lp.theta);
    PointSourceMagnifier pm;
    FiniteSourceMagnifier fm(to_settings(*rs));
    const double effective_q = lp.q != 0.0 ? 1.0 / lp.q : lp.q;
    for (long i = 0; i < n; ++i) {
        const double time = times[i];
        SourcePosition source;
        if (lp.piEN == 0.0 && lp.piEE == 0.0) {
            const double tn = (time - lp.t0) / lp.tE;
            const double beta = lp.umin;
            source.x = tn * cos_theta - beta * sin_theta;
            source.y = tn * sin_theta + beta * cos_theta;
        } else {
            source = trajectory.source_position(time, true, LCBI_XALLARAP_NONE);
        }
        const bool static_orbit = lp.orbital_motion_mode == LCBI_ORBIT_STATIC;
        const auto orbit = static_orbit ? OrbitalState {lp.sep, lp.theta, 0.0} : orbital_state(lp, time);
        if (!std::isfinite(orbit.separation) || !std::isfinite(orbit.angle)) return 2;
        SourcePosition sm = source;
        if (!static_orbit) {
            double tau = 0.0;
            double beta = 0.0;
            if (lp.piEN == 0.0 && lp.piEE == 0.0) {
                tau = (time - lp.t0) / lp.tE;
                beta = lp.umin;
            } else {
                tau = source.x * cos_theta + source.y * sin_theta;
                beta = -source.x * sin_theta + source.y * cos_theta;
            }
            sm = {tau * std::cos(orbit.angle) - beta * std::sin(orbit.angle),
                  beta * std::cos(orbit.angle) + tau * std::sin(orbit.angle)};
        }
        if (src) {
            src[3 * i] = sm.x;
            src[3 * i + 1] = sm.y;
            src[3 * i + 2] = orbit.separation;
        }
        if (lp.rho == 0.0) {
            out[i] = pm.binary_mag0(orbit.separation, effective_q, sm).magnification;
            if (method) method[i] = 0;
            continue;
        }
        const auto images = pm.binary_images(orbit.separation, effective_q, sm);
        double a_ps = 0.0;
        std::vector<SourcePosition> seeds;
        seeds.reserve(images.size());
        for (const auto& image : images) {
            a_ps += 1.0 / std::abs(image.jacobian_determinant);
            seeds.push_back(image.position);
        }
        const auto fs = fm.binary_mag(orbit.separation, effective_q, sm, std::abs(lp.rho), a_ps,
                                      &seeds, true);
        out[i] = fs.magnification;
        if (method) method[i] = static_cast<int>(fs.decision.method);
        if (!std::isfinite(fs.magnification)) return 2;
    }
    return 0;
}

// LensModel::magnification's binary finite-source step (lens_model.cpp:186-234,
// as ref_light_curve_dynamic runs it) at a given source position (x, y) in the
// binary frame and separation s: the sum over binary_images and the
// dispatcher with those images as seed hints.  Returns the magnification and
// the method.
double ref_lens_model_at(const RefSettings* rs, double s, double q, double x, double y, double rho, int* method)
{
    PointSourceMagnifier pm;
    FiniteSourceMagnifier fm(to_settings(*rs));
    const double effective_q = q != 0.0 ? 1.0 / q : q;
    const SourcePosition sm {x, y};
    const auto images = pm.binary_images(s, effective_q, sm);
    double a_ps = 0.0;
    std::vector<SourcePosition> seeds;
    seeds.reserve(images.size());
    for (const auto& image : images) {
        a_ps += 1.0 / std::abs(image.jacobian_determinant);
        seeds.push_back(image.position);
    }
    const auto fs = fm.binary_mag(s, effective_q, sm, std::abs(rho), a_ps, &seeds, true);
    if (method) *method = static_cast<int>(fs.decision.method);
    return fs.magnification;
}

// The geometry helpers' source_for_geometry (python/lcbinint_pybind.cpp:1840-1880),
// VBM frame, over the core's Trajectory and orbital_state (the binding's orbit
// cache, :803-925, evaluates orbital_state's expressions).  p as in
// ref_light_curve_dynamic; xy: x, y interleaved.
int ref_source_for_geometry(const double* p, int orbit_mode, const double* times, long n, double* xy)
{
    using namespace lcbinint::model;
    LensParameters lp;
    lp.t0 = p[0];
    lp.tE = p[1];
    lp.umin = p[2];
    lp.theta = p[3];
    lp.sep = p[4];
    lp.q = p[5];
    lp.rho = p[6];
    lp.piEN = p[7];
    lp.piEE = p[8];
    lp.ra = p[9];
    lp.dec = p[10];
    lp.tfix = p[11];
    lp.g1 = p[12];
    lp.g2 = p[13];
    lp.g3 = p[14];
    lp.lom_szs = p[15];
    lp.lom_ar = p[16];
    lp.orbital_motion_mode = static_cast<lcbi_orbital_motion_mode>(orbit_mode);
    const Trajectory trajectory(lp);
    for (long i = 0; i < n; ++i) {
        const double time = times[i];
        SourcePosition source;
        if (lp.piEN == 0.0 && lp.piEE == 0.0) {
            const double tau = (time - lp.t0) / lp.tE;
            const double beta = lp.umin;
            const double cos_theta = std::cos(lp.theta);
            const double sin_theta = std::sin(lp.theta);
            source.x = tau * cos_theta - beta * sin_theta;
            source.y = tau * sin_theta + beta * cos_theta;
        } else {
            source = trajectory.source_position(time, true, LCBI_XALLARAP_NONE);
        }
        if (lp.orbital_motion_mode != LCBI_ORBIT_STATIC) {
            const auto orbit = orbital_state(lp, time);
            const double cos_theta = std::cos(lp.theta);
            const double sin_theta = std::sin(lp.theta);
            const double tau = source.x * cos_theta + source.y * sin_theta;
            const double beta = -source.x * sin_theta + source.y * cos_theta;
            source = {tau * std::cos(orbit.angle) - beta * std::sin(orbit.angle),
                      beta * std::cos(orbit.angle) + tau * std::sin(orbit.angle)};
        }
        xy[2 * i] = source.x;
        xy[2 * i + 1] = source.y;
    }
    return 0;
}

int ref_sizeof_params(void) { return static_cast<int>(sizeof(lcbi_params)); }
int ref_sizeof_options(void) { return static_cast<int>(sizeof(lcbi_options)); }
int ref_sizeof_result(void) { return static_cast<int>(sizeof(lcbi_result)); }

} // extern "C"
