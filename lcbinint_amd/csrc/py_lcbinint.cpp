// pybind11 drop-in for the reference Python API on the finite-source hot path:
//
//   lcbinint.LightCurve(options=Options(coordinates=..., nbin=...),
//                       limb_darkening=LimbDarkening.linear(u))(times, params)
//
// Same class names, keyword arguments, defaults, aliases and error types as
// python/lcbinint_pybind.cpp (LimbDarkening :2828-2847, Options :2985-3112,
// LightCurve :3114-3181, dict keys :145-179).  The light curve is evaluated by
// the HIP library with the GIL released: the static binary lens through
// lcbg_light_curve, annual parallax and circular / Keplerian orbital motion
// through lcbg_batch_light_curves_dynamic (the LensModel path), with terrestrial
// parallax and xallarap there too; a binary source blends two such curves; the
// triple lens and the legacy coordinates raise as unsupported.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "lcbinint_amd.h"

namespace py = pybind11;

namespace {

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

// lcbi_orbital_motion_mode (include/lcbinint/lcbinint.h:15-19)
enum OrbitalMotionMode { STATIC = 0, CIRCULAR = 1, KEPLER = 2 };

struct PyLimbDarkening {
    double c = 0.0;
    double d = 0.0;
};

// The public option set (lcbi_options of the reference, python defaults of
// public_default_options, lcbinint_pybind.cpp:213-220).
struct PyOptions {
    lcbg_options o = lcbg_default_options();
    int vbm_compatible = 1;
    int center_of_mass = 0;
    int adaptive_source_bins = 0;
    int max_source_bins = 400;
    double finite_source_tol = 0.0;
    double finite_source_reltol = 0.0;
    std::string xallarap = "none";
};

void apply_param_type(PyOptions& p, const std::string& t)
{
    if (t == "auto" || t.empty() || t == "vbm" || t == "vbbl" || t == "standard") {
        p.center_of_mass = 0;
        p.vbm_compatible = 1;
    } else if (t == "lcbinint" || t == "original" || t == "legacy") {
        p.center_of_mass = 0;
        p.vbm_compatible = 0;
    } else if (t == "center_of_mass") {
        p.center_of_mass = 1;
        p.vbm_compatible = 0;
    } else if (t == "vbm_center_of_mass") {
        p.center_of_mass = 1;
        p.vbm_compatible = 1;
    } else {
        throw std::invalid_argument(
            "param_type must be 'vbm', 'lcbinint', 'original', or 'center_of_mass'");
    }
}

std::string coordinate_name(const PyOptions& p)
{
    if (p.vbm_compatible != 0 && p.center_of_mass == 0) return "vbm";
    if (p.center_of_mass != 0 && p.vbm_compatible != 0) return "vbm_center_of_mass";
    if (p.center_of_mass != 0) return "center_of_mass";
    return "lcbinint";
}

void apply_grid(PyOptions& p, const std::string& g)
{
    if (g == "auto" || g.empty()) {
        p.o.mode = 4;
    } else if (g == "cartesian" || g == "inverse_ray_cartesian") {
        p.o.mode = 1;
    } else if (g == "polar" || g == "inverse_ray_polar") {
        p.o.mode = 2;
    } else {
        throw std::invalid_argument("inverse_ray_grid must be 'auto', 'cartesian', or 'polar'");
    }
}

std::string grid_name(const PyOptions& p)
{
    if (p.o.mode == 4) return "auto";
    if (p.o.mode == 1) return "cartesian";
    if (p.o.mode == 2) return "polar";
    return "custom";
}

int int_or(const py::object& v, int d) { return v.is_none() ? d : py::cast<int>(v); }
double double_or(const py::object& v, double d) { return v.is_none() ? d : py::cast<double>(v); }
py::object positive_int(int v) { return v <= 0 ? py::object(py::none()) : py::object(py::int_(v)); }
py::object positive_double(double v)
{
    return v <= 0.0 ? py::object(py::none()) : py::object(py::float_(v));
}

double dict_get(const py::dict& d, const char* k, double dflt)
{
    const py::str key(k);
    return d.contains(key) ? py::cast<double>(d[key]) : dflt;
}

void raise_status(lcbg_status st)
{
    switch (st) {
    case LCBG_OK: return;
    case LCBG_NUMERICAL_ERROR: throw std::runtime_error("numerical error");
    case LCBG_UNSUPPORTED: throw std::runtime_error("unsupported");
    case LCBG_INVALID_ARGUMENT: throw std::invalid_argument("invalid argument");
    default: throw std::runtime_error(lcbg_status_string(st));
    }
}

// LightCurveInfo (python/lcbinint_pybind.cpp:2790-2804): per-epoch records.
struct PyLightCurveInfo {
    std::vector<double> times, magnifications, point_source_magnifications, finite_source_magnifications,
        finite_source_error_estimates, source_x, source_y;
    std::vector<int> image_counts, finite_source_methods, finite_source_refinement_levels;
    std::vector<std::string> finite_source_method_names;
    std::vector<bool> finite_source_converged;
};

const char* method_name(int m)
{
    switch (m) {
    case LCBG_METHOD_POINT_SOURCE: return "point_source";
    case LCBG_METHOD_HEXADECAPOLE: return "hexadecapole";
    case LCBG_METHOD_INVERSE_RAY_CARTESIAN: return "inverse_ray_cartesian";
    case LCBG_METHOD_INVERSE_RAY_POLAR: return "inverse_ray_polar";
    }
    return "unknown";
}

// EventCoordinates (python/lcbinint_pybind.cpp:95-101, 2848-2866)
struct PyEvent {
    double ra = 0.0, dec = 0.0, tfix = 0.0, obs_lat = 0.0, obs_lon = 0.0;
};

// GeometryCurve and GeometryBranches (python/lcbinint_pybind.cpp:45-54)
struct PyGeometryCurve {
    std::vector<double> times, x, y;
};

struct PyGeometryBranches {
    std::vector<std::vector<double>> x, y;
};

// The one-curve handle of the LensModel path, kept by the LightCurve across
// calls (its device buffers and caches are reused; the curve is updated).
struct DynHandle {
    std::mutex mu;
    lcbg_batch* batch = nullptr;
    lcbg_options options{};
    ~DynHandle()
    {
        if (batch) lcbg_batch_destroy(batch);
    }
};

class PyLightCurve {
public:
    PyLightCurve(std::string lens, std::string source, const PyEvent& event, const PyOptions& options,
                 const PyLimbDarkening& ld, int orbital_motion_mode, bool parallax,
                 bool terrestrial_parallax)
        : lens_(std::move(lens)), source_(std::move(source)), event_(event), options_(options), ld_(ld),
          orbit_(orbital_motion_mode), parallax_(parallax), terrestrial_(terrestrial_parallax),
          dyn_(std::make_shared<DynHandle>())
    {
        if (lens_ != "binary_lens" && lens_ != "triple_lens")
            throw std::invalid_argument("lens must be 'binary_lens' or 'triple_lens'");
        if (source_ != "single" && source_ != "binary")
            throw std::invalid_argument("source must be 'single' or 'binary'");
    }

    // the binary lens in VBM coordinates (static, annual and terrestrial
    // parallax, orbital motion, xallarap; single or binary source); triple
    // lenses and the legacy frame are outside this library
    void check_supported(double tE, double s, double q) const
    {
        const bool supported = options_.vbm_compatible != 0 && options_.center_of_mass == 0;
        if (!supported) throw std::runtime_error("unsupported");
        if (!std::isfinite(tE) || tE == 0.0 || !std::isfinite(q) || !std::isfinite(s))
            throw std::runtime_error("numerical error");
    }

    static std::vector<double> times_of(const py::array_t<double, py::array::c_style | py::array::forcecast>& times)
    {
        const py::buffer_info in = times.request();
        if (in.ndim != 1) throw std::invalid_argument("times must be one-dimensional");
        const auto n = static_cast<py::ssize_t>(in.shape[0]);
        return std::vector<double>(static_cast<const double*>(in.ptr), static_cast<const double*>(in.ptr) + n);
    }

    lcbg_options run_options() const
    {
        lcbg_options opts = options_.o;
        opts.finite_source_tol = options_.finite_source_tol;
        opts.finite_source_reltol = options_.finite_source_reltol;
        opts.adaptive_source_bins = options_.adaptive_source_bins;
        opts.max_source_bins = options_.max_source_bins;
        return opts;
    }

    // The parameters of binary_params_from_dict (python/lcbinint_pybind.cpp:145-179).
    struct CurveParams {
        double t0 = 0.0, tE = 1.0, u0 = 0.0, alpha = 0.0, s = 1.0, q = 1.0, rho = 0.0;
        double piEN = 0.0, piEE = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0, lom_szs = 0.0, lom_ar = 1.0;
        double u0_2 = 0.0, t0_2 = 0.0, flux_ratio = 0.0;
        double xi_1 = 0.0, xi_2 = 0.0, omega_xa = 0.0, inc_xa = 0.0, phi_xa = 0.0;
        double piEN_xa = 0.0, piEE_xa = 0.0, period_xa = 0.0, ecc_xa = 0.0, peri_xa = 0.0;
        double q2 = 0.0, sep2 = 0.0, ang = 0.0;
    };

    static CurveParams params_from_dict(const py::dict& d)
    {
        CurveParams p;
        const std::pair<const char*, double*> keys[] = {
            {"t0", &p.t0}, {"tE", &p.tE}, {"u0", &p.u0}, {"alpha", &p.alpha}, {"s", &p.s}, {"q", &p.q},
            {"rho", &p.rho}, {"piEN", &p.piEN}, {"piEE", &p.piEE}, {"g1", &p.g1}, {"g2", &p.g2}, {"g3", &p.g3},
            {"lom_szs", &p.lom_szs}, {"lom_ar", &p.lom_ar}, {"u0_2", &p.u0_2}, {"t0_2", &p.t0_2},
            {"flux_ratio", &p.flux_ratio}, {"xi_1", &p.xi_1}, {"xi_2", &p.xi_2}, {"omega_xa", &p.omega_xa},
            {"inc_xa", &p.inc_xa}, {"phi_xa", &p.phi_xa}, {"piEN_xa", &p.piEN_xa}, {"piEE_xa", &p.piEE_xa},
            {"period_xa", &p.period_xa}, {"ecc_xa", &p.ecc_xa}, {"peri_xa", &p.peri_xa}, {"q2", &p.q2},
            {"sep2", &p.sep2}, {"ang", &p.ang}};
        for (const auto& k : keys) *k.second = dict_get(d, k.first, *k.second);
        return p;
    }

    static CurveParams positional(double t0, double tE, double u0, double alpha, double s, double q, double rho,
                                  double piEN, double piEE, double g1, double g2, double g3, double lom_szs,
                                  double lom_ar)
    {
        CurveParams p;
        p.t0 = t0;
        p.tE = tE;
        p.u0 = u0;
        p.alpha = alpha;
        p.s = s;
        p.q = q;
        p.rho = rho;
        p.piEN = piEN;
        p.piEE = piEE;
        p.g1 = g1;
        p.g2 = g2;
        p.g3 = g3;
        p.lom_szs = lom_szs;
        p.lom_ar = lom_ar;
        return p;
    }

    // make_c_params_for_single_source (python/lcbinint_pybind.cpp:2720-2748):
    // parallax only with LightCurve(parallax=True), the observatory only with
    // terrestrial_parallax=True, xallarap in the options' form
    lcbg_dynamics dynamics_of(const CurveParams& p) const
    {
        lcbg_dynamics d{};
        d.piEN = parallax_ ? p.piEN : 0.0;
        d.piEE = parallax_ ? p.piEE : 0.0;
        d.ra = event_.ra;
        d.dec = event_.dec;
        d.tfix = event_.tfix;
        d.g1 = p.g1;
        d.g2 = p.g2;
        d.g3 = p.g3;
        d.lom_szs = p.lom_szs;
        d.lom_ar = p.lom_ar;
        d.obs_lat = terrestrial_ ? event_.obs_lat : 0.0;
        d.obs_lon = terrestrial_ ? event_.obs_lon : 0.0;
        d.xallarap_type = options_.xallarap == "angular_velocity" ? 1 : options_.xallarap == "orbital_elements" ? 2 : 0;
        d.xi_1 = p.xi_1;
        d.xi_2 = p.xi_2;
        d.omega_xa = p.omega_xa;
        d.inc_xa = p.inc_xa;
        d.phi_xa = p.phi_xa;
        d.piEN_xa = p.piEN_xa;
        d.piEE_xa = p.piEE_xa;
        d.period_xa = p.period_xa;
        d.ecc_xa = p.ecc_xa;
        d.peri_xa = p.peri_xa;
        return d;
    }

    // direct_static_binary of evaluate_binary_light_curve_numpy
    // (python/lcbinint_pybind.cpp:1103-1116): otherwise the LensModel path
    bool model_path(const lcbg_dynamics& d) const
    {
        return orbit_ != 0 || d.piEN != 0.0 || d.piEE != 0.0 || d.xi_1 != 0.0 || d.xi_2 != 0.0 ||
               d.piEN_xa != 0.0 || d.piEE_xa != 0.0;
    }

    // The triple lens (make_c_params_for_single_source :2722-2731 and
    // LensModel's triple branch, lens_model.cpp:123-150): the point-source
    // magnification on the static trajectory; a finite source, parallax,
    // xallarap or orbital motion with three lenses is outside this library.
    void check_triple(const CurveParams& p) const
    {
        if (!(p.q2 > 0.0)) throw py::value_error("triple_lens requires q2 > 0");
        const lcbg_dynamics d = dynamics_of(p);
        if (parallax_ || orbit_ != 0 || p.rho != 0.0 || d.xi_1 != 0.0 || d.xi_2 != 0.0 || d.piEN_xa != 0.0 ||
            d.piEE_xa != 0.0)
            throw std::runtime_error("unsupported");
    }

    std::vector<double> triple_curve(const std::vector<double>& t, const CurveParams& p,
                                     std::vector<int32_t>* counts = nullptr) const
    {
        check_supported(p.tE, p.s, p.q);
        check_triple(p);
        const lcbg_triple tr{p.t0, p.tE, p.u0, p.alpha, p.s, p.q, p.q2, p.sep2, p.ang};
        std::vector<double> out(t.size());
        std::vector<int32_t> cnt(t.size());
        lcbg_status st;
        {
            py::gil_scoped_release release;
            st = lcbg_triple_light_curve_point_source(&tr, t.data(), static_cast<int64_t>(t.size()), out.data(),
                                                      cnt.data());
        }
        raise_status(st);
        if (counts) *counts = cnt;
        return out;
    }

    // One source's light curve (the static VBM branch or the LensModel branch).
    std::vector<double> source_curve(const std::vector<double>& t, const CurveParams& p) const
    {
        if (lens_ == "triple_lens") return triple_curve(t, p);
        check_supported(p.tE, p.s, p.q);
        const lcbg_curve cv{p.t0, p.tE, p.u0, p.alpha, p.s, p.q, p.rho, ld_.c, ld_.d};
        const lcbg_options opts = run_options();
        const lcbg_dynamics dyn = dynamics_of(p);
        const int64_t n = static_cast<int64_t>(t.size());
        std::vector<double> out(t.size());
        lcbg_status st;
        {
            py::gil_scoped_release release;
            if (!model_path(dyn)) {
                st = lcbg_light_curve(&cv, &opts, t.data(), n, out.data(), nullptr);
            } else {
                std::lock_guard<std::mutex> lock(dyn_->mu);
                lcbg_batch* b = nullptr;
                st = dyn_batch(cv, opts, &b);
                const int64_t off[2] = {0, n};
                std::vector<int32_t> meth(t.size());
                if (st == LCBG_OK)
                    st = lcbg_batch_light_curves_dynamic(b, &dyn, orbit_, off, t.data(), n, out.data(), meth.data());
                if (st == LCBG_OK)
                    for (double v : out)
                        if (!std::isfinite(v)) st = LCBG_NUMERICAL_ERROR;
            }
        }
        raise_status(st);
        return out;
    }

    // The second source of a binary source: the same model at (u0_2, t0_2)
    // (python/lcbinint_pybind.cpp:2331-2353).
    static CurveParams second_source(const CurveParams& p)
    {
        CurveParams p2 = p;
        p2.u0 = p.u0_2;
        p2.t0 = p.t0_2;
        return p2;
    }

    // evaluate_binary_light_curve_info (python/lcbinint_pybind.cpp:1176-1290)
    PyLightCurveInfo records(const std::vector<double>& t, const CurveParams& p) const
    {
        if (lens_ == "triple_lens") {
            // point-source records: the magnification, the image count and
            // the source position of the static trajectory
            std::vector<int32_t> cnt;
            const std::vector<double> a = triple_curve(t, p, &cnt);
            const double ca = std::cos(p.alpha), sa = std::sin(p.alpha);
            PyLightCurveInfo out;
            out.times = t;
            for (std::size_t i = 0; i < t.size(); ++i) {
                const double tn = (t[i] - p.t0) / p.tE;
                out.magnifications.push_back(a[i]);
                out.point_source_magnifications.push_back(a[i]);
                out.finite_source_magnifications.push_back(a[i]);
                out.finite_source_error_estimates.push_back(0.0);
                out.source_x.push_back(tn * ca - p.u0 * sa);
                out.source_y.push_back(tn * sa + p.u0 * ca);
                out.image_counts.push_back(cnt[i]);
                out.finite_source_methods.push_back(LCBG_METHOD_POINT_SOURCE);
                out.finite_source_method_names.push_back(method_name(LCBG_METHOD_POINT_SOURCE));
                out.finite_source_refinement_levels.push_back(0);
                out.finite_source_converged.push_back(true);
            }
            return out;
        }
        check_supported(p.tE, p.s, p.q);
        const lcbg_curve cv{p.t0, p.tE, p.u0, p.alpha, p.s, p.q, p.rho, ld_.c, ld_.d};
        const lcbg_options opts = run_options();
        const lcbg_dynamics dyn = dynamics_of(p);
        const int64_t n = static_cast<int64_t>(t.size());
        std::vector<lcbg_result> rec(t.size());
        lcbg_status st;
        {
            py::gil_scoped_release release;
            if (!model_path(dyn)) {
                int64_t first = -1;
                st = lcbg_light_curve_results(&cv, &opts, t.data(), n, rec.data(), &first);
            } else {
                std::lock_guard<std::mutex> lock(dyn_->mu);
                lcbg_batch* b = nullptr;
                st = dyn_batch(cv, opts, &b);
                const int64_t off[2] = {0, n};
                if (st == LCBG_OK)
                    st = lcbg_batch_light_curves_dynamic_results(b, &dyn, orbit_, off, t.data(), n, rec.data());
            }
        }
        raise_status(st);
        PyLightCurveInfo out;
        out.times = t;
        for (const lcbg_result& r : rec) {
            out.magnifications.push_back(r.magnification);
            out.point_source_magnifications.push_back(r.point_source_magnification);
            out.finite_source_magnifications.push_back(r.finite_source_magnification);
            out.finite_source_error_estimates.push_back(r.error_estimate);
            out.source_x.push_back(r.source_x);
            out.source_y.push_back(r.source_y);
            out.image_counts.push_back(r.image_count);
            out.finite_source_methods.push_back(r.method);
            out.finite_source_method_names.push_back(method_name(r.method));
            out.finite_source_refinement_levels.push_back(r.refinement_level);
            out.finite_source_converged.push_back(r.converged != 0);
        }
        return out;
    }

    PyLightCurveInfo info(py::array_t<double, py::array::c_style | py::array::forcecast> times, double t0,
                          double tE, double u0, double alpha, double s, double q, double rho, double piEN,
                          double piEE, double g1, double g2, double g3, double lom_szs, double lom_ar) const
    {
        return records(times_of(times), positional(t0, tE, u0, alpha, s, q, rho, piEN, piEE, g1, g2, g3, lom_szs,
                                                   lom_ar));
    }

    // info_from_dict (python/lcbinint_pybind.cpp:2455-2479): a binary source
    // blends the two magnification columns, the other fields are the first
    // source's
    PyLightCurveInfo info_from_dict(py::array_t<double, py::array::c_style | py::array::forcecast> times,
                                    const py::dict& d) const
    {
        const std::vector<double> t = times_of(times);
        const CurveParams p = params_from_dict(d);
        PyLightCurveInfo out = records(t, p);
        if (source_ == "binary") {
            const std::vector<double> v2 = source_curve(t, second_source(p));
            const double fr = p.flux_ratio;
            const double inv = 1.0 / (1.0 + fr);
            for (std::size_t i = 0; i < out.magnifications.size(); ++i)
                out.magnifications[i] = (out.magnifications[i] + fr * v2[i]) * inv;
            for (std::size_t i = 0; i < out.finite_source_magnifications.size(); ++i)
                out.finite_source_magnifications[i] = (out.finite_source_magnifications[i] + fr * v2[i]) * inv;
        }
        return out;
    }

    static py::array_t<double> as_array(const std::vector<double>& v)
    {
        const auto n = static_cast<py::ssize_t>(v.size());
        auto out = py::array_t<double>({n}, {static_cast<py::ssize_t>(sizeof(double))});
        std::copy(v.begin(), v.end(), out.mutable_data());
        return out;
    }

    py::array_t<double> evaluate(py::array_t<double, py::array::c_style | py::array::forcecast> times,
                                 double t0, double tE, double u0, double alpha, double s, double q,
                                 double rho, double piEN, double piEE, double g1, double g2,
                                 double g3, double lom_szs, double lom_ar) const
    {
        return as_array(source_curve(times_of(times), positional(t0, tE, u0, alpha, s, q, rho, piEN, piEE, g1, g2,
                                                                 g3, lom_szs, lom_ar)));
    }

    // light_curve_from_dict (python/lcbinint_pybind.cpp:2325-2356): a binary
    // source is (A1 + f A2) / (1 + f), computed as (A1 + f A2) (1 / (1 + f))
    py::array_t<double> from_dict(py::array_t<double, py::array::c_style | py::array::forcecast> times,
                                  const py::dict& d) const
    {
        const std::vector<double> t = times_of(times);
        const CurveParams p = params_from_dict(d);
        std::vector<double> v = source_curve(t, p);
        if (source_ == "binary") {
            const std::vector<double> v2 = source_curve(t, second_source(p));
            const double fr = p.flux_ratio;
            const double inv = 1.0 / (1.0 + fr);
            for (std::size_t i = 0; i < v.size(); ++i) v[i] = (v[i] + fr * v2[i]) * inv;
        }
        return as_array(v);
    }

    // magnification_from_dict (python/lcbinint_pybind.cpp:2643-2655): one
    // epoch; a binary source is (A1 + f A2) / (1 + f)
    double magnification_from_dict(double time, const py::dict& d) const
    {
        const std::vector<double> t{time};
        const CurveParams p = params_from_dict(d);
        const double a1 = source_curve(t, p)[0];
        if (source_ != "binary") return a1;
        const double a2 = source_curve(t, second_source(p))[0];
        const double fr = p.flux_ratio;
        return (a1 + fr * a2) / (1.0 + fr);
    }

    // source_for_geometry / dynamic_source_trajectory (python/lcbinint_pybind.cpp:1819-1880)
    PyGeometryCurve trajectory(const std::vector<double>& times, double t0, double tE, double u0, double alpha,
                               double s, double q, double piEN, double piEE, double g1, double g2, double g3,
                               double lom_szs, double lom_ar) const
    {
        // the device trajectory is the VBM one (the legacy frame's formulas are not built)
        if (options_.vbm_compatible == 0) throw std::runtime_error("unsupported");
        PyGeometryCurve out;
        out.times = times;
        out.x.resize(times.size());
        out.y.resize(times.size());
        const lcbg_curve cv{t0, tE, u0, alpha, s, q, 0.0, ld_.c, ld_.d};
        const lcbg_dynamics dyn{piEN, piEE, event_.ra, event_.dec, event_.tfix, g1, g2, g3, lom_szs, lom_ar};
        lcbg_status st;
        {
            py::gil_scoped_release release;
            st = lcbg_source_trajectory(&cv, &dyn, orbit_, times.data(), static_cast<int64_t>(times.size()),
                                        out.x.data(), out.y.data());
        }
        raise_status(st);
        return out;
    }

    void require_parallax() const
    {
        if (!parallax_)
            throw std::invalid_argument(
                "piEN/piEE require LightCurve(parallax=True); the current trajectory model is static");
    }

    void require_binary_positional_api() const
    {
        if (lens_ != "binary_lens")
            throw std::invalid_argument("triple_lens requires a parameter dictionary with q2, sep2, and ang");
    }

    PyGeometryCurve source_trajectory_from_dict(const std::vector<double>& times, const py::dict& p) const
    {
        if (lens_ != "binary_lens") throw std::runtime_error("unsupported");
        return trajectory(times, dict_get(p, "t0", 0.0), dict_get(p, "tE", 1.0), dict_get(p, "u0", 0.0),
                          dict_get(p, "alpha", 0.0), dict_get(p, "s", 1.0), dict_get(p, "q", 1.0),
                          dict_get(p, "piEN", 0.0), dict_get(p, "piEE", 0.0), dict_get(p, "g1", 0.0),
                          dict_get(p, "g2", 0.0), dict_get(p, "g3", 0.0), dict_get(p, "lom_szs", 0.0),
                          dict_get(p, "lom_ar", 1.0));
    }

    PyGeometryCurve source_trajectory_static(const std::vector<double>& times, double t0, double tE, double u0,
                                             double alpha, double s, double q, double g1, double g2, double g3,
                                             double lom_szs, double lom_ar) const
    {
        require_binary_positional_api();
        return trajectory(times, t0, tE, u0, alpha, s, q, 0.0, 0.0, g1, g2, g3, lom_szs, lom_ar);
    }

    PyGeometryCurve source_trajectory(const std::vector<double>& times, double t0, double tE, double u0,
                                      double alpha, double s, double q, double piEN, double piEE, double g1,
                                      double g2, double g3, double lom_szs, double lom_ar) const
    {
        require_binary_positional_api();
        require_parallax();
        return trajectory(times, t0, tE, u0, alpha, s, q, piEN, piEE, g1, g2, g3, lom_szs, lom_ar);
    }

    // geometry_branches (python/lcbinint_pybind.cpp:1882-1926)
    PyGeometryBranches branches(double s, double q, int n_points, double time, double t0, double tE, double u0,
                                double alpha, double g1, double g2, double g3, double lom_szs, double lom_ar,
                                bool to_source) const
    {
        if (n_points < 8) throw std::invalid_argument("n_points must be at least 8");
        double separation = s;
        lcbg_status st = LCBG_OK;
        if (orbit_ != STATIC && std::isfinite(time)) {
            // the orbit's separation at `time` (host orbit constants, as the reference's orbit cache)
            const lcbg_curve cv{t0, tE, u0, alpha, s, q, 0.0, ld_.c, ld_.d};
            const lcbg_dynamics dyn{0.0, 0.0, event_.ra, event_.dec, event_.tfix, g1, g2, g3, lom_szs, lom_ar};
            double xys[3];
            st = lcbg_dynamic_sources(&cv, &dyn, orbit_, &time, 1, xys);
            separation = xys[2];
        }
        raise_status(st);
        const std::size_t n = static_cast<std::size_t>(n_points);
        std::vector<double> x(4 * n), y(4 * n);
        {
            py::gil_scoped_release release;
            st = lcbg_geometry_branches(separation, q, n_points, to_source ? 1 : 0, options_.vbm_compatible, x.data(),
                                        y.data());
        }
        raise_status(st);
        PyGeometryBranches out;
        for (std::size_t b = 0; b < 4; ++b) {
            out.x.emplace_back(x.begin() + b * n, x.begin() + (b + 1) * n);
            out.y.emplace_back(y.begin() + b * n, y.begin() + (b + 1) * n);
        }
        return out;
    }

    PyGeometryBranches caustics(double s, double q, int n_points, double time, double t0, double tE, double u0,
                                double alpha, double g1, double g2, double g3, double lom_szs, double lom_ar) const
    {
        require_binary_positional_api();
        return branches(s, q, n_points, time, t0, tE, u0, alpha, g1, g2, g3, lom_szs, lom_ar, true);
    }

    PyGeometryBranches critical_curves(double s, double q, int n_points, double time, double t0, double tE,
                                       double u0, double alpha, double g1, double g2, double g3, double lom_szs,
                                       double lom_ar) const
    {
        require_binary_positional_api();
        return branches(s, q, n_points, time, t0, tE, u0, alpha, g1, g2, g3, lom_szs, lom_ar, false);
    }

    // caustics_from_dict (python/lcbinint_pybind.cpp:2539-2556): the lens at
    // its reference separation
    PyGeometryBranches caustics_from_dict(const py::dict& p, int n_points) const
    {
        if (lens_ != "binary_lens") throw std::runtime_error("unsupported");
        return branches(dict_get(p, "s", 1.0), dict_get(p, "q", 1.0), n_points, kNaN, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0,
                        0.0, 0.0, 1.0, true);
    }

    const PyEvent& event() const { return event_; }

    const std::string& lens() const { return lens_; }
    const std::string& source() const { return source_; }
    const PyOptions& options() const { return options_; }
    const PyLimbDarkening& limb_darkening() const { return ld_; }
    bool parallax() const { return parallax_; }
    bool terrestrial_parallax() const { return terrestrial_; }
    int orbital_motion_mode() const { return orbit_; }

private:
    std::string lens_, source_;
    PyEvent event_;
    PyOptions options_;
    PyLimbDarkening ld_;
    int orbit_;
    bool parallax_, terrestrial_;
    std::shared_ptr<DynHandle> dyn_;

    // the kept one-curve handle, holding `cv` under `opts`; dyn_->mu held
    lcbg_status dyn_batch(const lcbg_curve& cv, const lcbg_options& opts, lcbg_batch** out) const
    {
        DynHandle& h = *dyn_;
        if (h.batch && std::memcmp(&h.options, &opts, sizeof(lcbg_options)) != 0) {
            lcbg_batch_destroy(h.batch);
            h.batch = nullptr;
        }
        lcbg_status st = LCBG_OK;
        if (!h.batch) {
            st = lcbg_batch_create(&cv, 1, &opts, &h.batch);
            h.options = opts;
        } else {
            st = lcbg_batch_update_curves(h.batch, &cv, nullptr);
        }
        *out = h.batch;
        return st;
    }
};

}  // namespace

PYBIND11_MODULE(_lcbinint, m)
{
    m.doc() = "MI355X (gfx950) binary-lens light curves behind the lcbinint Python API";

    py::enum_<lcbg_method>(m, "FiniteSourceMethod")
        .value("point_source", LCBG_METHOD_POINT_SOURCE)
        .value("hexadecapole", LCBG_METHOD_HEXADECAPOLE)
        .value("inverse_ray_cartesian", LCBG_METHOD_INVERSE_RAY_CARTESIAN)
        .value("inverse_ray_polar", LCBG_METHOD_INVERSE_RAY_POLAR);

    py::enum_<OrbitalMotionMode>(m, "OrbitalMotionMode")
        .value("STATIC", STATIC)
        .value("CIRCULAR", CIRCULAR)
        .value("KEPLER", KEPLER)
        .export_values();

    py::class_<PyLimbDarkening>(m, "LimbDarkening")
        .def("__repr__", [](const PyLimbDarkening& ld) {
            return "LimbDarkening(c=" + std::to_string(ld.c) + ", d=" + std::to_string(ld.d) + ")";
        })
        .def(py::init([](double c, double d) { return PyLimbDarkening{c, d}; }), py::arg("c") = 0.0,
             py::arg("d") = 0.0)
        .def_static("none", []() { return PyLimbDarkening{}; })
        .def_static("linear", [](double c) { return PyLimbDarkening{c, 0.0}; }, py::arg("c"))
        .def_static("quadratic", [](double c, double d) { return PyLimbDarkening{c, d}; }, py::arg("c"),
                    py::arg("d"))
        .def_readwrite("c", &PyLimbDarkening::c)
        .def_readwrite("d", &PyLimbDarkening::d);

    py::class_<PyOptions>(m, "Options")
        .def(py::init([](int source_bins, const py::object& nbin, const std::string& inverse_ray_grid,
                         int caustic_bins, double grid_ratio, const py::object& polar_nbin,
                         const py::object& polar_source_bins, const py::object& polar_grid_ratio,
                         double point_source_threshold, double hexadecapole_threshold,
                         double adaptive_hex_threshold, int adaptive_source_bins, int max_source_bins,
                         double finite_source_tol, double finite_source_reltol, double tol,
                         double reltol, const std::string& coordinates, double hex_tol,
                         const std::string& param_type, const std::string& xallarap_param_type) {
                 PyOptions p;
                 if (!param_type.empty() && param_type != "auto")
                     apply_param_type(p, param_type);
                 else
                     apply_param_type(p, coordinates);
                 if (xallarap_param_type != "none" && xallarap_param_type != "angular_velocity" &&
                     xallarap_param_type != "orbital_elements" && !xallarap_param_type.empty())
                     throw std::invalid_argument(
                         "xallarap_param_type must be 'none', 'angular_velocity', or 'orbital_elements'");
                 p.xallarap = xallarap_param_type.empty() ? "none" : xallarap_param_type;
                 apply_grid(p, inverse_ray_grid);
                 p.o.source_bins = int_or(nbin, source_bins);
                 p.o.caustic_bins = caustic_bins;
                 p.o.grid_ratio = grid_ratio;
                 p.o.polar_source_bins = int_or(polar_nbin, int_or(polar_source_bins, 0));
                 p.o.polar_grid_ratio = double_or(polar_grid_ratio, 0.0);
                 p.o.point_source_threshold = point_source_threshold;
                 p.o.hexadecapole_threshold = hexadecapole_threshold;
                 p.o.adaptive_hex_threshold = std::isnan(hex_tol) ? adaptive_hex_threshold : hex_tol;
                 p.max_source_bins = max_source_bins;
                 p.finite_source_tol = std::isnan(tol) ? finite_source_tol : tol;
                 p.finite_source_reltol = std::isnan(reltol) ? finite_source_reltol : reltol;
                 p.adaptive_source_bins = adaptive_source_bins < 0 ? 0 : adaptive_source_bins;
                 return p;
             }),
             py::arg("source_bins") = 50, py::arg("nbin") = py::none(),
             py::arg("inverse_ray_grid") = "auto", py::arg("caustic_bins") = 1400,
             py::arg("grid_ratio") = 4.0, py::arg("polar_nbin") = py::none(),
             py::arg("polar_source_bins") = py::none(), py::arg("polar_grid_ratio") = py::none(),
             py::arg("point_source_threshold") = 20.0, py::arg("hexadecapole_threshold") = 3.0,
             py::arg("adaptive_hex_threshold") = 0.001, py::arg("adaptive_source_bins") = -1,
             py::arg("max_source_bins") = 400, py::arg("finite_source_tol") = 0.0,
             py::arg("finite_source_reltol") = 0.0, py::arg("tol") = kNaN, py::arg("reltol") = kNaN,
             py::arg("coordinates") = "auto", py::arg("hex_tol") = kNaN,
             py::arg("param_type") = "auto", py::arg("xallarap_param_type") = "none")
        .def_property("source_bins", [](const PyOptions& p) { return p.o.source_bins; },
                      [](PyOptions& p, int v) { p.o.source_bins = v; })
        .def_property("nbin", [](const PyOptions& p) { return p.o.source_bins; },
                      [](PyOptions& p, int v) { p.o.source_bins = v; })
        .def_property("inverse_ray_grid", &grid_name, &apply_grid)
        .def_property_readonly("_mode", [](const PyOptions& p) { return p.o.mode; })
        .def_property("caustic_bins", [](const PyOptions& p) { return p.o.caustic_bins; },
                      [](PyOptions& p, int v) { p.o.caustic_bins = v; })
        .def_property("grid_ratio", [](const PyOptions& p) { return p.o.grid_ratio; },
                      [](PyOptions& p, double v) { p.o.grid_ratio = v; })
        .def_property("polar_nbin", [](const PyOptions& p) { return positive_int(p.o.polar_source_bins); },
                      [](PyOptions& p, const py::object& v) { p.o.polar_source_bins = int_or(v, 0); })
        .def_property("polar_source_bins",
                      [](const PyOptions& p) { return positive_int(p.o.polar_source_bins); },
                      [](PyOptions& p, const py::object& v) { p.o.polar_source_bins = int_or(v, 0); })
        .def_property("polar_grid_ratio",
                      [](const PyOptions& p) { return positive_double(p.o.polar_grid_ratio); },
                      [](PyOptions& p, const py::object& v) { p.o.polar_grid_ratio = double_or(v, 0.0); })
        .def_property("point_source_threshold", [](const PyOptions& p) { return p.o.point_source_threshold; },
                      [](PyOptions& p, double v) { p.o.point_source_threshold = v; })
        .def_property("hexadecapole_threshold", [](const PyOptions& p) { return p.o.hexadecapole_threshold; },
                      [](PyOptions& p, double v) { p.o.hexadecapole_threshold = v; })
        .def_property("adaptive_hex_threshold", [](const PyOptions& p) { return p.o.adaptive_hex_threshold; },
                      [](PyOptions& p, double v) { p.o.adaptive_hex_threshold = v; })
        .def_property("hex_tol", [](const PyOptions& p) { return p.o.adaptive_hex_threshold; },
                      [](PyOptions& p, double v) { p.o.adaptive_hex_threshold = v; })
        .def_readwrite("adaptive_source_bins", &PyOptions::adaptive_source_bins)
        .def_readwrite("max_source_bins", &PyOptions::max_source_bins)
        .def_readwrite("finite_source_tol", &PyOptions::finite_source_tol)
        .def_readwrite("finite_source_reltol", &PyOptions::finite_source_reltol)
        .def_property("tol", [](const PyOptions& p) { return p.finite_source_tol; },
                      [](PyOptions& p, double v) { p.finite_source_tol = v; })
        .def_property("reltol", [](const PyOptions& p) { return p.finite_source_reltol; },
                      [](PyOptions& p, double v) { p.finite_source_reltol = v; })
        .def_property("coordinates", &coordinate_name, &apply_param_type)
        .def_property("param_type", &coordinate_name, &apply_param_type)
        .def_property_readonly("xallarap_param_type", [](const PyOptions& p) { return p.xallarap; });

    py::class_<PyLightCurveInfo>(m, "LightCurveInfo")
        .def_readonly("times", &PyLightCurveInfo::times)
        .def_readonly("magnifications", &PyLightCurveInfo::magnifications)
        .def_readonly("point_source_magnifications", &PyLightCurveInfo::point_source_magnifications)
        .def_readonly("finite_source_magnifications", &PyLightCurveInfo::finite_source_magnifications)
        .def_readonly("finite_source_error_estimates", &PyLightCurveInfo::finite_source_error_estimates)
        .def_readonly("source_x", &PyLightCurveInfo::source_x)
        .def_readonly("source_y", &PyLightCurveInfo::source_y)
        .def_readonly("image_counts", &PyLightCurveInfo::image_counts)
        .def_readonly("finite_source_methods", &PyLightCurveInfo::finite_source_methods)
        .def_readonly("finite_source_method_names", &PyLightCurveInfo::finite_source_method_names)
        .def_readonly("finite_source_refinement_levels", &PyLightCurveInfo::finite_source_refinement_levels)
        .def_readonly("finite_source_converged", &PyLightCurveInfo::finite_source_converged)
        .def_property_readonly("all_converged",
                               [](const PyLightCurveInfo& c) {
                                   for (bool v : c.finite_source_converged)
                                       if (!v) return false;
                                   return true;
                               })
        .def_property_readonly("unconverged_indices", [](const PyLightCurveInfo& c) {
            std::vector<int> idx;
            for (std::size_t i = 0; i < c.finite_source_converged.size(); ++i)
                if (!c.finite_source_converged[i]) idx.push_back(static_cast<int>(i));
            return idx;
        });

    py::class_<PyEvent>(m, "EventCoordinates")
        .def(py::init([](double ra, double dec, double tfix, double obs_lat, double obs_lon) {
                 return PyEvent{ra, dec, tfix, obs_lat, obs_lon};
             }),
             py::arg("ra") = 0.0, py::arg("dec") = 0.0, py::arg("tfix") = 0.0, py::arg("obs_lat") = 0.0,
             py::arg("obs_lon") = 0.0)
        .def("__repr__", [](const PyEvent& e) {
            return "EventCoordinates(ra=" + std::to_string(e.ra) + ", dec=" + std::to_string(e.dec) +
                   ", tfix=" + std::to_string(e.tfix) + ", obs_lat=" + std::to_string(e.obs_lat) +
                   ", obs_lon=" + std::to_string(e.obs_lon) + ")";
        })
        .def_readwrite("ra", &PyEvent::ra)
        .def_readwrite("dec", &PyEvent::dec)
        .def_readwrite("tfix", &PyEvent::tfix)
        .def_readwrite("obs_lat", &PyEvent::obs_lat)
        .def_readwrite("obs_lon", &PyEvent::obs_lon);

    py::class_<PyGeometryCurve>(m, "GeometryCurve")
        .def_readonly("times", &PyGeometryCurve::times)
        .def_readonly("x", &PyGeometryCurve::x)
        .def_readonly("y", &PyGeometryCurve::y);

    py::class_<PyGeometryBranches>(m, "GeometryBranches")
        .def_readonly("x", &PyGeometryBranches::x)
        .def_readonly("y", &PyGeometryBranches::y);

    py::class_<PyLightCurve>(m, "LightCurve")
        .def(py::init([](const std::string& lens, const std::string& source, const PyEvent& event,
                         const PyOptions& options, const PyLimbDarkening& ld, OrbitalMotionMode orbital_motion_mode,
                         bool parallax, bool terrestrial_parallax) {
                 return PyLightCurve(lens, source, event, options, ld, static_cast<int>(orbital_motion_mode),
                                     parallax, terrestrial_parallax);
             }),
             py::kw_only(), py::arg("lens") = "binary_lens", py::arg("source") = "single",
             py::arg("event") = PyEvent{}, py::arg("options") = PyOptions{},
             py::arg("limb_darkening") = PyLimbDarkening{}, py::arg("orbital_motion_mode") = STATIC,
             py::arg("parallax") = false, py::arg("terrestrial_parallax") = false)
        .def_property_readonly("lens", &PyLightCurve::lens)
        .def_property_readonly("event", &PyLightCurve::event)
        .def_property_readonly("source", &PyLightCurve::source)
        .def_property_readonly("options", &PyLightCurve::options)
        .def_property_readonly("limb_darkening", &PyLightCurve::limb_darkening)
        .def_property_readonly("orbital_motion_mode", &PyLightCurve::orbital_motion_mode)
        .def_property_readonly("parallax", &PyLightCurve::parallax)
        .def_property_readonly("terrestrial_parallax", &PyLightCurve::terrestrial_parallax)
        .def("__call__", &PyLightCurve::from_dict, py::arg("times"), py::arg("params"))
        .def("__call__", &PyLightCurve::evaluate, py::arg("times"), py::kw_only(),
             py::arg("t0") = 0.0, py::arg("tE") = 1.0, py::arg("u0") = 0.0, py::arg("alpha") = 0.0,
             py::arg("s") = 1.0, py::arg("q") = 1.0, py::arg("rho") = 0.0, py::arg("piEN") = 0.0,
             py::arg("piEE") = 0.0, py::arg("g1") = 0.0, py::arg("g2") = 0.0, py::arg("g3") = 0.0,
             py::arg("lom_szs") = 0.0, py::arg("lom_ar") = 1.0)
        .def("light_curve", &PyLightCurve::from_dict, py::arg("times"), py::arg("params"))
        .def("magnification", &PyLightCurve::magnification_from_dict, py::arg("time"), py::arg("params"))
        .def("info", &PyLightCurve::info_from_dict, py::arg("times"), py::arg("params"))
        .def("info", &PyLightCurve::info, py::arg("times"), py::kw_only(), py::arg("t0") = 0.0,
             py::arg("tE") = 1.0, py::arg("u0") = 0.0, py::arg("alpha") = 0.0, py::arg("s") = 1.0,
             py::arg("q") = 1.0, py::arg("rho") = 0.0, py::arg("piEN") = 0.0, py::arg("piEE") = 0.0,
             py::arg("g1") = 0.0, py::arg("g2") = 0.0, py::arg("g3") = 0.0, py::arg("lom_szs") = 0.0,
             py::arg("lom_ar") = 1.0)
        // geometry helpers (python/lcbinint_pybind.cpp:3292-3364)
        .def("source_trajectory", &PyLightCurve::source_trajectory_from_dict, py::arg("times"), py::arg("params"))
        .def("source_trajectory", &PyLightCurve::source_trajectory_static, py::arg("times"), py::kw_only(),
             py::arg("t0") = 0.0, py::arg("tE") = 1.0, py::arg("u0") = 0.0, py::arg("alpha") = 0.0,
             py::arg("s") = 1.0, py::arg("q") = 1.0, py::arg("g1") = 0.0, py::arg("g2") = 0.0,
             py::arg("g3") = 0.0, py::arg("lom_szs") = 0.0, py::arg("lom_ar") = 1.0)
        .def("source_trajectory", &PyLightCurve::source_trajectory, py::arg("times"), py::kw_only(),
             py::arg("t0") = 0.0, py::arg("tE") = 1.0, py::arg("u0") = 0.0, py::arg("alpha") = 0.0,
             py::arg("s") = 1.0, py::arg("q") = 1.0, py::arg("piEN") = 0.0, py::arg("piEE") = 0.0,
             py::arg("g1") = 0.0, py::arg("g2") = 0.0, py::arg("g3") = 0.0, py::arg("lom_szs") = 0.0,
             py::arg("lom_ar") = 1.0)
        .def("caustics", &PyLightCurve::caustics_from_dict, py::arg("params"), py::arg("n_points") = 1000)
        .def("caustics", &PyLightCurve::caustics, py::kw_only(), py::arg("s") = 1.0, py::arg("q") = 1.0,
             py::arg("n_points") = 1400, py::arg("time") = kNaN, py::arg("t0") = 0.0, py::arg("tE") = 1.0,
             py::arg("u0") = 0.0, py::arg("alpha") = 0.0, py::arg("g1") = 0.0, py::arg("g2") = 0.0,
             py::arg("g3") = 0.0, py::arg("lom_szs") = 0.0, py::arg("lom_ar") = 1.0)
        .def("critical_curves", &PyLightCurve::critical_curves, py::kw_only(), py::arg("s") = 1.0,
             py::arg("q") = 1.0, py::arg("n_points") = 1400, py::arg("time") = kNaN, py::arg("t0") = 0.0,
             py::arg("tE") = 1.0, py::arg("u0") = 0.0, py::arg("alpha") = 0.0, py::arg("g1") = 0.0,
             py::arg("g2") = 0.0, py::arg("g3") = 0.0, py::arg("lom_szs") = 0.0, py::arg("lom_ar") = 1.0);

    // python/lcbinint_pybind.cpp:3405
    m.attr("LightCurveFunc") = m.attr("LightCurve");
}
