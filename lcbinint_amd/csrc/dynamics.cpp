// Source trajectories with annual parallax and binary orbital motion, on the
// host.  Compiled by g++ (see build.py) so that cos, sin, atan2, acos and sqrt
// are the glibc functions the reference calls: the source position and the
// projected separation of every epoch are then bitwise the reference's, and the
// device lattices that follow are the reference's lattices.
//
//  * Earth ephemeris: the Horizons vector-table CSV of load_earth_ephemeris
//    (src/lcbinint/model/trajectory.cpp:110-183), read once per process from
//    LCBININT_EARTH_EPHEMERIS or from lcbg_load_earth_ephemeris, linearly
//    interpolated (interpolate_earth_state :185-209);
//  * EarthOrbitalParallaxProjector (:221-285): sky basis (:80-98), five
//    light-time iterations, offsets relative to position + velocity * dt at
//    the reference time tfix (or t0), JD offset for reduced dates (:211-214);
//  * orbital_state (src/lcbinint/model/orbital_motion.cpp:70-199): circular
//    and Keplerian orbits;
//  * the source rotation into the instantaneous binary frame of
//    LensModel::magnification (src/lcbinint/model/lens_model.cpp:84-104,
//    164-178), VBM convention.
#include "lcbinint_amd.h"
#include "runtime.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace lcbg {
namespace {

constexpr double kPiHost = 3.14159265358979323846;
constexpr double kDeg = kPiHost / 180.0;
constexpr double kReducedJdOrigin = 2450000.0;
constexpr double kAuLightDays = 0.005775518331436995;
constexpr double kOrbitEps = 1.0e-8;
constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

using V3 = std::array<double, 3>;

struct EarthRow {
    double t;
    V3 p, v;
};

V3 v_add(V3 a, V3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
V3 v_sub(V3 a, V3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 v_scale(V3 a, double f) { return {a[0] * f, a[1] * f, a[2] * f}; }
double v_dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 v_cross(V3 a, V3 b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

// -------------------------------------------------------------- ephemeris

std::mutex g_earth_mu;
std::vector<EarthRow> g_earth;
bool g_earth_loaded = false;

std::string trimmed(const std::string& s)
{
    const auto b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    const auto e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

// The numeric fields of one table row: non-empty comma-separated tokens, the
// second one (the calendar date) skipped, tokens that do not parse dropped.
std::vector<double> row_fields(const std::string& line)
{
    std::vector<double> out;
    std::stringstream ss(line);
    std::string tok;
    int column = 0;
    while (std::getline(ss, tok, ',')) {
        tok = trimmed(tok);
        if (tok.empty()) continue;
        if (column++ == 1) continue;
        try {
            out.push_back(std::stod(tok));
        } catch (const std::exception&) {
        }
    }
    return out;
}

std::vector<EarthRow> read_table(const char* path)
{
    std::vector<EarthRow> rows;
    if (!path || !path[0]) return rows;
    std::ifstream in(path);
    if (!in) return rows;
    bool body = false;
    std::string line;
    while (std::getline(in, line)) {
        const std::string t = trimmed(line);
        if (!body) {
            body = t.rfind("$$SOE", 0) == 0;
            continue;
        }
        if (t.rfind("$$EOE", 0) == 0) break;
        const auto f = row_fields(t);
        if (f.size() < 7) continue;
        rows.push_back({f[0], {f[1], f[2], f[3]}, {f[4], f[5], f[6]}});
    }
    return rows;
}

const std::vector<EarthRow>& earth_table()
{
    std::lock_guard<std::mutex> lock(g_earth_mu);
    if (!g_earth_loaded) {
        g_earth = read_table(std::getenv("LCBININT_EARTH_EPHEMERIS"));
        g_earth_loaded = true;
    }
    return g_earth;
}

struct EarthAt {
    V3 p, v;
};

EarthAt earth_at(const std::vector<EarthRow>& rows, double t)
{
    if (rows.size() < 2 || t < rows.front().t || t > rows.back().t)
        return {{kNaN, kNaN, kNaN}, {kNaN, kNaN, kNaN}};
    auto hi = std::upper_bound(rows.begin(), rows.end(), t,
                               [](double value, const EarthRow& r) { return value < r.t; });
    if (hi == rows.begin()) hi = rows.begin() + 1;
    if (hi == rows.end()) hi = rows.end() - 1;  // t equals the last row's time
    const auto lo = hi - 1;
    const double dt = hi->t - lo->t;
    const double w = dt != 0.0 ? (t - lo->t) / dt : 0.0;
    return {v_add(lo->p, v_scale(v_sub(hi->p, lo->p), w)), v_add(lo->v, v_scale(v_sub(hi->v, lo->v), w))};
}

V3 normalized(V3 a)
{
    const double n = std::sqrt(v_dot(a, a));
    if (n == 0.0 || !std::isfinite(n)) return {kNaN, kNaN, kNaN};
    return v_scale(a, 1.0 / n);
}

// Annual-parallax offsets (tau, beta) of one event direction and reference time.
class Parallax {
public:
    Parallax(const std::vector<EarthRow>& rows, double ra_deg, double dec_deg, double ref_time)
        : rows_(rows)
    {
        offset_ = ref_time < kReducedJdOrigin ? kReducedJdOrigin : 0.0;
        eph_ref_ = ref_time + offset_;
        const double ra = ra_deg * kDeg;
        const double dec = dec_deg * kDeg;
        const double cos_dec = std::cos(dec);
        event_ = {std::cos(ra) * cos_dec, std::sin(ra) * cos_dec, std::sin(dec)};
        east_ = normalized(v_cross({0.0, 0.0, 1.0}, event_));
        north_ = v_cross(event_, east_);
        const EarthAt r = earth_at(rows_, emit_time(eph_ref_));
        ref_pn_ = -v_dot(r.p, north_);
        ref_pe_ = -v_dot(r.p, east_);
        ref_vn_ = -v_dot(r.v, north_);
        ref_ve_ = -v_dot(r.v, east_);
    }

    void offset(double t, double pi_n, double pi_e, double* dtau, double* dbeta) const
    {
        const double eph = t + offset_;
        const EarthAt s = earth_at(rows_, emit_time(eph));
        const double pn = -v_dot(s.p, north_);
        const double pe = -v_dot(s.p, east_);
        const double dt = eph - eph_ref_;
        const double dn = pn - ref_pn_ - ref_vn_ * dt;
        const double de = pe - ref_pe_ - ref_ve_ * dt;
        *dtau = pi_n * dn + pi_e * de;
        *dbeta = -pi_e * dn + pi_n * de;
    }

private:
    double emit_time(double obs) const
    {
        double emit = obs;
        for (int i = 0; i < 5; ++i) {
            const EarthAt s = earth_at(rows_, emit);
            emit = obs - v_dot(s.p, event_) * kAuLightDays;
        }
        return emit;
    }

    const std::vector<EarthRow>& rows_;
    double offset_ = 0.0, eph_ref_ = 0.0;
    V3 event_{}, east_{}, north_{};
    double ref_pn_ = 0.0, ref_pe_ = 0.0, ref_vn_ = 0.0, ref_ve_ = 0.0;
};

// ----------------------------------------------------------------- orbits

double kepler_e(double m, double e)
{
    double E = m + e * std::sin(m);
    for (int i = 0; i < 10; ++i) {
        const double f = E - e * std::sin(E) - m;
        const double fp = 1.0 - e * std::cos(E);
        const double d = f / fp;
        E -= d;
        if (std::abs(d) <= 1.0e-13) break;
    }
    return E;
}

// Per-curve constants of orbital_state; at() gives (s, angle) at a time.
struct Orbit {
    int mode = LCBG_ORBIT_STATIC;
    double sep = 1.0, angle = 0.0, t_ref = 0.0;
    // circular
    double w_orb = 0.0, phi0 = 0.0, c_inc = 1.0, s_true = 0.0, c_om = 0.0, s_om = 0.0;
    // Keplerian
    double n = 0.0, ecc = 0.0, t_peri = 0.0, a = 0.0, one_minus_e2 = 0.0;
    V3 xa{}, ya{};

    Orbit(int m, double s, double alpha, const lcbg_dynamics& d, double t0)
        : mode(m), sep(s), angle(alpha)
    {
        t_ref = d.tfix != 0.0 ? d.tfix : t0;
        if (mode == LCBG_ORBIT_CIRCULAR) {
            const double w13_sq = d.g1 * d.g1 + d.g3 * d.g3;
            const double w13 = std::sqrt(w13_sq);
            const double w123 = std::sqrt(w13_sq + d.g2 * d.g2);
            w_orb = d.g2;
            double inc = 0.0;
            if (w13 > kOrbitEps) {
                const double w3 = d.g3 > kOrbitEps ? d.g3 : kOrbitEps;
                w_orb = w3 * w123 / w13;
                inc = std::acos(std::clamp(d.g2 * w3 / (w13 * w123), -1.0, 1.0));
                phi0 = std::atan2(-d.g1 * w123, w3 * w13);
            }
            const double c_phi0 = std::cos(phi0);
            const double s_phi0 = std::sin(phi0);
            c_inc = std::cos(inc);
            const double den0 = std::sqrt(c_phi0 * c_phi0 + c_inc * c_inc * s_phi0 * s_phi0);
            s_true = sep / den0;
            const double c_a = std::cos(angle);
            const double s_a = std::sin(angle);
            c_om = (c_phi0 * c_a + c_inc * s_a * s_phi0) / den0;
            s_om = (c_phi0 * s_a - c_inc * c_a * s_phi0) / den0;
        } else if (mode == LCBG_ORBIT_KEPLER) {
            const double ar = d.lom_ar + kOrbitEps;
            const double szs = d.lom_szs;
            const double smix = 1.0 + szs * szs;
            const double sqsmix = std::sqrt(smix);
            const double w11 = d.g1 * d.g1, w22 = d.g2 * d.g2, w33 = d.g3 * d.g3;
            const double w12 = w11 + w22;
            const double wt2 = w12 + w33;
            const double arm1 = ar - 1.0;
            const double arm2 = 2.0 * ar - 1.0;
            n = std::sqrt(wt2 / arm2 / smix) / ar;
            const V3 za = normalized({-szs * d.g2, szs * d.g1 - d.g3, d.g2});
            xa = {-ar * w11 + arm1 * w22 - arm2 * szs * d.g1 * d.g3 + arm1 * w33,
                  -arm2 * d.g2 * (d.g1 + szs * d.g3),
                  arm1 * szs * w12 - arm2 * d.g1 * d.g3 - ar * szs * w33};
            const double xn = std::sqrt(std::max(0.0, v_dot(xa, xa)));
            xa = v_scale(xa, 1.0 / xn);
            ecc = xn / (ar * sqsmix * wt2);
            ya = v_cross(za, xa);
            const double conu = (xa[0] + xa[2] * szs) / sqsmix;
            const double cos_e0 = std::clamp((conu + ecc) / (1.0 + ecc * conu), -1.0, 1.0);
            const double sign = (ya[0] + ya[2] * szs) > 0.0 ? 1.0 : -1.0;
            const double e0 = std::acos(cos_e0) * sign;
            const double sin_e0 = std::sqrt(std::max(0.0, 1.0 - cos_e0 * cos_e0)) * sign;
            t_peri = t_ref - (e0 - ecc * sin_e0) / n;
            a = ar * sep * sqsmix;
            one_minus_e2 = std::sqrt(std::max(0.0, 1.0 - ecc * ecc));
        }
    }

    void at(double t, double* s_out, double* angle_out) const
    {
        if (mode == LCBG_ORBIT_CIRCULAR) {
            const double phi = w_orb * (t - t_ref) + phi0;
            const double c_phi = std::cos(phi);
            const double s_phi = std::sin(phi);
            const double den = std::sqrt(c_phi * c_phi + c_inc * c_inc * s_phi * s_phi);
            const double sin_a = (c_phi * s_om + c_inc * s_phi * c_om) / den;
            const double cos_a = (c_phi * c_om - c_inc * s_phi * s_om) / den;
            *s_out = s_true * den;
            *angle_out = std::atan2(sin_a, cos_a);
        } else if (mode == LCBG_ORBIT_KEPLER) {
            const double E = kepler_e(n * (t - t_peri), ecc);
            const double cos_e = std::cos(E);
            const double sin_e = std::sin(E);
            const double r0 = a * (cos_e - ecc);
            const double r1 = a * one_minus_e2 * sin_e;
            const double x0 = r0 * xa[0] + r1 * ya[0];
            const double x1 = r0 * xa[1] + r1 * ya[1];
            *s_out = std::sqrt(x0 * x0 + x1 * x1);
            *angle_out = angle + std::atan2(x1, x0);
        } else {
            *s_out = sep;
            *angle_out = angle;
        }
    }
};

}  // namespace

// Source position in the instantaneous binary frame and s(t) of every epoch of
// one curve (VBM convention); false if an epoch's orbit is not finite.
bool dynamic_sources(const lcbg_curve& cv, const lcbg_dynamics& d, int orbit_mode, const double* t,
                     int64_t n, double* xys)
{
    const bool parallax = d.piEN != 0.0 || d.piEE != 0.0;
    const std::vector<EarthRow>* rows = parallax ? &earth_table() : nullptr;
    std::unique_ptr<Parallax> px;
    if (parallax) px.reset(new Parallax(*rows, d.ra, d.dec, d.tfix != 0.0 ? d.tfix : cv.t0));
    const Orbit orbit(orbit_mode, cv.s, cv.alpha, d, cv.t0);
    const double cos_t = std::cos(cv.alpha);
    const double sin_t = std::sin(cv.alpha);
    bool ok = true;
    for (int64_t i = 0; i < n; ++i) {
        const double time = t[i];
        double tn = (time - cv.t0) / cv.tE;
        double beta = cv.u0;
        if (parallax) {
            double dtau, dbeta;
            px->offset(time, d.piEN, d.piEE, &dtau, &dbeta);
            tn += dtau;
            beta += dbeta;
        }
        double x = tn * cos_t - beta * sin_t;
        double y = tn * sin_t + beta * cos_t;
        double s_t, ang;
        orbit.at(time, &s_t, &ang);
        if (!std::isfinite(s_t) || !std::isfinite(ang)) ok = false;
        if (orbit_mode != LCBG_ORBIT_STATIC) {
            double tau, b;
            if (!parallax) {
                tau = (time - cv.t0) / cv.tE;
                b = cv.u0;
            } else {
                tau = x * cos_t + y * sin_t;
                b = -x * sin_t + y * cos_t;
            }
            const double ca = std::cos(ang);
            const double sa = std::sin(ang);
            x = tau * ca - b * sa;
            y = b * ca + tau * sa;
        }
        xys[3 * i] = x;
        xys[3 * i + 1] = y;
        xys[3 * i + 2] = s_t;
    }
    return ok;
}

// The same for every curve of a batch, curves spread over host threads.
bool dynamic_sources_batch(const lcbg_curve* curves, const lcbg_dynamics* dyn, int32_t n_curves,
                           int orbit_mode, const int64_t* offsets, const double* times, double* xys)
{
    if (n_curves <= 0) return true;
    bool any_parallax = false;
    for (int32_t c = 0; c < n_curves; ++c) any_parallax |= dyn[c].piEN != 0.0 || dyn[c].piEE != 0.0;
    if (any_parallax) (void)earth_table();  // load once, outside the threads
    unsigned hw = std::thread::hardware_concurrency();
    const int n_threads = static_cast<int>(std::min<unsigned>(hw ? hw : 1, 16u));
    std::vector<char> ok(static_cast<std::size_t>(n_curves), 1);
    auto work = [&](int k) {
        for (int32_t c = k; c < n_curves; c += n_threads)
            ok[static_cast<std::size_t>(c)] = dynamic_sources(curves[c], dyn[c], orbit_mode, times + offsets[c],
                                                              offsets[c + 1] - offsets[c], xys + 3 * offsets[c]);
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < n_threads; ++k) pool.emplace_back(work, k);
    work(0);
    for (auto& th : pool) th.join();
    return std::all_of(ok.begin(), ok.end(), [](char v) { return v != 0; });
}

}  // namespace lcbg

extern "C" lcbg_status lcbg_load_earth_ephemeris(const char* path)
{
    using namespace lcbg;
    const char* p = path ? path : std::getenv("LCBININT_EARTH_EPHEMERIS");
    std::vector<EarthRow> rows = read_table(p);
    std::lock_guard<std::mutex> lock(g_earth_mu);
    g_earth.swap(rows);
    g_earth_loaded = true;
    return g_earth.size() >= 2 ? LCBG_OK : LCBG_INVALID_ARGUMENT;
}

extern "C" lcbg_status lcbg_dynamic_sources(const lcbg_curve* curve, const lcbg_dynamics* dyn,
                                            int32_t orbital_motion_mode, const double* times,
                                            int64_t n, double* xys)
{
    if (!curve || !dyn || (n > 0 && (!times || !xys)) || n < 0) return LCBG_INVALID_ARGUMENT;
    if (orbital_motion_mode < LCBG_ORBIT_STATIC || orbital_motion_mode > LCBG_ORBIT_KEPLER)
        return LCBG_UNSUPPORTED;
    return lcbg::dynamic_sources(*curve, *dyn, orbital_motion_mode, times, n, xys) ? LCBG_OK
                                                                                   : LCBG_NUMERICAL_ERROR;
}
