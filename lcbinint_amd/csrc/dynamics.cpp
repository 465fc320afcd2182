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
#include "traj.hpp"

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
int g_earth_generation = 0;  // tables read so far

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
        ++g_earth_generation;
    }
    return g_earth;
}

// Forward-mode derivatives: the value part of every operation is the plain
// double operation, so a Dual run yields the double run's values bitwise.
template <int N>
struct Dual {
    double v = 0.0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}  // NOLINT: constants promote
};

template <int N> Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r(a.v + b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N> Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r(a.v - b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N> Dual<N> operator-(const Dual<N>& a)
{
    Dual<N> r(-a.v);
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r(a.v * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r(a.v / b.v);
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
template <int N> Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> Dual<N> operator*(const Dual<N>& a, double b)
{
    Dual<N> r(a.v * b);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N> Dual<N> operator/(const Dual<N>& a, double b)
{
    Dual<N> r(a.v / b);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
    return r;
}
template <int N> Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }

inline double val(double x) { return x; }
template <int N> double val(const Dual<N>& x) { return x.v; }

inline double f_sqrt(double x) { return std::sqrt(x); }
inline double f_cos(double x) { return std::cos(x); }
inline double f_sin(double x) { return std::sin(x); }
inline double f_acos(double x) { return std::acos(x); }
inline double f_atan2(double y, double x) { return std::atan2(y, x); }
template <int N> Dual<N> f_sqrt(const Dual<N>& a)
{
    Dual<N> r(std::sqrt(a.v));
    for (int k = 0; k < N; ++k) r.d[k] = r.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
    return r;
}
template <int N> Dual<N> f_cos(const Dual<N>& a)
{
    Dual<N> r(std::cos(a.v));
    const double s = -std::sin(a.v);
    for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
    return r;
}
template <int N> Dual<N> f_sin(const Dual<N>& a)
{
    Dual<N> r(std::sin(a.v));
    const double c = std::cos(a.v);
    for (int k = 0; k < N; ++k) r.d[k] = c * a.d[k];
    return r;
}
template <int N> Dual<N> f_acos(const Dual<N>& a)
{
    Dual<N> r(std::acos(a.v));
    const double den = std::sqrt(std::max(0.0, 1.0 - a.v * a.v));
    for (int k = 0; k < N; ++k) r.d[k] = den > 0.0 ? -a.d[k] / den : 0.0;
    return r;
}
template <int N> Dual<N> f_atan2(const Dual<N>& y, const Dual<N>& x)
{
    Dual<N> r(std::atan2(y.v, x.v));
    const double den = x.v * x.v + y.v * y.v;
    for (int k = 0; k < N; ++k) r.d[k] = den > 0.0 ? (x.v * y.d[k] - y.v * x.d[k]) / den : 0.0;
    return r;
}
// max(0, a) and clamp(a, lo, hi) as std::max / std::clamp choose
template <class T> T f_max0(const T& a) { return val(a) < 0.0 ? T(0.0) : a; }
template <class T> T f_clamp(const T& a, double lo, double hi)
{
    return val(a) < lo ? T(lo) : (hi < val(a) ? T(hi) : a);
}

template <class T>
struct Earth3 {
    T p[3], v[3];
};

template <class T>
Earth3<T> earth_at(const std::vector<EarthRow>& rows, const T& t)
{
    Earth3<T> out;
    const double tv = val(t);
    if (rows.size() < 2 || tv < rows.front().t || tv > rows.back().t) {
        for (int i = 0; i < 3; ++i) out.p[i] = out.v[i] = T(kNaN);
        return out;
    }
    auto hi = std::upper_bound(rows.begin(), rows.end(), tv,
                               [](double value, const EarthRow& r) { return value < r.t; });
    if (hi == rows.begin()) hi = rows.begin() + 1;
    if (hi == rows.end()) hi = rows.end() - 1;  // t equals the last row's time
    const auto lo = hi - 1;
    const double dt = hi->t - lo->t;
    const T w = dt != 0.0 ? (t - lo->t) / dt : T(0.0);
    for (int i = 0; i < 3; ++i) {
        out.p[i] = lo->p[i] + (hi->p[i] - lo->p[i]) * w;
        out.v[i] = lo->v[i] + (hi->v[i] - lo->v[i]) * w;
    }
    return out;
}

template <class T>
T dot3(const T* a, const V3& b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

V3 normalized(V3 a)
{
    const double n = std::sqrt(v_dot(a, a));
    if (n == 0.0 || !std::isfinite(n)) return {kNaN, kNaN, kNaN};
    return v_scale(a, 1.0 / n);
}

// Annual-parallax offsets (tau, beta) of one event direction and reference time.
template <class T>
class Parallax {
public:
    Parallax(const std::vector<EarthRow>& rows, double ra_deg, double dec_deg, const T& ref_time)
        : rows_(rows)
    {
        offset_ = val(ref_time) < kReducedJdOrigin ? kReducedJdOrigin : 0.0;
        eph_ref_ = ref_time + offset_;
        const double ra = ra_deg * kDeg;
        const double dec = dec_deg * kDeg;
        const double cos_dec = std::cos(dec);
        event_ = {std::cos(ra) * cos_dec, std::sin(ra) * cos_dec, std::sin(dec)};
        east_ = normalized(v_cross({0.0, 0.0, 1.0}, event_));
        north_ = v_cross(event_, east_);
        const Earth3<T> r = earth_at(rows_, emit_time(eph_ref_));
        ref_pn_ = -dot3(r.p, north_);
        ref_pe_ = -dot3(r.p, east_);
        ref_vn_ = -dot3(r.v, north_);
        ref_ve_ = -dot3(r.v, east_);
    }

    void offset(double t, const T& pi_n, const T& pi_e, T* dtau, T* dbeta) const
    {
        const double eph = t + offset_;
        const Earth3<T> s = earth_at(rows_, emit_time(T(eph)));
        const T pn = -dot3(s.p, north_);
        const T pe = -dot3(s.p, east_);
        const T dt = eph - eph_ref_;
        const T dn = pn - ref_pn_ - ref_vn_ * dt;
        const T de = pe - ref_pe_ - ref_ve_ * dt;
        *dtau = pi_n * dn + pi_e * de;
        *dbeta = -pi_e * dn + pi_n * de;
    }

    T emit_time(const T& obs) const
    {
        T emit = obs;
        for (int i = 0; i < 5; ++i) {
            const Earth3<T> s = earth_at(rows_, emit);
            emit = obs - dot3(s.p, event_) * kAuLightDays;
        }
        return emit;
    }

    const std::vector<EarthRow>& rows_;
    double offset_ = 0.0;
    T eph_ref_{};
    V3 event_{}, east_{}, north_{};
    T ref_pn_{}, ref_pe_{}, ref_vn_{}, ref_ve_{};
};

// ----------------------------------------------------------------- orbits

template <class T>
T kepler_e(const T& m, const T& e)
{
    T E = m + e * f_sin(m);
    for (int i = 0; i < 10; ++i) {
        const T f = E - e * f_sin(E) - m;
        const T fp = 1.0 - e * f_cos(E);
        const T d = f / fp;
        E = E - d;
        if (std::abs(val(d)) <= 1.0e-13) break;
    }
    return E;
}

// The orbit parameters a gradient may follow: g1, g2 (g3, lom_szs, lom_ar are
// held fixed).
template <class T>
struct OrbitIn {
    T g1, g2;
    double g3, szs, ar_in;
};

// Per-curve constants of orbital_state; at() gives (s, angle) at a time.
template <class T>
struct Orbit {
    int mode = LCBG_ORBIT_STATIC;
    T sep{}, angle{}, t_ref{};
    // circular
    T w_orb{}, phi0{}, c_inc{}, s_true{}, c_om{}, s_om{};
    // Keplerian
    T n{}, ecc{}, t_peri{}, a{}, one_minus_e2{};
    T xa[3], ya[3];

    Orbit(int m, const T& s, const T& alpha, const OrbitIn<T>& d, const T& tref)
        : mode(m), sep(s), angle(alpha), t_ref(tref)
    {
        if (mode == LCBG_ORBIT_CIRCULAR) {
            const T w13_sq = d.g1 * d.g1 + d.g3 * d.g3;
            const T w13 = f_sqrt(w13_sq);
            const T w123 = f_sqrt(w13_sq + d.g2 * d.g2);
            w_orb = d.g2;
            phi0 = T(0.0);
            T inc(0.0);
            if (val(w13) > kOrbitEps) {
                const double w3 = d.g3 > kOrbitEps ? d.g3 : kOrbitEps;
                w_orb = w3 * w123 / w13;
                inc = f_acos(f_clamp(d.g2 * w3 / (w13 * w123), -1.0, 1.0));
                phi0 = f_atan2(-d.g1 * w123, w3 * w13);
            }
            const T c_phi0 = f_cos(phi0);
            const T s_phi0 = f_sin(phi0);
            c_inc = f_cos(inc);
            const T den0 = f_sqrt(c_phi0 * c_phi0 + c_inc * c_inc * s_phi0 * s_phi0);
            s_true = sep / den0;
            const T c_a = f_cos(angle);
            const T s_a = f_sin(angle);
            c_om = (c_phi0 * c_a + c_inc * s_a * s_phi0) / den0;
            s_om = (c_phi0 * s_a - c_inc * c_a * s_phi0) / den0;
        } else if (mode == LCBG_ORBIT_KEPLER) {
            const double ar = d.ar_in + kOrbitEps;
            const double szs = d.szs;
            const double smix = 1.0 + szs * szs;
            const double sqsmix = std::sqrt(smix);
            const T w11 = d.g1 * d.g1, w22 = d.g2 * d.g2;
            const double w33 = d.g3 * d.g3;
            const T w12 = w11 + w22;
            const T wt2 = w12 + w33;
            const double arm1 = ar - 1.0;
            const double arm2 = 2.0 * ar - 1.0;
            n = f_sqrt(wt2 / arm2 / smix) / ar;
            // z axis (normalized)
            T za[3] = {-szs * d.g2, szs * d.g1 - d.g3, d.g2};
            {
                const T zn = f_sqrt(za[0] * za[0] + za[1] * za[1] + za[2] * za[2]);
                if (val(zn) == 0.0 || !std::isfinite(val(zn))) {
                    for (auto& z : za) z = T(kNaN);
                } else {
                    const T inv = 1.0 / zn;
                    for (auto& z : za) z = z * inv;
                }
            }
            xa[0] = -ar * w11 + arm1 * w22 - arm2 * szs * d.g1 * d.g3 + arm1 * w33;
            xa[1] = -arm2 * d.g2 * (d.g1 + szs * d.g3);
            xa[2] = arm1 * szs * w12 - arm2 * d.g1 * d.g3 - ar * szs * w33;
            const T xn = f_sqrt(f_max0(xa[0] * xa[0] + xa[1] * xa[1] + xa[2] * xa[2]));
            const T inv_xn = 1.0 / xn;
            for (auto& x : xa) x = x * inv_xn;
            ecc = xn / (ar * sqsmix * wt2);
            ya[0] = za[1] * xa[2] - za[2] * xa[1];
            ya[1] = za[2] * xa[0] - za[0] * xa[2];
            ya[2] = za[0] * xa[1] - za[1] * xa[0];
            const T conu = (xa[0] + xa[2] * szs) / sqsmix;
            const T cos_e0 = f_clamp((conu + ecc) / (1.0 + ecc * conu), -1.0, 1.0);
            const double sign = (val(ya[0]) + val(ya[2]) * szs) > 0.0 ? 1.0 : -1.0;
            const T e0 = f_acos(cos_e0) * sign;
            const T sin_e0 = f_sqrt(f_max0(1.0 - cos_e0 * cos_e0)) * sign;
            t_peri = t_ref - (e0 - ecc * sin_e0) / n;
            a = ar * sep * sqsmix;
            one_minus_e2 = f_sqrt(f_max0(1.0 - ecc * ecc));
        }
    }

    void at(double t, T* s_out, T* angle_out) const
    {
        if (mode == LCBG_ORBIT_CIRCULAR) {
            const T phi = w_orb * (t - t_ref) + phi0;
            const T c_phi = f_cos(phi);
            const T s_phi = f_sin(phi);
            const T den = f_sqrt(c_phi * c_phi + c_inc * c_inc * s_phi * s_phi);
            const T sin_a = (c_phi * s_om + c_inc * s_phi * c_om) / den;
            const T cos_a = (c_phi * c_om - c_inc * s_phi * s_om) / den;
            *s_out = s_true * den;
            *angle_out = f_atan2(sin_a, cos_a);
        } else if (mode == LCBG_ORBIT_KEPLER) {
            const T E = kepler_e(n * (t - t_peri), ecc);
            const T cos_e = f_cos(E);
            const T sin_e = f_sin(E);
            const T r0 = a * (cos_e - ecc);
            const T r1 = a * one_minus_e2 * sin_e;
            const T x0 = r0 * xa[0] + r1 * ya[0];
            const T x1 = r0 * xa[1] + r1 * ya[1];
            *s_out = f_sqrt(x0 * x0 + x1 * x1);
            *angle_out = angle + f_atan2(x1, x0);
        } else {
            *s_out = sep;
            *angle_out = angle;
        }
    }
};

// The parameters a trajectory depends on, as T (double, or Dual seeded on
// t0, tE, u0, alpha, s, piEN, piEE, g1, g2).
template <class T>
struct TrajIn {
    T t0, tE, u0, alpha, s, piEN, piEE, g1, g2;
};

// Source position in the instantaneous binary frame and s(t) of every epoch of
// one curve (VBM convention); false if an epoch's orbit is not finite.
template <class T>
bool trajectory(const TrajIn<T>& P, const lcbg_dynamics& d, int orbit_mode, const double* t, int64_t n,
                T* out)
{
    const bool parallax = d.piEN != 0.0 || d.piEE != 0.0;
    std::unique_ptr<Parallax<T>> px;
    if (parallax) px.reset(new Parallax<T>(earth_table(), d.ra, d.dec, d.tfix != 0.0 ? T(d.tfix) : P.t0));
    const OrbitIn<T> oin{P.g1, P.g2, d.g3, d.lom_szs, d.lom_ar};
    const Orbit<T> orbit(orbit_mode, P.s, P.alpha, oin, d.tfix != 0.0 ? T(d.tfix) : P.t0);
    const T cos_t = f_cos(P.alpha);
    const T sin_t = f_sin(P.alpha);
    bool ok = true;
    for (int64_t i = 0; i < n; ++i) {
        const double time = t[i];
        T tn = (time - P.t0) / P.tE;
        T beta = P.u0;
        if (parallax) {
            T dtau, dbeta;
            px->offset(time, P.piEN, P.piEE, &dtau, &dbeta);
            tn = tn + dtau;
            beta = beta + dbeta;
        }
        T x = tn * cos_t - beta * sin_t;
        T y = tn * sin_t + beta * cos_t;
        T s_t, ang;
        orbit.at(time, &s_t, &ang);
        if (!std::isfinite(val(s_t)) || !std::isfinite(val(ang))) ok = false;
        if (orbit_mode != LCBG_ORBIT_STATIC) {
            T tau, b;
            if (!parallax) {
                tau = (time - P.t0) / P.tE;
                b = P.u0;
            } else {
                tau = x * cos_t + y * sin_t;
                b = -x * sin_t + y * cos_t;
            }
            const T ca = f_cos(ang);
            const T sa = f_sin(ang);
            x = tau * ca - b * sa;
            y = b * ca + tau * sa;
        }
        out[3 * i] = x;
        out[3 * i + 1] = y;
        out[3 * i + 2] = s_t;
    }
    return ok;
}

// The per-curve constants of the device trajectory (traj.hpp), in U (double
// or TDual<9>) from the host computation in T (double or Dual<9>).
inline double to_u(double x, double*) { return x; }
inline TDual<kTrajParams> to_u(const Dual<kTrajParams>& x, TDual<kTrajParams>*)
{
    TDual<kTrajParams> r;
    r.v = x.v;
    for (int k = 0; k < kTrajParams; ++k) r.d[k] = x.d[k];
    return r;
}

// keplerian_position of the xallarap orbit (trajectory.cpp:318-333), host libm.
void xallarap_position(double time, double t0, double period, double ecc, double peri, double* x, double* y)
{
    const double mean = 2.0 * kPiHost / period * (time - t0) + peri;
    double E = mean + ecc * std::sin(mean);
    for (int i = 0; i < 10; ++i) {
        const double f = E - ecc * std::sin(E) - mean;
        const double fp = 1.0 - ecc * std::cos(E);
        const double delta = f / fp;
        E -= delta;
        if (std::abs(delta) < 1.0e-13) break;
    }
    const double cos_e = std::cos(E);
    const double sin_e = std::sin(E);
    const double r = 1.0 - ecc * cos_e;
    const double root = std::sqrt(std::max(0.0, 1.0 - ecc * ecc));
    const double cos_f = (cos_e - ecc) / r;
    const double sin_f = root * sin_e / r;
    *x = r * cos_f;
    *y = r * sin_f;
}

template <class T, class U>
void traj_consts(const TrajIn<T>& P, const lcbg_dynamics& d, int orbit_mode, TrajConst<U>* out)
{
    U* tag = nullptr;
    TrajConst<U>& K = *out;
    const bool parallax = d.piEN != 0.0 || d.piEE != 0.0;
    K.t0 = to_u(P.t0, tag);
    K.tE = to_u(P.tE, tag);
    K.u0 = to_u(P.u0, tag);
    K.cos_t = to_u(f_cos(P.alpha), tag);
    K.sin_t = to_u(f_sin(P.alpha), tag);
    K.parallax = parallax ? 1 : 0;
    K.mode = orbit_mode;
    K.sky_rotation = 0;
    K.offset = 0.0;
    for (int i = 0; i < 3; ++i) K.event[i] = K.east[i] = K.north[i] = 0.0;
    const U zero = to_u(T(0.0), tag);
    K.eph_ref = K.ref_pn = K.ref_pe = K.ref_vn = K.ref_ve = zero;
    K.piEN = to_u(P.piEN, tag);
    K.piEE = to_u(P.piEE, tag);
    if (parallax) {
        const Parallax<T> px(earth_table(), d.ra, d.dec, d.tfix != 0.0 ? T(d.tfix) : P.t0);
        K.offset = px.offset_;
        for (int i = 0; i < 3; ++i) {
            K.event[i] = px.event_[i];
            K.east[i] = px.east_[i];
            K.north[i] = px.north_[i];
        }
        K.eph_ref = to_u(px.eph_ref_, tag);
        K.ref_pn = to_u(px.ref_pn_, tag);
        K.ref_pe = to_u(px.ref_pe_, tag);
        K.ref_vn = to_u(px.ref_vn_, tag);
        K.ref_ve = to_u(px.ref_ve_, tag);
    }
    const OrbitIn<T> oin{P.g1, P.g2, d.g3, d.lom_szs, d.lom_ar};
    const Orbit<T> orbit(orbit_mode, P.s, P.alpha, oin, d.tfix != 0.0 ? T(d.tfix) : P.t0);
    K.sep = to_u(orbit.sep, tag);
    K.angle = to_u(orbit.angle, tag);
    K.t_ref = to_u(orbit.t_ref, tag);
    K.w_orb = to_u(orbit.w_orb, tag);
    K.phi0 = to_u(orbit.phi0, tag);
    K.c_inc = to_u(orbit.c_inc, tag);
    K.s_true = to_u(orbit.s_true, tag);
    K.c_om = to_u(orbit.c_om, tag);
    K.s_om = to_u(orbit.s_om, tag);
    K.n = to_u(orbit.n, tag);
    K.ecc = to_u(orbit.ecc, tag);
    K.t_peri = to_u(orbit.t_peri, tag);
    K.a = to_u(orbit.a, tag);
    K.one_minus_e2 = to_u(orbit.one_minus_e2, tag);
    for (int i = 0; i < 3; ++i) {
        K.xa[i] = to_u(orbit.xa[i], tag);
        K.ya[i] = to_u(orbit.ya[i], tag);
    }
    // terrestrial parallax: with annual parallax and an observatory position
    K.terrestrial = (parallax && (d.obs_lat != 0.0 || d.obs_lon != 0.0)) ? 1 : 0;
    K.obs_lon = K.obs_rc = K.obs_rs = 0.0;
    if (K.terrestrial) {
        const double lat = d.obs_lat * kDeg;
        K.obs_lon = d.obs_lon * kDeg;
        K.obs_rc = kTrajEarthRadiusAu * std::cos(lat);
        K.obs_rs = kTrajEarthRadiusAu * std::sin(lat);
    }
    // xallarap of the form the options name, when its amplitudes are set
    K.xallarap = kXaNone;
    K.xa_sinc = K.xa_phi = K.xa_cphi = K.xa_sphi = K.xa_omega = K.xa_xi1 = K.xa_xi2 = 0.0;
    K.xa_w = K.xa_peri = K.xa_ecc = K.xa_root = K.xa_p0x = K.xa_p0y = K.xa_vx = K.xa_vy = 0.0;
    K.xa_pin = K.xa_pie = 0.0;
    if (d.xallarap_type == kXaAngular && (d.xi_1 != 0.0 || d.xi_2 != 0.0)) {
        K.xallarap = kXaAngular;
        K.xa_sinc = std::sin(d.inc_xa);
        K.xa_phi = d.phi_xa;
        K.xa_cphi = std::cos(d.phi_xa);
        K.xa_sphi = std::sin(d.phi_xa);
        K.xa_omega = d.omega_xa;
        K.xa_xi1 = d.xi_1;
        K.xa_xi2 = d.xi_2;
    } else if (d.xallarap_type == kXaKepler && (d.piEN_xa != 0.0 || d.piEE_xa != 0.0) && d.period_xa > 0.0) {
        // the orbit at t0 and its velocity there by a forward difference of
        // period 1e-7 (apply_xallarap_orbital_elements, trajectory.cpp:335-361)
        const double t0 = val(P.t0);
        K.xallarap = kXaKepler;
        K.xa_sinc = std::sin(d.inc_xa);
        K.xa_w = 2.0 * kPiHost / d.period_xa;
        K.xa_peri = d.peri_xa;
        K.xa_ecc = d.ecc_xa;
        K.xa_root = std::sqrt(std::max(0.0, 1.0 - d.ecc_xa * d.ecc_xa));
        xallarap_position(t0, t0, d.period_xa, d.ecc_xa, d.peri_xa, &K.xa_p0x, &K.xa_p0y);
        const double dt_small = d.period_xa * 1.0e-7;
        double px = 0.0, py = 0.0;
        xallarap_position(t0 + dt_small, t0, d.period_xa, d.ecc_xa, d.peri_xa, &px, &py);
        K.xa_vx = (px - K.xa_p0x) / dt_small;
        K.xa_vy = (py - K.xa_p0y) / dt_small;
        K.xa_pin = d.piEN_xa;
        K.xa_pie = d.piEE_xa;
    }
}

}  // namespace

// Per-curve constants of the device trajectory: values, and with jac the
// derivatives with respect to t0, tE, u0, alpha, s, piEN, piEE, g1, g2.
void trajectory_constants(const lcbg_curve& cv, const lcbg_dynamics& d, int orbit_mode, TrajConst<double>* val_out,
                          TrajConst<TDual<kTrajParams>>* dual_out)
{
    if (val_out) {
        const TrajIn<double> P{cv.t0, cv.tE, cv.u0, cv.alpha, cv.s, d.piEN, d.piEE, d.g1, d.g2};
        traj_consts(P, d, orbit_mode, val_out);
    }
    if (dual_out) {
        using D = Dual<kTrajParams>;
        TrajIn<D> P;
        const double seed[kTrajParams] = {cv.t0, cv.tE, cv.u0, cv.alpha, cv.s, d.piEN, d.piEE, d.g1, d.g2};
        D* fields[kTrajParams] = {&P.t0, &P.tE, &P.u0, &P.alpha, &P.s, &P.piEN, &P.piEE, &P.g1, &P.g2};
        for (int k = 0; k < kTrajParams; ++k) {
            *fields[k] = D(seed[k]);
            fields[k]->d[k] = 1.0;
        }
        traj_consts(P, d, orbit_mode, dual_out);
    }
}

// The Earth table rows for the device (empty if none was read) and the
// number of tables read so far, which tells a device copy that it is stale.
std::vector<EarthRowD> earth_rows_device(int* generation)
{
    const std::vector<EarthRow>& rows = earth_table();
    std::lock_guard<std::mutex> lock(g_earth_mu);
    *generation = g_earth_generation;
    std::vector<EarthRowD> out(rows.size());
    for (std::size_t i = 0; i < rows.size(); ++i) {
        out[i].t = rows[i].t;
        for (int k = 0; k < 3; ++k) {
            out[i].p[k] = rows[i].p[k];
            out[i].v[k] = rows[i].v[k];
        }
    }
    return out;
}

bool dynamic_sources(const lcbg_curve& cv, const lcbg_dynamics& d, int orbit_mode, const double* t,
                     int64_t n, double* xys)
{
    const TrajIn<double> P{cv.t0, cv.tE, cv.u0, cv.alpha, cv.s, d.piEN, d.piEE, d.g1, d.g2};
    return trajectory(P, d, orbit_mode, t, n, xys);
}

// The same with the derivatives of (x, y, s) with respect to (t0, tE, u0,
// alpha, s, piEN, piEE, g1, g2): jac[27 i + 9 j + k] = d out_j / d param_k.
bool dynamic_sources_jacobian(const lcbg_curve& cv, const lcbg_dynamics& d, int orbit_mode,
                              const double* t, int64_t n, double* xys, double* jac)
{
    using D = Dual<kTrajParams>;
    TrajIn<D> P;
    const double seed[kTrajParams] = {cv.t0, cv.tE, cv.u0, cv.alpha, cv.s, d.piEN, d.piEE, d.g1, d.g2};
    D* fields[kTrajParams] = {&P.t0, &P.tE, &P.u0, &P.alpha, &P.s, &P.piEN, &P.piEE, &P.g1, &P.g2};
    for (int k = 0; k < kTrajParams; ++k) {
        *fields[k] = D(seed[k]);
        fields[k]->d[k] = 1.0;
    }
    std::vector<D> out(static_cast<std::size_t>(3 * n));
    const bool ok = trajectory(P, d, orbit_mode, t, n, out.data());
    for (int64_t i = 0; i < 3 * n; ++i) {
        xys[i] = out[static_cast<std::size_t>(i)].v;
        for (int k = 0; k < kTrajParams; ++k) jac[kTrajParams * i + k] = out[static_cast<std::size_t>(i)].d[k];
    }
    return ok;
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
    ++g_earth_generation;
    return g_earth.size() >= 2 ? LCBG_OK : LCBG_INVALID_ARGUMENT;
}

namespace {
// Terrestrial parallax or xallarap: effects the host computation leaves to
// the device trajectory.
bool device_only_effects(const lcbg_dynamics& d)
{
    const bool parallax = d.piEN != 0.0 || d.piEE != 0.0;
    return (parallax && (d.obs_lat != 0.0 || d.obs_lon != 0.0)) || d.xi_1 != 0.0 || d.xi_2 != 0.0 ||
           d.piEN_xa != 0.0 || d.piEE_xa != 0.0;
}
}  // namespace

extern "C" lcbg_status lcbg_dynamic_sources(const lcbg_curve* curve, const lcbg_dynamics* dyn,
                                            int32_t orbital_motion_mode, const double* times,
                                            int64_t n, double* xys)
{
    if (!curve || !dyn || (n > 0 && (!times || !xys)) || n < 0) return LCBG_INVALID_ARGUMENT;
    if (orbital_motion_mode < LCBG_ORBIT_STATIC || orbital_motion_mode > LCBG_ORBIT_KEPLER ||
        device_only_effects(*dyn))
        return LCBG_UNSUPPORTED;
    return lcbg::dynamic_sources(*curve, *dyn, orbital_motion_mode, times, n, xys) ? LCBG_OK
                                                                                   : LCBG_NUMERICAL_ERROR;
}

extern "C" lcbg_status lcbg_dynamic_sources_jacobian(const lcbg_curve* curve, const lcbg_dynamics* dyn,
                                                     int32_t orbital_motion_mode, const double* times,
                                                     int64_t n, double* xys, double* jac)
{
    if (!curve || !dyn || (n > 0 && (!times || !xys || !jac)) || n < 0) return LCBG_INVALID_ARGUMENT;
    if (orbital_motion_mode < LCBG_ORBIT_STATIC || orbital_motion_mode > LCBG_ORBIT_KEPLER ||
        device_only_effects(*dyn))
        return LCBG_UNSUPPORTED;
    return lcbg::dynamic_sources_jacobian(*curve, *dyn, orbital_motion_mode, times, n, xys, jac)
               ? LCBG_OK
               : LCBG_NUMERICAL_ERROR;
}
