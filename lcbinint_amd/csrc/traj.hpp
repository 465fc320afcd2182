// Source trajectories with annual parallax and binary orbital motion, the
// per-epoch part, for the device (and the host, which computes the per-curve
// constants with its own libm):
//
//   TrajConst   per-curve constants: the event's sky basis and the parallax
//               offsets at the reference time (EarthOrbitalParallaxProjector,
//               src/lcbinint/model/trajectory.cpp:221-285), the circular or
//               Keplerian orbit constants (orbital_state,
//               src/lcbinint/model/orbital_motion.cpp:70-199), cos/sin(alpha);
//   traj_epoch  one epoch: the parallax offset (five light-time iterations on
//               the linearly interpolated Earth table, :185-209, 259-266), the
//               orbit's separation and angle, and the rotation of the source
//               into the instantaneous binary frame (VBM convention,
//               src/lcbinint/model/lens_model.cpp:84-104, 164-178).
//
// T is double (values) or TDual<9> (values with the derivatives with respect to
// t0, tE, u0, alpha, s, piEN, piEE, g1, g2).  Every operation keeps the order
// of the host restatement in dynamics.cpp, so that without orbital motion
// (no trigonometry per epoch) the device's positions are bitwise the host's;
// the per-epoch cos, sin and atan2 of an orbit come from the device libm.
#pragma once

#include <hip/hip_runtime.h>

namespace lcbg {

constexpr int kTrajParams = 9;  // t0, tE, u0, alpha, s, piEN, piEE, g1, g2
constexpr double kTrajAuLightDays = 0.005775518331436995;

// value and N derivatives; the value part is the plain double operation
template <int N>
struct TDual {
    double v;
    double d[N];
};

template <int N>
__host__ __device__ inline TDual<N> tconst(double x)
{
    TDual<N> r;
    r.v = x;
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}

#define LCBG_TD template <int N> __host__ __device__ inline TDual<N>
LCBG_TD operator+(const TDual<N>& a, const TDual<N>& b)
{
    TDual<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
LCBG_TD operator-(const TDual<N>& a, const TDual<N>& b)
{
    TDual<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
LCBG_TD operator-(const TDual<N>& a)
{
    TDual<N> r;
    r.v = -a.v;
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
LCBG_TD operator*(const TDual<N>& a, const TDual<N>& b)
{
    TDual<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
LCBG_TD operator/(const TDual<N>& a, const TDual<N>& b)
{
    TDual<N> r;
    r.v = a.v / b.v;
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
LCBG_TD operator+(const TDual<N>& a, double b) { return a + tconst<N>(b); }
LCBG_TD operator+(double a, const TDual<N>& b) { return tconst<N>(a) + b; }
LCBG_TD operator-(const TDual<N>& a, double b) { return a - tconst<N>(b); }
LCBG_TD operator-(double a, const TDual<N>& b) { return tconst<N>(a) - b; }
LCBG_TD operator*(const TDual<N>& a, double b)
{
    TDual<N> r;
    r.v = a.v * b;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
LCBG_TD operator*(double a, const TDual<N>& b) { return b * a; }
LCBG_TD operator/(const TDual<N>& a, double b)
{
    TDual<N> r;
    r.v = a.v / b;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
    return r;
}
#undef LCBG_TD

__host__ __device__ inline double tval(double x) { return x; }
template <int N>
__host__ __device__ inline double tval(const TDual<N>& x) { return x.v; }
template <class T>
__host__ __device__ inline T tmake(double x);
template <>
__host__ __device__ inline double tmake<double>(double x) { return x; }
template <>
__host__ __device__ inline TDual<kTrajParams> tmake<TDual<kTrajParams>>(double x) { return tconst<kTrajParams>(x); }

__device__ inline double t_sqrt(double x) { return __builtin_sqrt(x); }
__device__ inline double t_cos(double x) { return cos(x); }
__device__ inline double t_sin(double x) { return sin(x); }
__device__ inline double t_atan2(double y, double x) { return atan2(y, x); }
template <int N>
__device__ inline TDual<N> t_sqrt(const TDual<N>& a)
{
    TDual<N> r;
    r.v = __builtin_sqrt(a.v);
    for (int k = 0; k < N; ++k) r.d[k] = r.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
    return r;
}
template <int N>
__device__ inline TDual<N> t_cos(const TDual<N>& a)
{
    TDual<N> r;
    double s, c;
    sincos(a.v, &s, &c);
    r.v = c;
    for (int k = 0; k < N; ++k) r.d[k] = -s * a.d[k];
    return r;
}
template <int N>
__device__ inline TDual<N> t_sin(const TDual<N>& a)
{
    TDual<N> r;
    double s, c;
    sincos(a.v, &s, &c);
    r.v = s;
    for (int k = 0; k < N; ++k) r.d[k] = c * a.d[k];
    return r;
}
template <int N>
__device__ inline TDual<N> t_atan2(const TDual<N>& y, const TDual<N>& x)
{
    TDual<N> r;
    r.v = atan2(y.v, x.v);
    const double den = x.v * x.v + y.v * y.v;
    for (int k = 0; k < N; ++k) r.d[k] = den > 0.0 ? (x.v * y.d[k] - y.v * x.d[k]) / den : 0.0;
    return r;
}

// One row of the Earth table: JD, position (AU), velocity (AU/day).
struct EarthRowD {
    double t, p[3], v[3];
};

enum : int { kTrajStatic = 0, kTrajCircular = 1, kTrajKepler = 2 };

template <class T>
struct TrajConst {
    T t0, tE, u0, cos_t, sin_t;
    int parallax, mode;
    int sky_rotation;  // with orbital motion, rotate the sky-frame source (x, y) even without
                       // parallax (the geometry helpers' source_for_geometry)
    // parallax
    double offset;                        // JD origin added to reduced dates
    double event[3], east[3], north[3];   // sky basis of the event
    T eph_ref, ref_pn, ref_pe, ref_vn, ref_ve, piEN, piEE;
    // orbit
    T sep, angle, t_ref;
    T w_orb, phi0, c_inc, s_true, c_om, s_om;            // circular
    T n, ecc, t_peri, a, one_minus_e2, xa[3], ya[3];     // Keplerian
    // terrestrial parallax (apply_terrestrial_parallax, trajectory.cpp:363-401):
    // the observatory's longitude (rad) and R cos(lat), R sin(lat) (AU)
    int terrestrial;
    double obs_lon, obs_rc, obs_rs;
    // xallarap (trajectory.cpp:287-361): kXaAngular or kXaKepler, or 0
    int xallarap;
    double xa_sinc;                                       // sin(inc_xa)
    double xa_phi, xa_cphi, xa_sphi, xa_omega, xa_xi1, xa_xi2;   // angular velocity
    double xa_w, xa_peri, xa_ecc, xa_root, xa_p0x, xa_p0y, xa_vx, xa_vy, xa_pin, xa_pie;  // orbital elements
};

enum : int { kXaNone = 0, kXaAngular = 1, kXaKepler = 2 };
constexpr double kTrajEarthRadiusAu = 4.2635212e-5;
constexpr double kTrajSiderealDegPerDay = 360.98564736629;
constexpr double kTrajDegree = 3.14159265358979323846 / 180.0;

// interpolate_earth_state (trajectory.cpp:185-209): rows sorted by time;
// NaN outside the table.
template <class T>
__device__ inline void traj_earth_at(const EarthRowD* rows, int n_rows, const T& t, T* p, T* v)
{
    const double tv = tval(t);
    if (n_rows < 2 || tv < rows[0].t || tv > rows[n_rows - 1].t) {
        const double nan = __builtin_nan("");
        for (int i = 0; i < 3; ++i) p[i] = v[i] = tmake<T>(nan);
        return;
    }
    // the first row with time > tv (upper_bound)
    int lo = 0, hi = n_rows;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (tv < rows[mid].t) hi = mid;
        else lo = mid + 1;
    }
    int h = lo;
    if (h == 0) h = 1;
    if (h == n_rows) h = n_rows - 1;
    const EarthRowD& a = rows[h - 1];
    const EarthRowD& b = rows[h];
    const double dt = b.t - a.t;
    const T w = dt != 0.0 ? (t - a.t) / dt : tmake<T>(0.0);
    for (int i = 0; i < 3; ++i) {
        p[i] = a.p[i] + (b.p[i] - a.p[i]) * w;
        v[i] = a.v[i] + (b.v[i] - a.v[i]) * w;
    }
}

template <class T>
__device__ inline T traj_dot3(const T* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <class T>
__device__ inline T traj_kepler_e(const T& m, const T& e)
{
    T E = m + e * t_sin(m);
    for (int i = 0; i < 10; ++i) {
        const T f = E - e * t_sin(E) - m;
        const T fp = 1.0 - e * t_cos(E);
        const T d = f / fp;
        E = E - d;
        if (fabs(tval(d)) <= 1.0e-13) break;
    }
    return E;
}

// solve_kepler of the xallarap orbit (trajectory.cpp:303-316): E from M.
template <class T>
__device__ inline T traj_kepler_xa(const T& m, double e)
{
    T E = m + e * t_sin(m);
    for (int i = 0; i < 10; ++i) {
        const T f = E - e * t_sin(E) - m;
        const T fp = 1.0 - e * t_cos(E);
        const T d = f / fp;
        E = E - d;
        if (fabs(tval(d)) < 1.0e-13) break;
    }
    return E;
}

// The terrestrial parallax and xallarap terms added to (tn, beta) after the
// annual parallax, in the reference's order (Trajectory::source_position,
// trajectory.cpp:436-448); the per-epoch cos and sin come from the device libm.
template <class T>
__device__ inline void traj_extra(const TrajConst<T>& K, double time, T& tn, T& beta)
{
    if (K.terrestrial) {
        const double jd = time + (time < 2450000.0 ? 2450000.0 : 0.0);
        const double gmst = (280.46061837 + kTrajSiderealDegPerDay * (jd - 2451545.0)) * kTrajDegree;
        const double ha = gmst + K.obs_lon;
        const double r[3] = {K.obs_rc * cos(ha), K.obs_rc * sin(ha), K.obs_rs};
        const double pn = -(r[0] * K.north[0] + r[1] * K.north[1] + r[2] * K.north[2]);
        const double pe = -(r[0] * K.east[0] + r[1] * K.east[1] + r[2] * K.east[2]);
        tn = tn + (K.piEN * pn + K.piEE * pe);
        beta = beta + (-K.piEE * pn + K.piEN * pe);
    }
    if (K.xallarap == kXaAngular) {
        const T phit = K.xa_omega * (time - K.t0);
        const T arg = K.xa_phi + phit;
        const T disp0 = K.xa_sinc * ((-K.xa_cphi + t_cos(arg)) + phit * K.xa_sphi);
        const T disp1 = (-phit * K.xa_cphi - K.xa_sphi) + t_sin(arg);
        tn = tn + (K.xa_xi1 * disp0 + K.xa_xi2 * disp1);
        beta = beta + (K.xa_xi2 * disp0 - K.xa_xi1 * disp1);
    } else if (K.xallarap == kXaKepler) {
        const T elapsed = time - K.t0;
        const T E = traj_kepler_xa(K.xa_w * elapsed + K.xa_peri, K.xa_ecc);
        const T ce = t_cos(E);
        const T se = t_sin(E);
        const T r = 1.0 - K.xa_ecc * ce;
        const T cf = (ce - K.xa_ecc) / r;
        const T sf = K.xa_root * se / r;
        const T dev_x = (r * cf - K.xa_p0x) - K.xa_vx * elapsed;
        const T dev_y = (r * sf - K.xa_p0y) - K.xa_vy * elapsed;
        const T disp0 = K.xa_sinc * dev_x;
        tn = tn + (K.xa_pin * disp0 + K.xa_pie * dev_y);
        beta = beta + (K.xa_pie * disp0 - K.xa_pin * dev_y);
    }
}

// One epoch: the source position (x, y) in the instantaneous binary frame and
// the separation s.  Returns false if the orbit is not finite
// (lens_model.cpp:117-120).
template <class T>
__device__ inline bool traj_epoch(const TrajConst<T>& K, const EarthRowD* rows, int n_rows, double time, T* x_out,
                                  T* y_out, T* s_out)
{
    T tn = (time - K.t0) / K.tE;
    T beta = K.u0;
    if (K.parallax) {
        const double eph = time + K.offset;
        const T obs = tmake<T>(eph);
        T emit = obs;
        T p[3], v[3];
        for (int i = 0; i < 5; ++i) {
            traj_earth_at(rows, n_rows, emit, p, v);
            emit = obs - traj_dot3(p, K.event) * kTrajAuLightDays;
        }
        traj_earth_at(rows, n_rows, emit, p, v);
        const T pn = -traj_dot3(p, K.north);
        const T pe = -traj_dot3(p, K.east);
        const T dt = eph - K.eph_ref;
        const T dn = pn - K.ref_pn - K.ref_vn * dt;
        const T de = pe - K.ref_pe - K.ref_ve * dt;
        const T dtau = K.piEN * dn + K.piEE * de;
        const T dbeta = -K.piEE * dn + K.piEN * de;
        tn = tn + dtau;
        beta = beta + dbeta;
    }
    traj_extra(K, time, tn, beta);
    T x = tn * K.cos_t - beta * K.sin_t;
    T y = tn * K.sin_t + beta * K.cos_t;
    T s_t, ang;
    if (K.mode == kTrajCircular) {
        const T phi = K.w_orb * (time - K.t_ref) + K.phi0;
        const T c_phi = t_cos(phi);
        const T s_phi = t_sin(phi);
        const T den = t_sqrt(c_phi * c_phi + K.c_inc * K.c_inc * s_phi * s_phi);
        const T sin_a = (c_phi * K.s_om + K.c_inc * s_phi * K.c_om) / den;
        const T cos_a = (c_phi * K.c_om - K.c_inc * s_phi * K.s_om) / den;
        s_t = K.s_true * den;
        ang = t_atan2(sin_a, cos_a);
    } else if (K.mode == kTrajKepler) {
        const T E = traj_kepler_e(K.n * (time - K.t_peri), K.ecc);
        const T cos_e = t_cos(E);
        const T sin_e = t_sin(E);
        const T r0 = K.a * (cos_e - K.ecc);
        const T r1 = K.a * K.one_minus_e2 * sin_e;
        const T x0 = r0 * K.xa[0] + r1 * K.ya[0];
        const T x1 = r0 * K.xa[1] + r1 * K.ya[1];
        s_t = t_sqrt(x0 * x0 + x1 * x1);
        ang = K.angle + t_atan2(x1, x0);
    } else {
        s_t = K.sep;
        ang = K.angle;
    }
    const bool ok = isfinite(tval(s_t)) && isfinite(tval(ang));
    if (K.mode != kTrajStatic) {
        T tau, b;
        if (!K.parallax && !K.sky_rotation) {
            tau = (time - K.t0) / K.tE;
            b = K.u0;
        } else {
            tau = x * K.cos_t + y * K.sin_t;
            b = -x * K.sin_t + y * K.cos_t;
        }
        const T ca = t_cos(ang);
        const T sa = t_sin(ang);
        x = tau * ca - b * sa;
        y = b * ca + tau * sa;
    }
    *x_out = x;
    *y_out = y;
    *s_out = s_t;
    return ok;
}

}  // namespace lcbg
