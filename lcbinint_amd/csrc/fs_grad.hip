// Value-and-gradient kernels of the finite-source light curve.
//
// Per epoch the derivative of the magnification is taken for the method the
// dispatcher chose in the value run, with forward-mode tangents (dual.hpp) in
// six slots: source x, y, s, q_eff, rho and the linear limb-darkening c.  The
// reference has no gradients; these follow the mathematics of each method:
//
//  * point source: each physical image z of the source y solves the lens
//    equation w(z; s, q) = y; its tangent is J^-1 (dy - dw/dp dp) (implicit
//    differentiation), and A = sum 1 / |det J(z)| is differentiated through z,
//    s and q;
//  * hexadecapole: the 13 point-source samples of hexadecapole_binary
//    (finite_source_magnifier.cpp:792-847) at y + rho e_k, each with its own
//    implicit-differentiation tangent, combined by the same formula;
//  * inverse ray: an antialiased quadrature on a lattice anchored at the
//    origin of the lens frame, cell (i, j) at z = h (i + 1/2, j + 1/2) with
//    h the power of two at or below rho / nbin, fixed while the parameters
//    move (a lattice that scaled with rho would sweep every edge cell across
//    the edge |z| / rho times faster than the source moves).  With
//    g = rho^2 - |w(z) - y|^2 the source brightness is
//    (1 - c) [g > 0] + c sqrt(max(g, 0)) / rho; each cell weighs it averaged
//    over its footprint, g linearised over the cell and smeared by a Gaussian
//    of kSmear cells (the antialiased edge), and A = sum W h^2 / (pi rho^2
//    (1 - c/3)).  The cells with nonzero weight are enumerated by a scan-line
//    flood fill from the value run's image seeds.  This is a differentiable
//    stand-in for the plain quadrature, which it matches to its
//    discretisation error; the value run's magnifications are unchanged.
//
// k_grad_chain turns the six tangents into derivatives with respect to the
// curve parameters (t0, tE, u0, alpha, s, q, rho, u_LD) and reduces chi^2 and
// its gradient per curve.
#include "dual.hpp"
#include "fs_launch.hpp"
#include "fs_math.hpp"
#include "fs_wave.hpp"

namespace lcbg {

namespace {

constexpr double kSqrtHalfG = 0.70710678118654752440;
// Width of the antialiasing Gaussian in lattice cells.  A wider kernel
// smooths more but moves the quadrature further from the plain one (median
// |mag_aa / mag - 1| 1e-3 at one cell, 2e-3 at 1.5 cells on 256 x 1000
// curves), so the width stays one cell; the roughness that was left at one
// cell came from the critical curve, where the width sigma |grad g| vanishes,
// and is removed by the floor of width_floor.
constexpr double kSmear = 1.0;

struct DLens {
    Dual a, m1, m2;
};

// make_lens with tangents
__device__ __forceinline__ DLens make_dlens(const Dual& s, const Dual& q)
{
    const bool small = q.v < 1.0;
    const Dual qq = small ? q : 1.0 / q;
    DLens L;
    L.a = small ? -s : s;
    L.m1 = 1.0 / (1.0 + qq);
    L.m2 = qq * L.m1;
    return L;
}

template <class T>
struct MapOut {
    T wx, wy, kr, ki;  // w(z), and kappa = dw/d(conj z) = m1/(zbar-a)^2 + m2/zbar^2
};

// Real lens map w = z - m1 (z - a)/|z - a|^2 - m2 z/|z|^2 - a m1 and kappa;
// the Jacobian of (x, y) -> w is [[1 + kr, ki], [ki, 1 - kr]].  Divisions are
// written as products with reciprocals so that the value-only and the tangent
// evaluations round identically.
template <class T, class LT>
__device__ __forceinline__ MapOut<T> lens_map_k(const LT& L, const T& x, const T& y)
{
    const T xa = x - L.a;
    const T y2 = y * y;
    const T i1 = 1.0 / (xa * xa + y2);
    const T i2 = 1.0 / (x * x + y2);
    MapOut<T> m;
    m.wx = x - L.m1 * xa * i1 - L.m2 * x * i2 - L.a * L.m1;
    m.wy = y - L.m1 * y * i1 - L.m2 * y * i2;
    const T i1s = i1 * i1;
    const T i2s = i2 * i2;
    m.kr = L.m1 * (xa * xa - y2) * i1s + L.m2 * (x * x - y2) * i2s;
    m.ki = 2.0 * (L.m1 * xa * y * i1s + L.m2 * x * y * i2s);
    return m;
}

// Floor of the smoothing variance, in units of sigma^4: 128 phi(|kappa|^2)
// with phi(k) = 2 k / (1 + k^2), which is 1 on the critical curve
// (|kappa| = 1), where grad g vanishes at the points of the source edge that
// touch the caustic and the linear width sigma |grad g| would leave the weight
// sharp, and falls to 0 away from it, both towards the lenses (where kappa
// diverges) and far from them.  The second-order variance of g over the
// footprint on the critical curve is 16 sigma^4 (sigma^4 |H|_F^2 / 2 with
// g_zzbar = -(1 + |kappa|^2) = -2); the floor is eight times that.  Measured
// on 256 x 1000 curves against central differences of the same quadrature
// (99th percentile of the relative error of dA/dt0 over inverse-ray epochs):
// 9e-4 without the floor, 8e-5 at 16 sigma^4, 7e-6 at 64 sigma^4; the floor
// changes the value by far less than the width does.
template <class T>
__device__ __forceinline__ T width_floor(const MapOut<T>& m)
{
    const T k = m.kr * m.kr + m.ki * m.ki;
    return 256.0 * k / (1.0 + k * k);
}

// Magnification 1/|det J| of the image at z of source y, with tangents: the
// image's tangent is J^-1 (dy - dw/dp dp) (implicit differentiation of
// w(z; s, q) = y).
__device__ Dual image_dual(const DLens& L, double zxv, double zyv, const Dual& yx, const Dual& yy)
{
    const Dual zx0 = dconst(zxv);
    const Dual zy0 = dconst(zyv);
    const MapOut<Dual> m0 = lens_map_k(L, zx0, zy0);  // tangents: dw/dp at fixed z
    const double j11 = 1.0 + m0.kr.v, j12 = m0.ki.v, j22 = 1.0 - m0.kr.v;
    const double det = j11 * j22 - j12 * j12;
    const double idet = 1.0 / det;
    Dual zx = zx0, zy = zy0;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) {
        const double rx = yx.d[k] - m0.wx.d[k];
        const double ry = yy.d[k] - m0.wy.d[k];
        zx.d[k] = (j22 * rx - j12 * ry) * idet;
        zy.d[k] = (-j12 * rx + j11 * ry) * idet;
    }
    const MapOut<Dual> m = lens_map_k(L, zx, zy);
    const Dual detd = 1.0 - (m.kr * m.kr + m.ki * m.ki);
    return 1.0 / dabs(detd);
}

// Point-source magnification with tangents.
__device__ Dual ps_dual(const Dual& yx, const Dual& yy, const Dual& s, const Dual& q)
{
    if (s.v == 0.0 || !(q.v > 0.0)) return dconst(0.0);
    const Geom g = make_geom(s.v, q.v, yx.v, yy.v);
    cx roots[5];
    solve_quintic(g, roots);
    const unsigned mask = select_physical(g, roots);
    const DLens L = make_dlens(s, q);
    Dual A = dconst(0.0);
    for (int i = 0; i < 5; ++i) {
        if (!(mask & (1u << i))) continue;
        A += image_dual(L, roots[i].re, roots[i].im, yx, yy);
    }
    return A;
}

// hexadecapole_binary (finite_source_magnifier.cpp:792-847) with tangents; the
// linear limb-darkening Gamma = 10 c / (15 - 5 c), Lambda = 0 (:349-367).
__device__ Dual hex_dual(const Dual& yx, const Dual& yy, const Dual& s, const Dual& q,
                         const Dual& rho, const Dual& c)
{
    const double sineta[8] = {0.0, 1.0, 0.0, -1.0, kSqrtHalfG, kSqrtHalfG, -kSqrtHalfG, -kSqrtHalfG};
    const double coseta[8] = {1.0, 0.0, -1.0, 0.0, kSqrtHalfG, -kSqrtHalfG, -kSqrtHalfG, kSqrtHalfG};
    const Dual a0 = ps_dual(yx, yy, s, q);
    Dual a1p = dconst(0.0), a2p = dconst(0.0), a1c = dconst(0.0);
    for (int i = 0; i < 4; ++i) {
        a1p += ps_dual(yx + rho * coseta[i], yy + rho * sineta[i], s, q);
        a2p += ps_dual(yx + (0.5 * rho) * coseta[i], yy + (0.5 * rho) * sineta[i], s, q);
        a1c += ps_dual(yx + rho * coseta[i + 4], yy + rho * sineta[i + 4], s, q);
    }
    a1p = a1p * 0.25 - a0;
    a2p = a2p * 0.25 - a0;
    a1c = a1c * 0.25 - a0;
    const Dual a2rho2 = (16.0 * a2p - a1p) * (1.0 / 3.0);
    const Dual a4rho4 = (a1p + a1c) * 0.5 - a2rho2;
    const Dual den = 15.0 - 5.0 * c;
    const Dual gamma = den.v == 0.0 ? dconst(0.0) : (10.0 * c) / den;
    const Dual quad = 0.5 * a2rho2 * (1.0 - 0.2 * gamma);
    const Dual hex = a4rho4 * (1.0 / 3.0) * (1.0 - (11.0 / 35.0) * gamma);
    return a0 + quad + hex;
}

// ------------------------------------------------ antialiased inverse ray

struct GradCtx {
    Lens L;        // values
    DLens Ld;      // with tangents
    Dual yx, yy, rho, c;
    double q_eff;  // lens-frame mass ratio (value)
    double h;      // cell side: the power of two at or below rho / nbin
    int lane;
};

// 64-point Gauss-Legendre nodes and weights on [-1, 1] (numpy.polynomial.legendre.leggauss(64))
__constant__ double kGLx[64] = {
    -0.9993050417357722, -0.9963401167719552, -0.9910133714767443, -0.983336253884626,
    -0.973326827789911, -0.9610087996520538, -0.9464113748584028, -0.9295691721319396,
    -0.9105221370785028, -0.8893154459951141, -0.8659993981540928, -0.8406292962525803,
    -0.8132653151227975, -0.7839723589433414, -0.7528199072605319, -0.7198818501716108,
    -0.6852363130542333, -0.6489654712546573, -0.6111553551723933, -0.571895646202634,
    -0.5312794640198946, -0.48940314570705296, -0.4463660172534641, -0.4022701579639916,
    -0.3572201583376681, -0.31132287199021097, -0.2646871622087674, -0.21742364374000708,
    -0.1696444204239928, -0.12146281929612056, -0.07299312178779904, -0.02435029266342443,
    0.02435029266342443, 0.07299312178779904, 0.12146281929612056, 0.1696444204239928,
    0.21742364374000708, 0.2646871622087674, 0.31132287199021097, 0.3572201583376681,
    0.4022701579639916, 0.4463660172534641, 0.48940314570705296, 0.5312794640198946,
    0.571895646202634, 0.6111553551723933, 0.6489654712546573, 0.6852363130542333,
    0.7198818501716108, 0.7528199072605319, 0.7839723589433414, 0.8132653151227975,
    0.8406292962525803, 0.8659993981540928, 0.8893154459951141, 0.9105221370785028,
    0.9295691721319396, 0.9464113748584028, 0.9610087996520538, 0.973326827789911,
    0.983336253884626, 0.9910133714767443, 0.9963401167719552, 0.9993050417357722,
};
__constant__ double kGLw[64] = {
    0.0017832807216942152, 0.004147033260562923, 0.006504457968979654, 0.008846759826364391,
    0.011168139460131466, 0.013463047896718231, 0.015726030476025082, 0.0179517157756973,
    0.020134823153530094, 0.022270173808383007, 0.024352702568710853, 0.026377469715054627,
    0.028339672614259702, 0.030234657072402495, 0.03205792835485145, 0.03380516183714179,
    0.03547221325688232, 0.03705512854024015, 0.03855015317861559, 0.03995374113272035,
    0.041262563242623486, 0.0424735151236536, 0.043583724529323464, 0.044590558163756545,
    0.045491627927418114, 0.046284796581314375, 0.04696818281621, 0.0475401657148303,
    0.04799938859645832, 0.048344762234802954, 0.048575467441503456, 0.04869095700913975,
    0.04869095700913975, 0.048575467441503456, 0.048344762234802954, 0.04799938859645832,
    0.0475401657148303, 0.04696818281621, 0.046284796581314375, 0.045491627927418114,
    0.044590558163756545, 0.043583724529323464, 0.0424735151236536, 0.041262563242623486,
    0.03995374113272035, 0.03855015317861559, 0.03705512854024015, 0.03547221325688232,
    0.03380516183714179, 0.03205792835485145, 0.030234657072402495, 0.028339672614259702,
    0.026377469715054627, 0.024352702568710853, 0.022270173808383007, 0.020134823153530094,
    0.0179517157756973, 0.015726030476025082, 0.013463047896718231, 0.011168139460131466,
    0.008846759826364391, 0.006504457968979654, 0.004147033260562923, 0.0017832807216942152,
};
// E[sqrt(1 + Z/t)] = sum_k binom(1/2, 2k) E[Z^2k] t^(-2k), k = 0..7
constexpr double kSqrtSeries[8] = {1.0, -0.125, -0.11718750000000001, -0.30761718750000006, -1.3746643066406252, -8.763484954833984, -72.8464686870575, -747.9771338403226};

// F(t) = E[sqrt(max(t + Z, 0))] and F'(t) for a standard normal Z:
// F = int_0^inf 2 v^2 phi(v^2 - t) dv, F' = int_0^inf phi(v^2 - t) dv
// (u = t + Z = v^2), by 64-point Gauss-Legendre on [0, sqrt(max(t, 0) + 9)]
// for t < 12 (relative error below 1e-15), and by the asymptotic series
// sqrt(t) E[sqrt(1 + Z/t)] beyond (below 6e-14).
__device__ __forceinline__ void sqrt_gauss(double t, double* F, double* dF)
{
    const double kInvSqrt2Pi = 0.39894228040143267794;
    if (t >= 12.0) {
        const double it2 = 1.0 / (t * t);
        double p = 1.0, f = 0.0, df = 0.0;
        for (int k = 0; k < 8; ++k) {
            f += kSqrtSeries[k] * p;
            df += kSqrtSeries[k] * (0.5 - 2.0 * k) * p;
            p *= it2;
        }
        const double st = __builtin_sqrt(t);
        *F = st * f;
        *dF = df / st;
        return;
    }
    const double V = __builtin_sqrt((t > 0.0 ? t : 0.0) + 9.0);
    double f = 0.0, df = 0.0;
    for (int k = 0; k < 64; ++k) {
        const double v = 0.5 * V * (kGLx[k] + 1.0);
        const double a = v * v - t;
        const double e = kGLw[k] * exp(-0.5 * a * a);
        f += e * v * v;
        df += e;
    }
    *F = 0.5 * V * kInvSqrt2Pi * 2.0 * f;
    *dF = 0.5 * V * kInvSqrt2Pi * df;
}

// Cells the fill enumerates: those within eight Gaussian widths outside the
// edge, g + 8 (sigma |grad g|_1 + sigma^2 sqrt(floor)) > 0 with
// sigma = kSmear h (the 1-norm bounds the 2-norm from above, and the sum bounds
// the width of aa_weight).
//
// The test only bounds the set of cells with nonzero weight (its margin is
// eight widths), so it is evaluated with the hardware reciprocal and reciprocal
// square root refined once (relative error ~1e-16 instead of IEEE rounding): a
// cell whose test value lies within that error of zero has a weight below
// erfc(8) either way.  The cells' weights (aa_weight) are exact.
__device__ __forceinline__ double rcp_refined(double d)
{
    const double r = __builtin_amdgcn_rcp(d);
    const double e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ bool aa_in(const GradCtx& G, int i, int j)
{
    const double rho = G.rho.v;
    const double x = (static_cast<double>(i) + 0.5) * G.h;
    const double y = (static_cast<double>(j) + 0.5) * G.h;
    const Lens& L = G.L;
    const double xa = x - L.a;
    const double y2 = y * y;
    const double i1 = rcp_refined(xa * xa + y2);
    const double i2 = rcp_refined(x * x + y2);
    const double wx = x - L.m1 * xa * i1 - L.m2 * x * i2 - L.a * L.m1;
    const double wy = y - L.m1 * y * i1 - L.m2 * y * i2;
    const double i1s = i1 * i1;
    const double i2s = i2 * i2;
    const double kr = L.m1 * (xa * xa - y2) * i1s + L.m2 * (x * x - y2) * i2s;
    const double ki = 2.0 * (L.m1 * xa * y * i1s + L.m2 * x * y * i2s);
    const double dx = wx - G.yx.v;
    const double dy = wy - G.yy.v;
    const double g = rho * rho - (dx * dx + dy * dy);
    const double gx = -2.0 * (dx * (1.0 + kr) + dy * ki);
    const double gy = -2.0 * (dx * ki + dy * (1.0 - kr));
    const double sig = kSmear * G.h;
    const double k = kr * kr + ki * ki;
    const double fl = 256.0 * k * rcp_refined(1.0 + k * k);
    const double root = fl > 0.0 ? fl * __builtin_amdgcn_rsq(fl) : 0.0;
    const double curv = sig * sig * root;
    return g + 8.0 * (sig * (dabs(gx) + dabs(gy)) + curv) > 0.0;
}

// Weight of cell (i, j) with tangents: the source brightness
// I = (1 - c) [g > 0] + c sqrt(max(g, 0)) / rho  (g = rho^2 - |w - y|^2,
// linear limb darkening) averaged over a Gaussian footprint of kSmear cells
// (sigma = kSmear h per axis): g -> g + s Z, s^2 = sigma^2 |grad g|^2 +
// sigma^4 width_floor.  The floor keeps the width positive where grad g
// vanishes on the critical curve, so that the weight stays smooth there.
// With t = g / s,
// W = (1 - c) Phi(t) + (c / rho) sqrt(s) F(t).  Unlike a box footprint, the
// Gaussian makes the lattice sum smooth at the cell scale (its ripple scales
// with the kernel's transform at the lattice frequency, ~exp(-2 pi^2)).
__device__ Dual aa_weight(const GradCtx& G, int i, int j)
{
    const Dual x = dconst((static_cast<double>(i) + 0.5) * G.h);
    const Dual y = dconst((static_cast<double>(j) + 0.5) * G.h);
    const MapOut<Dual> m = lens_map_k(G.Ld, x, y);
    const Dual dx = m.wx - G.yx;
    const Dual dy = m.wy - G.yy;
    const Dual g = G.rho * G.rho - (dx * dx + dy * dy);
    const Dual gx = -2.0 * (dx * (1.0 + m.kr) + dy * m.ki);
    const Dual gy = -2.0 * (dx * m.ki + dy * (1.0 - m.kr));
    const double sig2 = (kSmear * G.h) * (kSmear * G.h);
    const Dual sc = dsqrt(sig2 * (gx * gx + gy * gy) + (sig2 * sig2) * width_floor(m));
    if (!(sc.v > 0.0)) return dconst(g.v > 0.0 ? 1.0 - G.c.v + G.c.v * __builtin_sqrt(g.v) / G.rho.v : 0.0);
    const Dual t = g / sc;
    if (t.v < -9.0) return dconst(0.0);
    Dual phi = t;
    phi.v = 0.5 * erfc(-t.v * 0.70710678118654752440);
    const double dens = 0.39894228040143267794 * exp(-0.5 * t.v * t.v);
#pragma unroll
    for (int k = 0; k < kTangents; ++k) phi.d[k] = dens * t.d[k];
    double F, dF;
    sqrt_gauss(t.v, &F, &dF);
    Dual Fd = t;
    Fd.v = F;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) Fd.d[k] = dF * t.d[k];
    return (1.0 - G.c) * phi + G.c * (1.0 / G.rho) * dsqrt(sc) * Fd;
}

struct AaFill {
    int* head;   // [kPhiHash] hashed run-chain heads, keyed on the row
    int* runs;   // [run*4]: row, left, right, next
    int n_runs;
    int run_cap;
    bool overflow;
    long long steps;
    long long cells;  // cells of the recorded runs
};

__device__ __forceinline__ bool aa_visited(const AaFill& F, int i, int j)
{
    for (int r = F.head[j & (kPhiHash - 1)]; r >= 0; r = F.runs[4 * r + 3]) {
        if (F.runs[4 * r] == j && i >= F.runs[4 * r + 1] && i <= F.runs[4 * r + 2]) return true;
    }
    return false;
}

// The runs already recorded in row j, one per lane (lane k < n holds run k),
// gathered by one walk of the row's hash chain; `many` when the row has more
// runs than lanes (the checks then walk the chain per cell).
struct RowRuns {
    int l, r;
    int n;
    bool many;
};

__device__ __forceinline__ int lane_int(int v, int k) { return __builtin_amdgcn_readlane(v, k); }

__device__ RowRuns row_runs(const AaFill& F, int j)
{
    RowRuns V{0, -1, 0, false};
    const int lane = lane_id();
    for (int p = uni(F.head[j & (kPhiHash - 1)]); p >= 0;) {
        const int row = uni(F.runs[4 * p + 0]);
        const int l = uni(F.runs[4 * p + 1]);
        const int r = uni(F.runs[4 * p + 2]);
        const int next = uni(F.runs[4 * p + 3]);
        if (row == j) {
            if (V.n < kWave) {
                if (lane == V.n) {
                    V.l = l;
                    V.r = r;
                }
                V.n = V.n + 1;
            } else {
                V.many = true;
            }
        }
        p = next;
    }
    return V;
}

__device__ __forceinline__ void row_runs_add(RowRuns& V, int l, int r)
{
    if (V.n < kWave) {
        if (lane_id() == V.n) {
            V.l = l;
            V.r = r;
        }
        V.n = V.n + 1;
    } else {
        V.many = true;
    }
}

__device__ __forceinline__ bool row_visited(const AaFill& F, const RowRuns& V, int i, int j)
{
    if (V.many) return aa_visited(F, i, j);
    bool v = false;
    for (int k = 0; k < V.n; ++k) v = v || (i >= lane_int(V.l, k) && i <= lane_int(V.r, k));
    return v;
}

// The first cell at or after i (uniform) that no recorded run of the row holds.
__device__ __forceinline__ int skip_runs(const RowRuns& V, int i)
{
    if (V.many) return i;
    for (bool moved = true; moved;) {
        moved = false;
        for (int k = 0; k < V.n; ++k) {
            const int l = lane_int(V.l, k), r = lane_int(V.r, k);
            if (i >= l && i <= r) {
                i = r + 1;
                moved = true;
            }
        }
    }
    return i;
}

__device__ __forceinline__ bool aa_step(AaFill& F)
{
    F.steps = uni(F.steps + 1);
    if (F.steps > kStepBudget) F.overflow = true;
    return !F.overflow;
}

// Cells k = 0, 1, ... of the row from `from` in direction `dir` that are in
// the fill and unvisited, 64 per wave step: the count of such cells before the
// first that is not.
__device__ int extend_cells(const GradCtx& G, AaFill& F, const RowRuns& V, int from, int dir, int j)
{
    int n = 0;
    for (;;) {
        if (!aa_step(F)) return n;
        const int c = from + dir * (n + G.lane);
        const bool ok = !row_visited(F, V, c, j) && aa_in(G, c, j);
        const int stop = first_set(~ballot(ok));
        if (stop < 0) {
            n += kWave;
            continue;
        }
        return n + stop;
    }
}

// Records run [left, right] of row j.
__device__ void record_run(AaFill& F, RowRuns& V, int left, int right, int j)
{
    if (F.n_runs >= F.run_cap) {
        F.overflow = true;
        return;
    }
    const int older = uni(F.head[j & (kPhiHash - 1)]);
    wave_sync();
    {
        const int r = F.n_runs;
        wave_store(&F.runs[4 * r + 0], j);
        wave_store(&F.runs[4 * r + 1], left);
        wave_store(&F.runs[4 * r + 2], right);
        wave_store(&F.runs[4 * r + 3], older);
        wave_store(&F.head[j & (kPhiHash - 1)], r);
    }
    F.n_runs = uni(F.n_runs + 1);
    F.cells = uni(F.cells + static_cast<long long>(right - left + 1));
    row_runs_add(V, left, right);
    wave_sync();
}

// Extends cell (i, j) to its run along the row and records it.
__device__ void aa_add_run(const GradCtx& G, AaFill& F, int i, int j)
{
    RowRuns V = row_runs(F, j);
    const int left = i - extend_cells(G, F, V, i - 1, -1, j);
    const int right = i + extend_cells(G, F, V, i + 1, 1, j);
    if (F.overflow) return;
    record_run(F, V, left, right, j);
}

// The runs of row j that touch cells [lo, hi] (8-connected neighbours of a
// run of the next or previous row).  A row with one recorded run overlapping
// [lo, hi] (most often the row the growing run was found from) first has the
// cells of [lo, hi] outside that run tested in one gathered step, which
// usually finds none in the fill and ends the row.  Otherwise the range is evaluated in
// chunks of 64 consecutive cells, stretches held by recorded runs skipped
// without evaluation, the first chunk reaching kLookLeft cells left of lo; a
// run found in a chunk takes its cells from that chunk's evaluation and is
// extended past it only where it reaches the chunk's edge.  A run that starts
// right of lo has an evaluated neighbour on its left outside the fill (a
// visited neighbour would have been taken into its run when that run was
// extended, the fill being a fixed set), so it needs no left extension.
constexpr int kLookLeft = 8;

__device__ void grow_row(const GradCtx& G, AaFill& F, int lo, int hi, int j)
{
    RowRuns V = row_runs(F, j);
    // the row's recorded runs that overlap [lo, hi]: one of them (m == 1) at k1
    int m = 0, k1 = 0;
    for (int k = 0; k < V.n && !V.many; ++k) {
        if (lane_int(V.l, k) <= hi && lane_int(V.r, k) >= lo) {
            ++m;
            k1 = k;
        }
    }
    if (!V.many && m == 1) {
        const int l1 = lane_int(V.l, k1), r1 = lane_int(V.r, k1);
        const int lend = l1 - 1 < hi ? l1 - 1 : hi;
        const int nl = lend >= lo ? lend - lo + 1 : 0;    // [lo, min(l1 - 1, hi)]
        const int rs = r1 + 1 > lo ? r1 + 1 : lo;
        const int nr = hi >= rs ? hi - rs + 1 : 0;         // [max(r1 + 1, lo), hi]
        if (nl + nr <= kWave) {
            if (!aa_step(F)) return;
            bool ok = false;
            if (G.lane < nl + nr) ok = aa_in(G, G.lane < nl ? lo + G.lane : rs + (G.lane - nl), j);
            if (!ballot(ok)) return;
        }
    }
    int c = lo;
    bool first = true;
    while (!F.overflow) {
        c = skip_runs(V, c);
        if (c > hi || !aa_step(F)) break;
        const int b = (first && c == lo) ? lo - kLookLeft : c;  // cell of lane 0
        first = false;
        const int cc = b + G.lane;
        const bool ok = !row_visited(F, V, cc, j) && aa_in(G, cc, j);
        const uint64_t okm = ballot(ok);
        const int from = c - b, to = hi - b;  // lanes that may start a run
        uint64_t pending = okm & (to >= kWave - 1 ? ~0ull : ((2ull << to) - 1ull)) & ~((1ull << from) - 1ull);
        int next = b + kWave;
        while (pending && !F.overflow) {
            const int sb = first_set(pending);
            const int s = b + sb;
            int left = s;
            if (s == lo) {
                const uint64_t below = ~okm & ((1ull << sb) - 1ull);  // cells left of s outside
                left = below ? b + (63 - __builtin_clzll(below)) + 1 : b - extend_cells(G, F, V, b - 1, -1, j);
            }
            const uint64_t rest = ~(okm >> sb);  // zero bits: the run's cells in this chunk
            const int len = rest ? first_set(rest) : kWave - sb;
            int right = s + len - 1;
            if (sb + len == kWave) {
                right = right + extend_cells(G, F, V, b + kWave, 1, j);
                next = right + 1;
            }
            if (F.overflow) break;
            record_run(F, V, left, right, j);
            const int done = sb + len;  // bits [0, done] are settled (done: not in the fill)
            pending = done >= kWave - 1 ? 0ull : pending & ~((2ull << done) - 1ull);
        }
        c = next;
    }
}

// The cells of one epoch's antialiased quadrature, flood-filled from the
// seeds: the runs of cells with nonzero weight are left in F.runs (F.n_runs of
// them, F.cells cells) and their weights are summed later by k_grad_cells.
// Images smaller than the lattice can resolve (linear size rho sqrt(A_i)
// under four cells) enter as point images with their exact tangents, the
// returned sum, and the seeds near them start no fill.  F.overflow is set past
// a capacity or the budget.
__device__ Dual aa_enumerate(const GradCtx& G, AaFill& F, const Pt* seeds, int n_seeds)
{
    F.n_runs = 0;
    F.overflow = false;
    F.steps = 0;
    F.cells = 0;
    const double h = G.h;
    Pt unresolved[5];
    int n_unres = 0;
    Dual a_points = dconst(0.0);
    {
        const Geom g = make_geom(G.Ld.a.v < 0.0 ? -G.Ld.a.v : G.Ld.a.v, G.q_eff, G.yx.v, G.yy.v);
        cx roots[5];
        solve_quintic(g, roots);
        const unsigned mask = select_physical(g, roots);
        for (int i = 0; i < 5; ++i) {
            if (!(mask & (1u << i))) continue;
            const double ai = 1.0 / fabs(jacobian_det(g, roots[i]));
            if (G.rho.v * __builtin_sqrt(ai) < 4.0 * h) {
                unresolved[n_unres++] = Pt{roots[i].re, roots[i].im};
                a_points += image_dual(G.Ld, roots[i].re, roots[i].im, G.yx, G.yy);
            }
        }
    }
    for (int sidx = 0; sidx < n_seeds && !F.overflow; ++sidx) {
        const Pt sd = seeds[sidx];
        bool near_point = false;
        for (int u = 0; u < n_unres; ++u)
            near_point = near_point || dist2(sd, unresolved[u]) < 64.0 * h * h;
        if (near_point) continue;
        const int i0 = static_cast<int>(floor(sd.x / h));
        const int j0 = static_cast<int>(floor(sd.y / h));
        // start: the first cell with nonzero weight in a 9 x 9 block around
        // the seed's cell, nearest first (|di| + |dj|, then row, then column)
        int found = -1;
        for (int base = 0; base < 81 && found < 0; base += kWave) {
            const int o = base + G.lane;
            bool in = false;
            if (o < 81) {
                const int di = (o % 9) - 4;
                const int dj = (o / 9) - 4;
                in = aa_in(G, i0 + di, j0 + dj);
            }
            // order candidates by Manhattan distance, then index
            const int key = in ? ((abs((o % 9) - 4) + abs((o / 9) - 4)) * 128 + o) : 0x7fffffff;
            int best = key;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const int ob = __shfl_xor(best, off, kWave);
                best = ob < best ? ob : best;
            }
            best = uni(best);
            if (best != 0x7fffffff) found = best & 127;
        }
        if (found < 0) continue;
        const int si = i0 + (found % 9) - 4;
        const int sj = j0 + (found / 9) - 4;
        if (aa_visited(F, si, sj)) continue;
        aa_add_run(G, F, si, sj);
    }
    // grow: each recorded run visits its two neighbour rows, diagonal
    // neighbours included (8-connected), so that a thin diagonal image stays
    // one component while its fringe cells' weights go to zero
    for (int q = 0; q < F.n_runs && !F.overflow; ++q) {
        const int row = uni(F.runs[4 * q + 0]);
        const int L = uni(F.runs[4 * q + 1]) - 1;
        const int R = uni(F.runs[4 * q + 2]) + 1;
        grow_row(G, F, L, R, row - 1);
        grow_row(G, F, L, R, row + 1);
    }
    wave_sync();
    for (int q = G.lane; q < F.n_runs; q += kWave) F.head[F.runs[4 * q] & (kPhiHash - 1)] = -1;
    wave_sync();
    return a_points;
}


__device__ __forceinline__ void epoch_inputs(const CurveRec& cv, Pt src, Dual& yx, Dual& yy, Dual& s,
                                             Dual& q, Dual& rho, Dual& c)
{
    yx = dvar(src.x, kTx);
    yy = dvar(src.y, kTy);
    s = dvar(cv.s, kTs);
    q = dvar(cv.q_eff, kTq);
    rho = dvar(cv.rho, kTrho);
    c = dvar(cv.ld_c, kTc);
}

__device__ __forceinline__ void store_grad(const FsEpochs& ep, double* mag_aa, int64_t e, const Dual& A)
{
    if (mag_aa) mag_aa[e] = A.v;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) ep.dint[e * kTangents + k] = A.d[k];
}

}  // namespace

// Point-source and hexadecapole epochs: one thread each.
__global__ void __launch_bounds__(256) k_grad_point(const CurveRec* __restrict__ curves, FsEpochs ep,
                                                    double* mag_aa)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= ep.n) return;
    const int method = ep.method[e];
    if (method != LCBG_METHOD_POINT_SOURCE && method != LCBG_METHOD_HEXADECAPOLE) return;
    const CurveRec cv = curves[ep.curve_of[e]];
    Dual yx, yy, s, q, rho, c;
    epoch_inputs(cv, ep.src[e], yx, yy, s, q, rho, c);
    const Dual A = method == LCBG_METHOD_HEXADECAPOLE ? hex_dual(yx, yy, s, q, rho, c)
                                                      : ps_dual(yx, yy, s, q);
    store_grad(ep, mag_aa, e, A);
}


// The lattice of an epoch's antialiased quadrature and its inputs with tangents.
__device__ __forceinline__ GradCtx grad_ctx(const CurveRec& cv, Pt src, const FsOptions& opt)
{
    GradCtx G;
    G.L = make_lens(cv.s, cv.q_eff);
    Dual s, q;
    epoch_inputs(cv, src, G.yx, G.yy, s, q, G.rho, G.c);
    G.Ld = make_dlens(s, q);
    G.q_eff = cv.q_eff;
    const int nb = cv.source_bins > 0 ? cv.source_bins : opt.source_bins;
    // a cell side that does not move with rho: the power of two at or
    // below rho / nbin (between one and two cells per rho / nbin)
    G.h = exp2(floor(log2(cv.rho / static_cast<double>(nb > 1 ? nb : 1))));
    G.lane = lane_id();
    return G;
}

// Exclusive prefix sum over the wave of a per-lane int (Hillis-Steele).
__device__ __forceinline__ int wave_exclusive_scan(int v, int* total)
{
    int x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane_id() >= off) x += y;
    }
    *total = __shfl(x, kWave - 1, kWave);
    return x - v;
}

// The weight sums of chunk c (cells 64c .. 64c + 63 of the pool, all of one
// item) into W.partial: each lane finds its cell's run by binary search over
// the item's runs and evaluates aa_weight.  Chunks of an item that failed after
// its chunks were counted are skipped.
__device__ void grad_chunk(const CurveRec* __restrict__ curves, const FsOptions& opt, const FsEpochs& ep,
                           const IrLists& Lst, const GradWork& W, int c)
{
    const int item = uni(W.chunk_item[c]);
    const int run_off = uni(W.run_off[item]);
    if (run_off < 0) return;
    const int run_n = uni(W.run_n[item]);
    const int cells = uni(W.cells[item]);
    const int t = (c - uni(W.chunk_off[item])) * kWave + lane_id();
    const int64_t e = Lst.ir_sorted[item];
    const CurveRec cv = curves[ep.curve_of[e]];
    GradCtx G = grad_ctx(cv, ep.src[e], opt);
    Dual w = dconst(0.0);
    if (t < cells) {
        // the last run whose first cell index is <= t
        int lo = 0, hi = run_n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (W.run_cells[run_off + mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int* run = W.runs + 4 * static_cast<size_t>(run_off + lo);
        const int i = run[1] + (t - W.run_cells[run_off + lo]);
        w = aa_weight(G, i, run[0]);
    }
    double sums[kGradSlots];
    sums[0] = wave_sum(w.v);
#pragma unroll
    for (int k = 0; k < kTangents; ++k) sums[1 + k] = wave_sum(w.d[k]);
#pragma unroll
    for (int k = 0; k < kGradSlots; ++k) wave_store(&W.partial[static_cast<size_t>(c) * kGradSlots + k], sums[k]);
}

__device__ __forceinline__ unsigned agent_load(const unsigned* p)
{
    return uni(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ int agent_load(const int* p)
{
    return uni(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Bounded waits of the task consumers, in wall-clock ticks (100 MHz): for a
// reserved task to be published (its producer writes it right after the
// reservation) and for new tasks while items are still being enumerated.
constexpr long long kTaskPublishWait = 10000000;  // 0.1 s
constexpr long long kTaskIdleWait = 200000000;    // 2 s

// The enumeration publishes an item's chunks as tasks of kGradTaskChunks
// chunks once its runs are in W (the item's data, a release fence, then the
// ready flags), and counts the item as done afterwards.
__device__ void publish_tasks(const FsEpochs& ep, const GradWork& W, int chunk_base, int n_chunks)
{
    if (n_chunks > 0) {
        const int n_tasks = (n_chunks + kGradTaskChunks - 1) / kGradTaskChunks;
        const int t0 = wave_fetch_add(&ep.counters[kQGradTasks], static_cast<unsigned>(n_tasks));
        const long long cap = static_cast<long long>(W.task_cap);
        const int lim = static_cast<int>(t0 >= cap ? 0 : (t0 + n_tasks <= cap ? n_tasks : cap - t0));
        for (int k = lane_id(); k < lim; k += kWave) {
            const int first = k * kGradTaskChunks;
            W.tasks[2 * static_cast<size_t>(t0 + k)] = chunk_base + first;
            W.tasks[2 * static_cast<size_t>(t0 + k) + 1] =
                n_chunks - first < kGradTaskChunks ? n_chunks - first : kGradTaskChunks;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        for (int k = lane_id(); k < lim; k += kWave)
            __hip_atomic_store(&W.task_ready[t0 + k], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const unsigned inc = lane_id() == 0 ? 1u : 0u;
    __hip_atomic_fetch_add(&ep.counters[kQGradItemsDone], inc, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    wave_sync();
}

// Once the item queue is empty, a wave of the enumeration evaluates published
// tasks, flagging each evaluated chunk in W.chunk_done, until every item has
// been enumerated.  A wave that waits past a bound leaves too; the chunks no
// wave has evaluated are k_grad_cells' (which evaluates every unflagged chunk
// the same way), so the results do not depend on the timing.
__device__ void take_tasks(const CurveRec* __restrict__ curves, const FsOptions& opt, const FsEpochs& ep,
                           const IrLists& Lst, const GradWork& W, int count)
{
    unsigned* cursor = &ep.counters[kQGradTaskCursor];
    const int chunk_cap = uni(static_cast<int>(W.chunk_cap));
    int t = wave_fetch_inc(cursor);
    long long idle_from = static_cast<long long>(wall_clock64());
    for (;;) {
        const unsigned reserved = agent_load(&ep.counters[kQGradTasks]);
        if (static_cast<unsigned>(t) < reserved) {
            if (static_cast<unsigned>(t) >= W.task_cap) return;
            const long long w0 = static_cast<long long>(wall_clock64());
            bool ready = agent_load(&W.task_ready[t]) != 0;
            while (!ready && static_cast<long long>(wall_clock64()) - w0 < kTaskPublishWait) {
                __builtin_amdgcn_s_sleep(2);
                ready = agent_load(&W.task_ready[t]) != 0;
            }
            if (!ready) return;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const int c0 = uni(W.tasks[2 * static_cast<size_t>(t)]);
            const int nc = uni(W.tasks[2 * static_cast<size_t>(t) + 1]);
            for (int c = c0; c < c0 + nc && c >= 0 && c < chunk_cap; ++c) {
                if (uni(W.chunk_item[c]) >= count) break;  // not an item of this run
                grad_chunk(curves, opt, ep, Lst, W, c);
                wave_store(&W.chunk_done[c], uint8_t{1});
            }
            wave_sync();
            if (static_cast<int>(agent_load(&ep.counters[kQGradItemsDone])) >= count) return;
            t = wave_fetch_inc(cursor);
            idle_from = static_cast<long long>(wall_clock64());
            continue;
        }
        // once every item is enumerated, the tasks left go to k_grad_cells,
        // which evaluates them at a higher occupancy
        const unsigned done = uni(__hip_atomic_load(&ep.counters[kQGradItemsDone], __ATOMIC_ACQUIRE,
                                                    __HIP_MEMORY_SCOPE_AGENT));
        if (static_cast<int>(done) >= count) return;
        if (static_cast<long long>(wall_clock64()) - idle_from > kTaskIdleWait) return;
        __builtin_amdgcn_s_sleep(8);
    }
}

// Cell enumeration of every inverse-ray item, one wave per item, persistent
// waves over the items in cost order; the runs go to the pools of W.  In the
// first pass a wave that finds the item queue empty takes chunk tasks while
// the other waves still enumerate (the enumeration's tail is a few large
// items, which leave most of the device idle).
__global__ void __launch_bounds__(kWave) k_grad_ir(const CurveRec* __restrict__ curves,
                                                    FsOptions opt, FsEpochs ep, IrLists Lst,
                                                    PolarScratch scr, int pass, GradWork W)
{
    const bool retry = pass == kPassRetry;
    const bool tasks = !retry && W.task_cap > 0;
    AaFill F;
    F.head = scr.phi_head + static_cast<size_t>(blockIdx.x) * kPhiHash;
    F.runs = scr.runs + static_cast<size_t>(blockIdx.x) * scr.run_cap * 4;
    F.run_cap = scr.run_cap;
    const int count = uni(static_cast<int>(ep.counters[retry ? kQGradRetryCount : kQIrCount]));
    unsigned* cursor = &ep.counters[retry ? kQGradRetryCursor : kQGradCursor];
    for (;;) {
        const int qi = next_item(cursor);
        if (qi >= count) break;
        const int item = retry ? uni(Lst.grad_retry[qi]) : qi;
        const int64_t e = Lst.ir_sorted[item];
        const long long t_begin = static_cast<long long>(wall_clock64());
        const int n_seeds = uni(Lst.seed_n[item]);
        const CurveRec cv = curves[ep.curve_of[e]];
        const GradCtx G = grad_ctx(cv, ep.src[e], opt);
        bool failed = n_seeds < 0;
        Dual pts = dconst(0.0);
        F.overflow = false;
        F.n_runs = 0;
        if (!failed) {
            pts = aa_enumerate(G, F, Lst.seed_pool + uni(Lst.seed_off[item]), n_seeds);
            failed = F.overflow;
        }
        const bool cap = F.overflow && F.n_runs >= F.run_cap;
        if (ep.cost) {
            // enumeration ticks, runs, cells and wave steps (added: a capacity
            // retry adds its own)
            const int64_t n = ep.n;
            const long long ticks = static_cast<long long>(wall_clock64()) - t_begin;
            wave_store(&ep.cost[11 * n + e], ep.cost[11 * n + e] + static_cast<int64_t>(ticks));
            wave_store(&ep.cost[12 * n + e], ep.cost[12 * n + e] + static_cast<int64_t>(F.n_runs));
            wave_store(&ep.cost[13 * n + e], ep.cost[13 * n + e] + static_cast<int64_t>(F.cells));
            wave_store(&ep.cost[14 * n + e], ep.cost[14 * n + e] + static_cast<int64_t>(F.steps));
        }
        if (cap && !retry) {
            const int pos = wave_fetch_inc(&ep.counters[kQGradRetryCount]);
            wave_store(&Lst.grad_retry[pos], item);
            wave_sync();
            if (tasks) publish_tasks(ep, W, 0, 0);
            continue;
        }
        if (F.overflow) atomicOr(&ep.counters[kQFlags], cap ? kErrPolarCap : kErrBudget);
        int run_base = -1, chunk_base = 0;
        const int n = uni(F.n_runs);
        const long long cells = uni(F.cells);
        const int n_chunks = static_cast<int>((cells + kWave - 1) / kWave);
        if (!failed) {
            run_base = wave_fetch_add(&ep.counters[kQGradRuns], static_cast<unsigned>(n));
            chunk_base = wave_fetch_add(&ep.counters[kQGradChunks], static_cast<unsigned>(n_chunks));
            // every allocated chunk inside the pool names its item, also when the
            // item fails here (k_grad_cells then skips it)
            const unsigned end = static_cast<unsigned>(chunk_base) + static_cast<unsigned>(n_chunks);
            const int lim = static_cast<int>((end < W.chunk_cap ? end : W.chunk_cap) - static_cast<unsigned>(chunk_base));
            for (int k = G.lane; k < lim; k += kWave) W.chunk_item[chunk_base + k] = item;
            if (static_cast<unsigned>(run_base) + static_cast<unsigned>(n) > W.run_cap || end > W.chunk_cap ||
                cells > 0x7fffffffll) {
                atomicOr(&ep.counters[kQFlags], kErrGradPool);
                failed = true;
                run_base = -1;
            }
        }
        if (!failed) {
            // the runs, each with the count of the item's cells before it
            int before = 0;
            for (int q0 = 0; q0 < n; q0 += kWave) {
                const int q = q0 + G.lane;
                int len = 0;
                if (q < n) {
                    const int row = F.runs[4 * q + 0];
                    const int l = F.runs[4 * q + 1];
                    const int r = F.runs[4 * q + 2];
                    len = r - l + 1;
                    int* dst = W.runs + 4 * static_cast<size_t>(run_base + q);
                    dst[0] = row;
                    dst[1] = l;
                    dst[2] = r;
                    dst[3] = item;
                }
                int total = 0;
                const int off = wave_exclusive_scan(len, &total);
                if (q < n) W.run_cells[run_base + q] = before + off;
                before = uni(before + total);
            }
        }
        wave_sync();
        wave_store(&W.run_off[item], run_base);
        wave_store(&W.run_n[item], n);
        wave_store(&W.chunk_off[item], chunk_base);
        wave_store(&W.cells[item], static_cast<int>(failed ? 0 : cells));
        wave_store(&W.h[item], G.h);
        wave_store(&W.points[static_cast<size_t>(item) * kGradSlots], pts.v);
#pragma unroll
        for (int k = 0; k < kTangents; ++k)
            wave_store(&W.points[static_cast<size_t>(item) * kGradSlots + 1 + k], pts.d[k]);
        wave_sync();
        if (tasks) publish_tasks(ep, W, chunk_base, failed ? 0 : n_chunks);
    }
    if (tasks) take_tasks(curves, opt, ep, Lst, W, count);
}

// No chunk task published and no chunk evaluated: every flag of the task and
// chunk pools cleared at the start of a gradient run (so that a run that ended
// early leaves nothing behind for the next).
__global__ void __launch_bounds__(256) k_grad_clear(GradWork W)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t t0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t t = t0; t < W.task_cap; t += stride) W.task_ready[t] = 0;
    // the chunk flags 16 at a time (the pool's size is a multiple of 16)
    uint4* flags = reinterpret_cast<uint4*>(W.chunk_done);
    for (size_t t = t0; t < W.chunk_cap / 16; t += stride) flags[t] = make_uint4(0u, 0u, 0u, 0u);
}

// Weights of every chunk of 64 enumerated cells not evaluated by a task of the
// enumeration, one wave per chunk, the waves taking 64 consecutive chunks at a
// time.
__global__ void __launch_bounds__(kWave) k_grad_cells(const CurveRec* __restrict__ curves, FsOptions opt,
                                                       FsEpochs ep, IrLists Lst, GradWork W)
{
    const unsigned total = ep.counters[kQGradChunks];
    const int n_chunks = uni(static_cast<int>(total < W.chunk_cap ? total : W.chunk_cap));
    const int stride = static_cast<int>(gridDim.x) * kWave;
    for (int base = static_cast<int>(blockIdx.x) * kWave; base < n_chunks; base += stride) {
        const int c = base + lane_id();
        const bool todo = c < n_chunks && W.chunk_done[c] == 0;
        uint64_t m = ballot(todo);
        while (m) {
            const int k = first_set(m);
            m &= m - 1;
            grad_chunk(curves, opt, ep, Lst, W, base + k);
        }
    }
}

// The antialiased magnification of each inverse-ray item: its chunks' sums,
// lane-strided in chunk order and reduced over the wave (a fixed order, so
// repeated runs agree bitwise), sum W h^2 / (pi rho^2 (1 - c/3)) plus the
// unresolved point images.
__global__ void __launch_bounds__(kWave) k_grad_finish(const CurveRec* __restrict__ curves, FsOptions opt,
                                                        FsEpochs ep, IrLists Lst, GradWork W, double* mag_aa)
{
    const int count = uni(static_cast<int>(ep.counters[kQIrCount]));
    for (int item = static_cast<int>(blockIdx.x); item < count; item += static_cast<int>(gridDim.x)) {
        const int64_t e = Lst.ir_sorted[item];
        const int run_off = uni(W.run_off[item]);
        if (run_off < 0) {
            wave_sync();
            store_grad(ep, mag_aa, e, dconst(__builtin_nan("")));
            wave_sync();
            continue;
        }
        const int c0 = uni(W.chunk_off[item]);
        const int nc = uni((W.cells[item] + kWave - 1) / kWave);
        double acc[kGradSlots];
#pragma unroll
        for (int k = 0; k < kGradSlots; ++k) acc[k] = 0.0;
        for (int c = lane_id(); c < nc; c += kWave) {
#pragma unroll
            for (int k = 0; k < kGradSlots; ++k) acc[k] += W.partial[static_cast<size_t>(c0 + c) * kGradSlots + k];
        }
        Dual total;
        total.v = wave_sum(acc[0]);
#pragma unroll
        for (int k = 0; k < kTangents; ++k) total.d[k] = wave_sum(acc[1 + k]);
        const CurveRec cv = curves[ep.curve_of[e]];
        const GradCtx G = grad_ctx(cv, ep.src[e], opt);
        const double h = uni(W.h[item]);
        Dual pts;
        pts.v = uni(W.points[static_cast<size_t>(item) * kGradSlots]);
#pragma unroll
        for (int k = 0; k < kTangents; ++k) pts.d[k] = uni(W.points[static_cast<size_t>(item) * kGradSlots + 1 + k]);
        const Dual flux = (1.0 - G.c * (1.0 / 3.0)) * (kPi * G.rho * G.rho);
        const Dual A = total * (h * h) / flux + pts;
        wave_sync();
        store_grad(ep, mag_aa, e, A);  // uniform values, stored from every lane
        wave_sync();
    }
}

// d magnification / d (t0, tE, u0, alpha, s, q, rho, u_LD) per epoch from the six
// tangents, and per curve chi^2 = sum ((F - fs A - fb) / sigma)^2 with its
// gradient.  One workgroup per curve.
__global__ void __launch_bounds__(256) k_grad_chain(const CurveRec* __restrict__ curves, FsEpochs ep,
                                                    const double* __restrict__ flux,
                                                    const double* __restrict__ sigma, double fs,
                                                    double fb, double* dmag, double* chi2,
                                                    double* grad)
{
    const int c = blockIdx.x;
    const CurveRec cv = curves[c];
    const int64_t lo = ep.offsets[c];
    const int64_t hi = ep.offsets[c + 1];
    double acc[kGradParams + 1];
#pragma unroll
    for (int k = 0; k <= kGradParams; ++k) acc[k] = 0.0;
    for (int64_t e = lo + threadIdx.x; e < hi; e += blockDim.x) {
        const double* t = ep.dint + e * kTangents;
        const double tau = (ep.times[e] - cv.t0) * cv.inverse_tE;
        const double u0 = ep.src_y ? ep.src_y[e] : cv.u0;
        const double yx = tau * cv.cos_alpha - u0 * cv.sin_alpha;
        const double yy = tau * cv.sin_alpha + u0 * cv.cos_alpha;
        double g[kGradParams];
        const double along = t[kTx] * cv.cos_alpha + t[kTy] * cv.sin_alpha;
        g[0] = -along * cv.inverse_tE;                              // t0
        g[1] = -along * tau * cv.inverse_tE;                        // tE
        g[2] = -t[kTx] * cv.sin_alpha + t[kTy] * cv.cos_alpha;      // u0
        g[3] = -t[kTx] * yy + t[kTy] * yx;                          // alpha
        g[4] = t[kTs];                                              // s
        g[5] = -t[kTq] * cv.q_eff * cv.q_eff;                       // q (q_eff = 1/q)
        g[6] = t[kTrho];                                            // rho
        g[7] = t[kTc];                                              // u_LD
        if (dmag) {
#pragma unroll
            for (int k = 0; k < kGradParams; ++k) dmag[e * kGradParams + k] = g[k];
        }
        if (chi2) {
            const double r = (flux[e] - (fs * ep.mag[e] + fb)) / sigma[e];
            acc[kGradParams] += r * r;
            const double w = -2.0 * r / sigma[e] * fs;
#pragma unroll
            for (int k = 0; k < kGradParams; ++k) acc[k] += w * g[k];
        }
    }
    if (!chi2) return;
    __shared__ double red[kGradParams + 1][256];
#pragma unroll
    for (int k = 0; k <= kGradParams; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) {
#pragma unroll
            for (int k = 0; k <= kGradParams; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        chi2[c] = red[kGradParams][0];
#pragma unroll
        for (int k = 0; k < kGradParams; ++k) grad[static_cast<int64_t>(c) * kGradParams + k] = red[k][0];
    }
}

lcbg_status launch_grad_enumerate(const CurveRec* d_curves, const FsOptions& opt, const FsEpochs& ep,
                                  const IrLists& lists, const PolarScratch& slots,
                                  const PolarScratch& large_slots, const GradWork& work, hipStream_t st)
{
    hipLaunchKernelGGL(k_grad_clear, dim3(1024), dim3(256), 0, st, work);
    // one wave per slot: waves that only take tasks would slow the enumeration
    hipLaunchKernelGGL(k_grad_ir, dim3(static_cast<unsigned>(slots.n_slots)), dim3(kWave), 0, st, d_curves,
                       opt, ep, lists, slots, int{kPassSmall}, work);
    hipLaunchKernelGGL(k_grad_ir, dim3(static_cast<unsigned>(large_slots.n_slots)), dim3(kWave), 0, st,
                       d_curves, opt, ep, lists, large_slots, int{kPassRetry}, work);
    return check_launch();
}

lcbg_status launch_grad_rest(const CurveRec* d_curves, int n_curves, const FsOptions& opt,
                             const FsEpochs& ep, const IrLists& lists, const GradWork& work, int n_cus,
                             double* mag_aa, double* dmag, const double* flux, const double* sigma,
                             double fs, double fb, double* chi2, double* grad, hipStream_t st)
{
    const unsigned blocks = static_cast<unsigned>((ep.n + 255) / 256);
    hipLaunchKernelGGL(k_grad_point, dim3(blocks), dim3(256), 0, st, d_curves, ep, mag_aa);
    // chunks across the device (16 waves per CU), then the per-item sums
    hipLaunchKernelGGL(k_grad_cells, dim3(static_cast<unsigned>(16 * (n_cus > 0 ? n_cus : 256))), dim3(kWave),
                       0, st, d_curves, opt, ep, lists, work);
    hipLaunchKernelGGL(k_grad_finish, dim3(static_cast<unsigned>(4 * (n_cus > 0 ? n_cus : 256))), dim3(kWave),
                       0, st, d_curves, opt, ep, lists, work, mag_aa);
    hipLaunchKernelGGL(k_grad_chain, dim3(static_cast<unsigned>(n_curves)), dim3(256), 0, st, d_curves,
                       ep, flux, sigma, fs, fb, dmag, chi2, grad);
    return check_launch();
}

}  // namespace lcbg
