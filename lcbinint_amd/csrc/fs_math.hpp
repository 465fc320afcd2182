// FP64 building blocks of the finite-source path (gfx950).
//
// Each function follows one reference routine expression by expression (the
// library is built with -ffp-contract=off) so that lattice coordinates, seed
// positions and caustic points are bitwise the reference's:
//   - std::complex arithmetic of the critical-curve polynomial and the complex
//     lens equation: the product is the naive formula GCC emits inline, the
//     quotient is libgcc's __divdc3 (Smith's method);
//   - the three rounding forms of the real lens map (finite_source_magnifier.cpp
//     :65-80 and :278-291 divide, :121-153 multiplies by reciprocals, :196-209
//     divides by squared denominators);
//   - limb-darkening nodes (finite_source_magnifier.cpp:369-375, 3979-4012);
//   - glibc's hypot algorithm.
#pragma once

#include "binary_lens.hpp"
#include "fs_types.hpp"

namespace lcbg {

constexpr double kPi = 3.14159265358979323846;

// std::clamp / std::max / std::min argument order and NaN behaviour.
__device__ __forceinline__ double dclamp(double v, double lo, double hi)
{
    return (v < lo) ? lo : (hi < v) ? hi : v;
}
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

using Pt = Pt2;

__device__ __forceinline__ double dist2(Pt a, Pt b)
{
    const double dx = a.x - b.x;
    const double dy = a.y - b.y;
    return dx * dx + dy * dy;
}

// libgcc __divdc3 (complex division, GCC 12): Smith's algorithm with power-of-two
// prescaling; the prescaling is exact, so it is kept only where it changes the
// result (subnormal ratio).
__device__ __forceinline__ cx std_cdiv(double a, double b, double c, double d)
{
    const double RMIN = 0x1p-1022;
    double ratio, denom, x, y;
    if (fabs(c) < fabs(d)) {
        ratio = c / d;
        denom = (c * ratio) + d;
        if (fabs(ratio) > RMIN) {
            x = ((a * ratio) + b) / denom;
            y = ((b * ratio) - a) / denom;
        } else {
            x = ((c * (a / d)) + b) / denom;
            y = ((c * (b / d)) - a) / denom;
        }
    } else {
        ratio = d / c;
        denom = (d * ratio) + c;
        if (fabs(ratio) > RMIN) {
            x = ((b * ratio) + a) / denom;
            y = (b - (a * ratio)) / denom;
        } else {
            x = (a + (d * (b / c))) / denom;
            y = (b - (d * (a / c))) / denom;
        }
    }
    return cx{x, y};
}

// Lens placement of make_binary_lens_mapper (finite_source_magnifier.cpp:43-52),
// identical to make_fast_vbm_geometry: heavier lens at z = a, lighter at 0.
struct Lens {
    double a, m1, m2;
};

__device__ __forceinline__ Lens make_lens(double separation, double mass_ratio)
{
    const double s = fabs(separation);
    const double q_input = fabs(mass_ratio);
    const double q = q_input < 1.0 ? q_input : 1.0 / q_input;
    Lens L;
    L.a = q_input < 1.0 ? -s : s;
    L.m1 = 1.0 / (1.0 + q);
    L.m2 = q * L.m1;
    return L;
}

// critical_curve_polynomial_coefficients (finite_source_magnifier.cpp:528-545)
// with phase = (cos phi, sin phi); std::complex products carried literally,
// zero imaginary parts included.
__device__ __forceinline__ void critical_quartic(const Lens& L, cx phase, cx* c)
{
    const cx a{L.a, 0.0};
    const cx m1{L.m1, 0.0};
    const cx m2{L.m2, 0.0};
    const cx a2 = a * a;
    c[0] = m2 * a2;
    c[1] = (-2.0 * m2) * a;
    c[2] = (m1 + m2) - phase * a2;
    c[3] = (2.0 * phase) * a;
    c[4] = cx{-phase.re, -phase.im};
}

// Source-plane image of a critical point through the complex lens equation,
// binary_lens_equation (point_source_magnifier.cpp:780-792).
__device__ __forceinline__ Pt lens_equation_complex(const Lens& L, cx z)
{
    const cx zc = conj(z);
    const cx d1{zc.re - L.a, zc.im - 0.0};
    const cx t1 = std_cdiv(L.m1, 0.0, d1.re, d1.im);
    const cx t2 = std_cdiv(L.m2, 0.0, zc.re, zc.im);
    const cx shifted{(z.re - t1.re) - t2.re, (z.im - t1.im) - t2.im};
    const cx am1{L.a * L.m1 - 0.0 * 0.0, L.a * 0.0 + 0.0 * L.m1};
    return Pt{shifted.re - am1.re, shifted.im - am1.im};
}

// map_binary_lens_real (finite_source_magnifier.cpp:278-291), divide form.
__device__ __forceinline__ Pt map_real(const Lens& L, double x, double y)
{
    const double xa = x - L.a;
    const double den1 = xa * xa + y * y;
    const double den2 = x * x + y * y;
    return Pt{x - L.m1 * xa / den1 - L.m2 * x / den2 - L.a * L.m1,
              y - L.m1 * y / den1 - L.m2 * y / den2};
}

// mapped_binary_lens_distance2 (finite_source_magnifier.cpp:65-80)
__device__ __forceinline__ double mapped_dist2(const Lens& L, double x, double y, Pt src)
{
    const Pt w = map_real(L, x, y);
    const double dx = w.x - src.x;
    const double dy = w.y - src.y;
    return dx * dx + dy * dy;
}

// evaluate_binary_lens_cell (finite_source_magnifier.cpp:121-153), reciprocal
// form with the Jacobian; used for fold-guarded Cartesian cells.
__device__ __forceinline__ double mapped_dist2_jac(const Lens& L, double x, double y, Pt src,
                                                   double* jac)
{
    const double xa = x - L.a;
    const double y2 = y * y;
    const double den1 = xa * xa + y2;
    const double den2 = x * x + y2;
    const double inv_den1 = 1.0 / den1;
    const double inv_den2 = 1.0 / den2;
    const double mx = x - L.m1 * xa * inv_den1 - L.m2 * x * inv_den2 - L.a * L.m1;
    const double my = y - L.m1 * y * inv_den1 - L.m2 * y * inv_den2;
    const double dx = mx - src.x;
    const double dy = my - src.y;
    double j = 0.0;
    if (den1 >= 1.0e-20 && den2 >= 1.0e-20) {
        const double i1 = inv_den1 * inv_den1;
        const double i2 = inv_den2 * inv_den2;
        const double re_f = L.m1 * (xa * xa - y2) * i1 + L.m2 * (x * x - y2) * i2;
        const double im_f = -2.0 * y * (L.m1 * xa * i1 + L.m2 * x * i2);
        j = 1.0 - re_f * re_f - im_f * im_f;
    }
    *jac = j;
    return dx * dx + dy * dy;
}

// binary_jacobian (finite_source_magnifier.cpp:196-209)
__device__ __forceinline__ double jacobian_real(const Lens& L, double x, double y)
{
    const double xa = x - L.a;
    const double den1 = xa * xa + y * y;
    const double den2 = x * x + y * y;
    if (den1 < 1.0e-20 || den2 < 1.0e-20) return 0.0;
    const double den1sq = den1 * den1;
    const double den2sq = den2 * den2;
    const double re_f = L.m1 * (xa * xa - y * y) / den1sq + L.m2 * (x * x - y * y) / den2sq;
    const double im_f = -2.0 * y * (L.m1 * xa / den1sq + L.m2 * x / den2sq);
    return 1.0 - re_f * re_f - im_f * im_f;
}

__device__ __forceinline__ int sign_of(double v) { return v > 0.0 ? 1 : v < 0.0 ? -1 : 0; }

// Limb-darkening brightness at a normalised squared radius, as the nearest node
// of the reference's 20001-entry table (finite_source_magnifier.cpp:4001-4012),
// the node value computed by source_surface_brightness (:369-375) on the fly.
__device__ __forceinline__ double ld_node_value(int i, double c, double d)
{
    const double r2 = static_cast<double>(i) / 20000.0;
    const double b = dclamp(r2, 0.0, 1.0);
    const double mu = __builtin_sqrt(dmax(0.0, 1.0 - b));
    // with d = 0 the last term is +0, and subtracting +0 changes no value
    if (d == 0.0) return 1.0 - c * (1.0 - mu);
    return 1.0 - c * (1.0 - mu) - d * (1.0 - __builtin_sqrt(mu));
}

__device__ __forceinline__ double ld_brightness(double r2, double c, double d)
{
    int i;
    if (r2 <= 0.0)
        i = 0;
    else if (r2 >= 1.0)
        i = 20000;
    else
        i = static_cast<int>(r2 * 20000.0 + 0.5);
    return ld_node_value(i, c, d);
}

// source_flux (finite_source_magnifier.cpp:377-384)
__device__ __forceinline__ double source_flux(double rho, double c, double d)
{
    const double ff = 1.0 - c / 3.0 - d / 5.0;
    if (ff <= 0.0 || !isfinite(ff)) return __builtin_nan("");
    return kPi * rho * rho * ff;
}

// hypot(x, y) as glibc computes it (the reference calls std::hypot): the
// square root of the sum of squares corrected by one Newton-type step, the
// non-FMA kernel of Borges, "An Improved Algorithm for hypot(a,b)" (2019),
// with the scaling thresholds of glibc's implementation.  Bitwise equal to the
// host libm on 2e7 random pairs, including nearly equal arguments where a
// correctly rounded hypot would differ.
__device__ __forceinline__ double hypot_kernel(double ax, double ay)
{
    double t1, t2;
    double h = __builtin_sqrt(ax * ax + ay * ay);
    if (h <= 2.0 * ay) {
        const double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        const double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}

__device__ __forceinline__ double hypot_glibc(double x, double y)
{
    if (!isfinite(x) || !isfinite(y)) {
        if (isinf(x) || isinf(y)) return __builtin_inf();
        return x + y;
    }
    x = fabs(x);
    y = fabs(y);
    const double ax = x < y ? y : x;
    const double ay = x < y ? x : y;
    const double kScale = 0x1p-600, kLarge = 0x1p+511, kTiny = 0x1p-511, kEps = 0x1p-54;
    if (ax > kLarge) {
        if (ay <= ax * kEps) return ax + ay;
        return hypot_kernel(ax * kScale, ay * kScale) / kScale;
    }
    if (ay < kTiny) {
        if (ax >= ay / kEps) return ax + ay;
        return hypot_kernel(ax / kScale, ay / kScale) * kScale;
    }
    if (ay <= ax * kEps) return ax + ay;
    return hypot_kernel(ax, ay);
}

// point_segment_distance (finite_source_magnifier.cpp:330-342)
__device__ __forceinline__ double point_segment_distance(Pt p, Pt a, Pt b)
{
    const double dx = b.x - a.x;
    const double dy = b.y - a.y;
    const double length2 = dx * dx + dy * dy;
    if (length2 == 0.0) return __builtin_sqrt(dist2(p, a));
    const double t = dclamp(((p.x - a.x) * dx + (p.y - a.y) * dy) / length2, 0.0, 1.0);
    return hypot_glibc(p.x - (a.x + t * dx), p.y - (a.y + t * dy));
}

// Quartic of the critical curve at one phase, solved as PolynomialRootSolver::solve
// does (polynomial_roots.cpp:35-73: degree 4, polish, zero start).
__device__ __forceinline__ void critical_roots(const Lens& L, cx phase, cx* roots)
{
    cx c[5];
    critical_quartic(L, phase, c);
    roots_gen<4>(roots, c, false);
}

}  // namespace lcbg
