// Binary-lens point-source layer for gfx950, FP64, one thread per source
// position: VBM-convention geometry, the lens-equation quintic, its roots, the
// physical-image selection rules and the magnification sum.
//
// Reference algorithms: point_source_magnifier.cpp (make_fast_vbm_geometry :105,
// binary_polynomial_coefficients :152, residual_squared :187,
// jacobian_determinant :205, select_physical_roots :225,
// point_source_result_from_roots :281, derivative_error_indicator_from_roots :313,
// binary_image_candidates :739).  Expressions keep the reference's operation
// order so results are bitwise the reference's (built with -ffp-contract=off).
#pragma once

#include "sg_solver.hpp"

namespace lcbg {

struct Geom {
    double a, m1, m2, a2, a3, m2_2;
    cx y;
};

// make_fast_vbm_geometry (point_source_magnifier.cpp:105-115): the lighter lens
// (mass m2) sits at the origin, the heavier at z = a.
__device__ __forceinline__ Geom make_geom(double separation, double mass_ratio, double sx,
                                          double sy)
{
    Geom g;
    const double s = fabs(separation);
    const double q_input = fabs(mass_ratio);
    const double q = q_input < 1.0 ? q_input : 1.0 / q_input;
    g.a = q_input < 1.0 ? -s : s;
    g.m1 = 1.0 / (1.0 + q);
    g.m2 = q * g.m1;
    g.a2 = g.a * g.a;
    g.a3 = g.a2 * g.a;
    g.m2_2 = g.m2 * g.m2;
    g.y = cx{sx + g.a * g.m1, sy};
    return g;
}

// binary_polynomial_coefficients (point_source_magnifier.cpp:152-185)
__device__ __forceinline__ void quintic_coefficients(const Geom& g, cx* c)
{
    const double a = g.a, m2 = g.m2, a2 = g.a2, a3 = g.a3;
    const cx y = g.y;
    const cx yc = conj(y);
    const double c9 = a2 * g.m2_2;
    const double c10 = a * m2;
    const cx c12 = cx{a, 0.0} - yc;
    const cx c13 = cx{a, 0.0} + y;
    const cx c14 = c13 + y;
    const cx c15 = conj(c14);
    const cx c16 = a * y;
    const cx c17 = conj(c16);
    const cx c18 = conj(c12);
    c[0] = c9 * y;
    c[1] = cx{-c9, 0.0} + c10 * (cx{a, 0.0} + ((2.0 * c17 - cx{2.0 + a2, 0.0}) * y));
    c[2] = c10 * ((cx{1.0, 0.0} + c16) - (2.0 * yc) * c13) -
           (c17 - cx{1.0, 0.0}) * (c16 * c12 - c18);
    c[3] = (c10 * c15 + ((cx{a3, 0.0} + (2.0 * (1.0 + a2)) * y) - c17 * c14) * yc) - a * c13;
    c[4] = cx{-c10, 0.0} - c12 * (yc * (c13 + cx{a, 0.0}) - cx{1.0, 0.0});
    c[5] = yc * c12;
}

// residual_squared (point_source_magnifier.cpp:187-203)
__device__ __forceinline__ double residual_squared(const Geom& g, cx z)
{
    const double u1 = z.re - g.a;
    const double v1 = -z.im;
    const double d1 = u1 * u1 + v1 * v1;
    const double u2 = z.re;
    const double v2 = -z.im;
    const double d2 = u2 * u2 + v2 * v2;
    const double r_re = g.y.re - z.re + g.m1 * u1 / d1 + g.m2 * u2 / d2;
    const double r_im = g.y.im - z.im - g.m1 * v1 / d1 - g.m2 * v2 / d2;
    return r_re * r_re + r_im * r_im;
}

// jacobian_determinant (point_source_magnifier.cpp:205-223)
__device__ __forceinline__ double jacobian_det(const Geom& g, cx z)
{
    const double u1 = z.re - g.a;
    const double v1 = z.im;
    const double d1 = u1 * u1 + v1 * v1;
    const double d1_2 = d1 * d1;
    const double u2 = z.re;
    const double v2 = z.im;
    const double d2 = u2 * u2 + v2 * v2;
    const double d2_2 = d2 * d2;
    const double der_re = g.m1 * (u1 * u1 - v1 * v1) / d1_2 + g.m2 * (u2 * u2 - v2 * v2) / d2_2;
    const double der_im = -2.0 * g.m1 * u1 * v1 / d1_2 - 2.0 * g.m2 * u2 * v2 / d2_2;
    return 1.0 - (der_re * der_re + der_im * der_im);
}

// select_physical_roots (point_source_magnifier.cpp:225-279): drop the two
// largest-residual roots when the residual gap says there are three images.
// Returns a 5-bit mask of physical roots.
__device__ __forceinline__ unsigned select_physical(const Geom& g, const cx* roots)
{
    double res[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) res[i] = residual_squared(g, roots[i]);
    int w1 = 0, w2 = 0, w3 = 0;
    double r1 = res[0], r2 = 0.0, r3 = 0.0;
    // i == 1
    if (res[1] > r1) {
        w2 = w1; r2 = r1; w1 = 1; r1 = res[1];
    } else {
        w2 = 1; r2 = res[1];
    }
    // i == 2
    if (res[2] > r1) {
        w3 = w2; r3 = r2; w2 = w1; r2 = r1; w1 = 2; r1 = res[2];
    } else if (res[2] > r2) {
        w3 = w2; r3 = r2; w2 = 2; r2 = res[2];
    } else {
        w3 = 2; r3 = res[2];
    }
#pragma unroll
    for (int i = 3; i < 5; ++i) {
        if (res[i] > r1) {
            w3 = w2; r3 = r2; w2 = w1; r2 = r1; w1 = i; r1 = res[i];
        } else if (res[i] > r2) {
            w3 = w2; r3 = r2; w2 = i; r2 = res[i];
        } else if (res[i] > r3) {
            w3 = i; r3 = res[i];
        }
    }
    const bool three = r2 * 1.0e-8 > r3 + 1.0e-24;
    unsigned mask = 0x1fu;
    if (three) mask &= ~((1u << w1) | (1u << w2));
    return mask;
}

__device__ __forceinline__ void solve_quintic(const Geom& g, cx* roots)
{
    cx c[6];
    quintic_coefficients(g, c);
    roots_gen<5>(roots, c, false);
}

struct PointMag {
    double mag;
    int images;
};

// point_source_result_from_roots (point_source_magnifier.cpp:281-297)
__device__ __forceinline__ PointMag ps_from_roots(const Geom& g, const cx* roots, unsigned mask)
{
    double mag = 0.0;
    int n = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        if (mask & (1u << i)) {
            mag += 1.0 / fabs(jacobian_det(g, roots[i]));
            ++n;
        }
    }
    return PointMag{mag, n};
}

// Point-source magnification as LensModel::magnification sums it
// (src/lcbinint/model/lens_model.cpp:199-205) over binary_images
// (point_source_magnifier.cpp:723-778): roots sorted by residual (a stable
// sort, as std::sort's insertion sort orders five elements), three images if
// r[3] * 1e-4 > r[2] + 1e-12, else five; sum of 1 / |J| in that order.
__device__ __forceinline__ double ps_from_images(const Geom& g, const cx* roots)
{
    double res[5];
    int ord[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        res[i] = __builtin_sqrt(residual_squared(g, roots[i]));
        ord[i] = i;
    }
    for (int i = 1; i < 5; ++i) {
        const int v = ord[i];
        const double rv = res[v];
        int j = i;
        while (j > 0 && rv < res[ord[j - 1]]) {
            ord[j] = ord[j - 1];
            --j;
        }
        ord[j] = v;
    }
    const int physical = (res[ord[3]] * 1.0e-4 > res[ord[2]] + 1.0e-12) ? 3 : 5;
    double mag = 0.0;
    for (int i = 0; i < physical; ++i) mag += 1.0 / fabs(jacobian_det(g, roots[ord[i]]));
    return mag;
}

// Number of images binary_images keeps (point_source_magnifier.cpp:723-778):
// three if r[3] * 1e-4 > r[2] + 1e-12 over the residuals sorted ascending, else five.
__device__ __forceinline__ int binary_images_count(const Geom& g, const cx* roots)
{
    double res[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) res[i] = __builtin_sqrt(residual_squared(g, roots[i]));
    // the third and fourth smallest residuals of five
    for (int i = 1; i < 5; ++i) {
        const double v = res[i];
        int j = i;
        while (j > 0 && v < res[j - 1]) {
            res[j] = res[j - 1];
            --j;
        }
        res[j] = v;
    }
    return (res[3] * 1.0e-4 > res[2] + 1.0e-12) ? 3 : 5;
}

// binary_mag0 (point_source_magnifier.cpp:567 -> 652-693), no root cache.
__device__ __forceinline__ PointMag binary_mag0(double s, double q, double x, double y)
{
    if (s == 0.0 || q <= 0.0) return PointMag{0.0, 0};
    const Geom g = make_geom(s, q, x, y);
    cx roots[5];
    solve_quintic(g, roots);
    return ps_from_roots(g, roots, select_physical(g, roots));
}

// derivative_error_indicator_from_roots (point_source_magnifier.cpp:313-359)
__device__ __forceinline__ double derivative_indicator(const Geom& g, const cx* roots,
                                                       unsigned mask)
{
    double indicator = 0.0;
    const cx a{g.a, 0.0}, m1{g.m1, 0.0}, m2{g.m2, 0.0};
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        if (!(mask & (1u << i))) continue;
        const cx z = roots[i];
        const cx dza = z - a;
        const cx za2 = dza * dza;
        const cx zb2 = z * z;
        const cx j1 = m1 / za2 + m2 / zb2;
        const cx j1c = conj(j1);
        const double det_j = (1.0 - j1 * j1c).re;
        if (det_j == 0.0 || !isfinite(det_j)) return __builtin_inf();
        cx j2 = -2.0 * (m1 / (za2 * dza) + m2 / (zb2 * z));
        cx j3 = 6.0 * (m1 / (za2 * za2) + m2 / (zb2 * zb2));
        const double det_j2 = det_j * det_j;
        const cx j1c2 = j1c * j1c;
        j3 = j3 * j1c2;
        const double ob2 = (j2.re * j2.re + j2.im * j2.im) * (6.0 - 6.0 * det_j + det_j2);
        j2 = ((j2 * j2) * j1c2) * j1c;
        const double denominator = fabs(det_j * det_j2 * det_j2);
        if (denominator == 0.0 || !isfinite(denominator)) return __builtin_inf();
        const double cq =
            0.5 * (fabs(ob2 - 6.0 * j2.re - 2.0 * j3.re * det_j) + 3.0 * fabs(j2.im)) /
            denominator;
        if (isfinite(cq))
            indicator += cq;
        else
            return __builtin_inf();
    }
    return indicator;
}

}  // namespace lcbg
