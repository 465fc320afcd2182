/*
 * lcb_oracle.c - CPU restatement of the reference lcbinint binary-lens hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never by the product.  Built with
 * -ffp-contract=off (no fused multiply-add), like the reference's x86-64 build,
 * so that every operation rounds as the reference's does.
 *
 * Each function restates the reference routine cited above it; the operation
 * order follows the reference expression by expression because the parity
 * contract is bitwise for the point-source layer.
 */
#define _GNU_SOURCE
#include "lcb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- counting */

static double g_flops = 0.0;
double lcbo_flop_count(void) { return g_flops; }
void lcbo_flop_reset(void) { g_flops = 0.0; }
#define FL(n) (g_flops += (double)(n))

/* ------------------------------------------------- SG complex (SkowronGould.h:10-83) */

typedef struct { double re, im; } cx;

static inline cx C(double re, double im) { cx z = {re, im}; return z; }
static inline cx cadd(cx a, cx b) { FL(2); return C(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { FL(2); return C(a.re - b.re, a.im - b.im); }
static inline cx cmul(cx a, cx b)
{
    FL(6);
    return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx cdiv(cx p1, cx p2) /* SkowronGould.h:50-55 */
{
    FL(11);
    double md = p2.re * p2.re + p2.im * p2.im;
    return C((p1.re * p2.re + p1.im * p2.im) / md, (p1.im * p2.re - p1.re * p2.im) / md);
}
static inline cx cscale(double a, cx z) { FL(2); return C(a * z.re, a * z.im); }
static inline cx caddd(double a, cx z) { FL(1); return C(z.re + a, z.im); }
static inline cx cdsub(double a, cx z) { FL(1); return C(a - z.re, -z.im); }
static inline cx cneg(cx z) { return C(-z.re, -z.im); }
static inline cx cconj(cx z) { return C(z.re, -z.im); }
static inline double cabs2(cx z) { FL(3); return z.re * z.re + z.im * z.im; }
static inline double cabsv(cx z) { FL(4); return sqrt(z.re * z.re + z.im * z.im); }
static inline int ceq(cx a, cx b) { return a.re == b.re && a.im == b.im; }
static inline cx csqrt_sg(cx z) /* SkowronGould.h:23-30 */
{
    FL(10);
    double md = sqrt(z.re * z.re + z.im * z.im);
    return (md > 0.0) ? C(sqrt((md + z.re) / 2.0),
                          sqrt((md - z.re) / 2.0) * ((z.im > 0.0) ? 1.0 : -1.0))
                      : C(0.0, 0.0);
}
static inline cx cexp_sg(cx z) /* SkowronGould.h:39-43 */
{
    double r = exp(z.re);
    return C(r * cos(z.im), r * sin(z.im));
}
/* real(conj(a) * b), written as the reference's operator chain evaluates it */
static inline double creal_conj_mul(cx a, cx b)
{
    cx c = cconj(a);
    FL(3);
    return c.re * b.re - c.im * b.im;
}

#define MAXIT 80
static const double FRAC_JUMPS[10] = {0.64109297, 0.91577881, 0.25921289, 0.50487203,
                                      0.08177045, 0.13653241, 0.306162,   0.37794326,
                                      0.04618805, 0.75132137};

/* cmplx_laguerre (SkowronGould.cpp:472-624) */
static void sg_laguerre(const cx* poly, int degree, cx* root, int* success)
{
    const double FRAC_ERR = 2.0e-15;
    if (ceq(poly[degree], C(0.0, 0.0))) {
        if (degree == 0) return;
        sg_laguerre(poly, degree - 1, root, success);
    }
    if (degree <= 1) {
        if (degree == 0) {
            *success = 0;
            return;
        }
        *root = cdiv(cneg(poly[0]), poly[1]);
        return;
    }
    int good_to_go = 0;
    double one_nth = 1.0 / degree;
    double n_1_nth = (degree - 1.0) * one_nth;
    double two_n_div_n_1 = 2.0 / n_1_nth;
    cx c_one_nth = C(one_nth, 0.0);
    cx c_one = C(1.0, 0.0);
    cx zero = C(0.0, 0.0);
    cx fac_newton = zero;
    FL(3);
    for (int i = 1; i <= MAXIT; i++) {
        double ek = cabsv(poly[degree]);
        double absroot = cabsv(*root);
        cx p = poly[degree], dp = zero, d2p_half = zero;
        for (int k = degree - 1; k >= 0; k--) {
            d2p_half = cadd(dp, cmul(d2p_half, *root));
            dp = cadd(p, cmul(dp, *root));
            p = cadd(poly[k], cmul(p, *root));
            ek = absroot * ek + cabsv(p);
            FL(2);
        }
        double abs2p = creal_conj_mul(p, p);
        if (abs2p == 0) return;
        double sc = FRAC_ERR * ek;
        double stopping_crit2 = sc * sc;
        FL(2);
        if (abs2p < stopping_crit2) {
            FL(1);
            if (abs2p < 0.01 * stopping_crit2) return;
            good_to_go = 1;
        } else {
            good_to_go = 0;
        }
        cx denom = zero;
        if (!ceq(dp, zero)) {
            fac_newton = cdiv(p, dp);
            cx fac_extra = cdiv(d2p_half, dp);
            cx F_half = cmul(fac_newton, fac_extra);
            cx denom_sqrt = csqrt_sg(csub(c_one, cscale(two_n_div_n_1, F_half)));
            if (denom_sqrt.re >= 0.0)
                denom = cadd(c_one_nth, cscale(n_1_nth, denom_sqrt));
            else
                denom = csub(c_one_nth, cscale(n_1_nth, denom_sqrt));
        }
        cx dx;
        if (ceq(denom, C(0.0, 0.0))) {
            dx = cscale(absroot + 1.0, cexp_sg(C(0.0, FRAC_JUMPS[i % 10] * 2 * M_PI)));
        } else {
            dx = cdiv(fac_newton, denom);
        }
        cx newroot = csub(*root, dx);
        if (ceq(newroot, *root)) return;
        if (good_to_go) {
            *root = newroot;
            return;
        }
        if (i % 10 == 0) {
            double faq = FRAC_JUMPS[(i / 10 - 1) % 10];
            newroot = csub(*root, cscale(faq, dx));
        }
        *root = newroot;
    }
    *success = 0;
}

/* cmplx_newton_spec (SkowronGould.cpp:626-762) */
static void sg_newton_spec(const cx* poly, int degree, cx* root, int* success)
{
    const double FRAC_ERR = 2e-15;
    cx zero = C(0.0, 0.0);
    *success = 1;
    if (ceq(poly[degree], zero)) {
        /* the reference recurses with the same degree here; this branch is
           unreachable for the lens polynomials (leading coefficient != 0). */
        return;
    }
    if (degree <= 1) {
        if (degree == 0) {
            *success = 0;
            return;
        }
        *root = cdiv(cneg(poly[0]), poly[1]);
        return;
    }
    int good_to_go = 0;
    double stopping_crit2 = 0.0;
    for (int i = 1; i <= MAXIT; i++) {
        cx p = poly[degree], dp = zero;
        if (i % 10 == 1) {
            double ek = cabsv(poly[degree]);
            double absroot = cabsv(*root);
            for (int k = degree - 1; k >= 0; k--) {
                dp = cadd(p, cmul(dp, *root));
                p = cadd(poly[k], cmul(p, *root));
                ek = absroot * ek + cabsv(p);
                FL(2);
            }
            double sc = FRAC_ERR * ek;
            stopping_crit2 = sc * sc;
            FL(2);
        } else {
            for (int k = degree - 1; k >= 0; k--) {
                dp = cadd(p, cmul(dp, *root));
                p = cadd(poly[k], cmul(p, *root));
            }
        }
        double abs2p = creal_conj_mul(p, p);
        if (abs2p == 0.0) return;
        if (abs2p < stopping_crit2) {
            if (ceq(dp, zero)) return;
            FL(1);
            if (abs2p < 0.01 * stopping_crit2) return;
            good_to_go = 1;
        } else {
            good_to_go = 0;
        }
        cx dx;
        if (ceq(dp, zero)) {
            FL(1);
            dx = cscale(cabsv(*root) + 1.0, cexp_sg(C(0.0, FRAC_JUMPS[i % 10] * 2 * M_PI)));
        } else {
            dx = cdiv(p, dp);
        }
        cx newroot = csub(*root, dx);
        if (ceq(newroot, *root)) return;
        if (good_to_go) {
            *root = newroot;
            return;
        }
        if (i % 10 == 0) {
            double faq = FRAC_JUMPS[(i / 10 - 1) % 10];
            newroot = csub(*root, cscale(faq, dx));
        }
        *root = newroot;
    }
    *success = 0;
}

/* cmplx_laguerre2newton (SkowronGould.cpp:764-1130) with starting_mode 2 */
static void sg_laguerre2newton(const cx* poly, int degree, cx* root, int* success)
{
    const double FRAC_ERR = 2.0e-15;
    cx zero = C(0.0, 0.0);
    cx c_one = C(1.0, 0.0);
    int iter = 0;
    *success = 1;
    double stopping_crit2 = 0;
    int j = 1;
    int good_to_go = 0;
    int mode = 2;
    int i;
    double ek, absroot, abs2p;
    cx p, dp, d2p_half, dx, newroot, fac_netwon = zero, fac_extra, F_half;
    for (;;) {
        if (mode >= 2) {
            double one_nth = 1.0 / (degree);
            double n_1_nth = (degree - 1) * one_nth;
            double two_n_div_n_1 = 2.0 / n_1_nth;
            cx c_one_nth = C(one_nth, 0.0);
            FL(3);
            for (i = 1; i <= MAXIT; i++) {
                ek = cabsv(poly[degree]);
                absroot = cabsv(*root);
                p = poly[degree];
                dp = zero;
                d2p_half = zero;
                for (int k = degree; k >= 1; k--) {
                    d2p_half = cadd(dp, cmul(d2p_half, *root));
                    dp = cadd(p, cmul(dp, *root));
                    p = cadd(poly[k - 1], cmul(p, *root));
                    ek = absroot * ek + cabsv(p);
                    FL(2);
                }
                abs2p = creal_conj_mul(p, p);
                iter = iter + 1;
                if (abs2p == 0) return;
                {
                    double sc = FRAC_ERR * ek;
                    stopping_crit2 = sc * sc;
                    FL(2);
                }
                if (abs2p < stopping_crit2) {
                    FL(1);
                    if (abs2p < 0.01 * stopping_crit2) return;
                    good_to_go = 1;
                } else {
                    good_to_go = 0;
                }
                cx denom = zero;
                if (!ceq(dp, zero)) {
                    fac_netwon = cdiv(p, dp);
                    fac_extra = cdiv(d2p_half, dp);
                    F_half = cmul(fac_netwon, fac_extra);
                    double abs2_F_half = creal_conj_mul(F_half, F_half);
                    if (abs2_F_half <= 0.0625) {
                        if (abs2_F_half <= 0.000625)
                            mode = 0;
                        else
                            mode = 1;
                    }
                    cx denom_sqrt = csqrt_sg(csub(c_one, cscale(two_n_div_n_1, F_half)));
                    if (denom_sqrt.re > 0.0)
                        denom = cadd(c_one_nth, cscale(n_1_nth, denom_sqrt));
                    else
                        denom = csub(c_one_nth, cscale(n_1_nth, denom_sqrt));
                }
                if (ceq(denom, zero)) {
                    /* the reference adds (not multiplies) the jump here */
                    FL(1);
                    dx = caddd(cabsv(*root) + 1.0,
                               cexp_sg(C(0.0, FRAC_JUMPS[i % 10] * 2 * M_PI)));
                } else {
                    dx = cdiv(fac_netwon, denom);
                }
                newroot = csub(*root, dx);
                if (ceq(newroot, *root)) return;
                if (good_to_go) {
                    *root = newroot;
                    return;
                }
                if (mode != 2) {
                    *root = newroot;
                    j = i + 1;
                    break;
                }
                if ((i % 10) == 0) {
                    double faq = FRAC_JUMPS[((i / 10 - 1) % 10)];
                    newroot = csub(*root, cscale(faq, dx));
                }
                *root = newroot;
            }
            if (i >= MAXIT) {
                *success = 0;
                return;
            }
        }
        if (mode == 1) {
            for (i = j; i <= MAXIT; i++) {
                p = poly[degree];
                dp = zero;
                d2p_half = zero;
                if ((i - j) % 10 == 0) {
                    ek = cabsv(poly[degree]);
                    absroot = cabsv(*root);
                    for (int k = degree; k >= 1; k--) {
                        d2p_half = cadd(dp, cmul(d2p_half, *root));
                        dp = cadd(p, cmul(dp, *root));
                        p = cadd(poly[k - 1], cmul(p, *root));
                        ek = absroot * ek + cabsv(p);
                        FL(2);
                    }
                    double sc = FRAC_ERR * ek;
                    stopping_crit2 = sc * sc;
                    FL(2);
                } else {
                    for (int k = degree; k >= 1; k--) {
                        d2p_half = cadd(dp, cmul(d2p_half, *root));
                        dp = cadd(p, cmul(dp, *root));
                        p = cadd(poly[k - 1], cmul(p, *root));
                    }
                }
                abs2p = creal_conj_mul(p, p);
                iter = iter + 1;
                if (abs2p == 0.0) return;
                if (abs2p < stopping_crit2) {
                    if (ceq(dp, zero)) return;
                    FL(1);
                    if (abs2p < 0.01 * stopping_crit2) return;
                    good_to_go = 1;
                } else {
                    good_to_go = 0;
                }
                if (ceq(dp, zero)) {
                    FL(1);
                    dx = cscale(cabsv(*root) + 1.0,
                                cexp_sg(C(0.0, FRAC_JUMPS[i % 10] * 2 * M_PI)));
                } else {
                    fac_netwon = cdiv(p, dp);
                    fac_extra = cdiv(d2p_half, dp);
                    F_half = cmul(fac_netwon, fac_extra);
                    double abs2_F_half = creal_conj_mul(F_half, F_half);
                    if (abs2_F_half <= 0.000625) mode = 0;
                    dx = cmul(fac_netwon, cadd(c_one, F_half));
                }
                newroot = csub(*root, dx);
                if (ceq(newroot, *root)) return;
                if (good_to_go) {
                    *root = newroot;
                    return;
                }
                if (mode != 1) {
                    *root = newroot;
                    j = i + 1;
                    break;
                }
                if ((i % 10) == 0) {
                    double faq = FRAC_JUMPS[(i / 10 - 1) % 10];
                    newroot = csub(*root, cscale(faq, dx));
                }
                *root = newroot;
            }
            if (i >= MAXIT) {
                *success = 0;
                return;
            }
        }
        if (mode == 0) {
            for (i = j; i <= j + 10; i++) {
                p = poly[degree];
                dp = zero;
                if (i == j) {
                    ek = cabsv(poly[degree]);
                    absroot = cabsv(*root);
                    for (int k = degree; k >= 1; k--) {
                        dp = cadd(p, cmul(dp, *root));
                        p = cadd(poly[k - 1], cmul(p, *root));
                        ek = absroot * ek + cabsv(p);
                        FL(2);
                    }
                    double sc = FRAC_ERR * ek;
                    stopping_crit2 = sc * sc;
                    FL(2);
                } else {
                    for (int k = degree; k >= 1; k--) {
                        dp = cadd(p, cmul(dp, *root));
                        p = cadd(poly[k - 1], cmul(p, *root));
                    }
                }
                abs2p = creal_conj_mul(p, p);
                iter = iter + 1;
                if (abs2p == 0.0) return;
                if (abs2p < stopping_crit2) {
                    if (ceq(dp, zero)) return;
                    FL(1);
                    if (abs2p < 0.01 * stopping_crit2) return;
                    good_to_go = 1;
                } else {
                    good_to_go = 0;
                }
                if (ceq(dp, zero)) {
                    FL(1);
                    dx = cscale(cabsv(*root) + 1.0,
                                cexp_sg(C(0.0, 2 * M_PI * FRAC_JUMPS[i % 10])));
                } else {
                    dx = cdiv(p, dp);
                }
                newroot = csub(*root, dx);
                if (ceq(newroot, *root)) return;
                if (good_to_go) {
                    *root = newroot;
                    return;
                }
                *root = newroot;
            }
            if (iter >= MAXIT) {
                *success = 0;
                return;
            }
            mode = 2;
        }
    }
}

/* solve_quadratic_eq (SkowronGould.cpp:1132-1154) */
static void sg_solve_quadratic(cx* x0, cx* x1, const cx* poly)
{
    cx a = poly[2], b = poly[1], c = poly[0];
    cx b2 = cmul(b, b);
    cx four_a = C(4 * a.re, 4 * a.im);
    FL(2);
    cx delta = csqrt_sg(csub(b2, cmul(four_a, c)));
    if (creal_conj_mul(b, delta) >= 0)
        *x0 = cscale(-0.5, cadd(b, delta));
    else
        *x0 = cscale(-0.5, csub(b, delta));
    if (ceq(*x0, C(0.0, 0.0))) {
        *x1 = C(0.0, 0.0);
    } else {
        *x1 = cdiv(c, *x0);
        *x0 = cdiv(*x0, a);
    }
}

/* cmplx_roots_gen (SkowronGould.cpp:79-168) */
static void sg_roots_gen(cx* roots, const cx* poly, int degree, int polish, int use_start)
{
    cx poly2[30];
    int success = 1;
    if (!use_start) {
        for (int jj = 0; jj < degree; jj++) roots[jj] = C(0.0, 0.0);
    }
    for (int j = 0; j <= degree; j++) poly2[j] = poly[j];
    if (degree <= 1) {
        if (degree == 1) roots[0] = cdiv(cneg(poly[0]), poly[1]);
        return;
    }
    for (int n = degree; n >= 3; n--) {
        int ismallest = n - 1;
        double abssmall = cabs2(roots[ismallest]);
        for (int ii = 0; ii < n - 1; ii++) {
            double v = cabs2(roots[ii]);
            if (v < abssmall) {
                ismallest = ii;
                abssmall = v;
            }
        }
        cx coef = roots[ismallest];
        roots[ismallest] = roots[n - 1];
        roots[n - 1] = coef;
        sg_laguerre2newton(poly2, n, &roots[n - 1], &success);
        if (!success) {
            roots[n - 1] = C(0.0, 0.0);
            sg_laguerre(poly2, n, &roots[n - 1], &success);
        }
        coef = poly2[n];
        for (int i = n - 1; i >= 0; i--) {
            cx prev = poly2[i];
            poly2[i] = coef;
            coef = cadd(prev, cmul(roots[n - 1], coef));
        }
    }
    sg_solve_quadratic(&roots[1], &roots[0], poly2);
    if (polish) {
        for (int n = 0; n < degree - 1; n++) sg_newton_spec(poly, degree, &roots[n], &success);
    }
}

void lcbo_cmplx_roots_gen(const double* coeffs, int degree, int polish, int use_start,
                          double* roots)
{
    cx poly[30], r[30];
    for (int i = 0; i <= degree; i++) poly[i] = C(coeffs[2 * i], coeffs[2 * i + 1]);
    for (int i = 0; i < degree; i++) r[i] = C(roots[2 * i], roots[2 * i + 1]);
    sg_roots_gen(r, poly, degree, polish, use_start);
    for (int i = 0; i < degree; i++) {
        roots[2 * i] = r[i].re;
        roots[2 * i + 1] = r[i].im;
    }
}

/* ------------------------------------------ binary point source (PSM) */

typedef struct {
    double a, m1, m2, a2, a3, m2_2;
    cx y;
} geom;

/* make_fast_vbm_geometry (point_source_magnifier.cpp:105-115) */
static geom make_geom(double separation, double mass_ratio, double sx, double sy)
{
    geom g;
    const double s = fabs(separation);
    const double q_input = fabs(mass_ratio);
    const double q = q_input < 1.0 ? q_input : 1.0 / q_input;
    g.a = q_input < 1.0 ? -s : s;
    g.m1 = 1.0 / (1.0 + q);
    g.m2 = q * g.m1;
    g.a2 = g.a * g.a;
    g.a3 = g.a2 * g.a;
    g.m2_2 = g.m2 * g.m2;
    g.y = C(sx + g.a * g.m1, sy);
    FL(9);
    return g;
}

/* binary_polynomial_coefficients (point_source_magnifier.cpp:152-185) */
static void quintic_coefficients(const geom* g, cx* c)
{
    const double a = g->a, m2 = g->m2, a2 = g->a2, a3 = g->a3;
    const cx y = g->y;
    const cx yc = cconj(y);
    const double c9 = a2 * g->m2_2;
    const double c10 = a * m2;
    FL(2);
    const cx c12 = csub(C(a, 0.0), yc);
    const cx c13 = cadd(C(a, 0.0), y);
    const cx c14 = cadd(c13, y);
    const cx c15 = cconj(c14);
    const cx c16 = cscale(a, y);
    const cx c17 = cconj(c16);
    const cx c18 = cconj(c12);
    c[0] = cscale(c9, y);
    {
        cx t = csub(cscale(2.0, c17), C(2.0 + a2, 0.0));
        FL(1);
        c[1] = cadd(C(-c9, 0.0), cscale(c10, cadd(C(a, 0.0), cmul(t, y))));
    }
    c[2] = csub(cscale(c10, csub(cadd(C(1.0, 0.0), c16), cmul(cscale(2.0, yc), c13))),
                cmul(csub(c17, C(1.0, 0.0)), csub(cmul(c16, c12), c18)));
    {
        double k = 2.0 * (1.0 + a2);
        FL(2);
        cx inner = csub(cadd(C(a3, 0.0), cscale(k, y)), cmul(c17, c14));
        c[3] = csub(cadd(cscale(c10, c15), cmul(inner, yc)), cscale(a, c13));
    }
    c[4] = csub(C(-c10, 0.0), cmul(c12, csub(cmul(yc, cadd(c13, C(a, 0.0))), C(1.0, 0.0))));
    c[5] = cmul(yc, c12);
}

/* residual_squared (point_source_magnifier.cpp:187-203) */
static double residual_squared(const geom* g, cx z)
{
    const double z_re = z.re, z_im = z.im;
    const double u1 = z_re - g->a;
    const double v1 = -z_im;
    const double d1 = u1 * u1 + v1 * v1;
    const double u2 = z_re;
    const double v2 = -z_im;
    const double d2 = u2 * u2 + v2 * v2;
    const double r_re = g->y.re - z_re + g->m1 * u1 / d1 + g->m2 * u2 / d2;
    const double r_im = g->y.im - z_im - g->m1 * v1 / d1 - g->m2 * v2 / d2;
    FL(25);
    return r_re * r_re + r_im * r_im;
}

/* jacobian_determinant (point_source_magnifier.cpp:205-223) */
static double jacobian_det(const geom* g, cx z)
{
    const double u1 = z.re - g->a;
    const double v1 = z.im;
    const double d1 = u1 * u1 + v1 * v1;
    const double d1_2 = d1 * d1;
    const double u2 = z.re;
    const double v2 = z.im;
    const double d2 = u2 * u2 + v2 * v2;
    const double d2_2 = d2 * d2;
    const double der_re = g->m1 * (u1 * u1 - v1 * v1) / d1_2 + g->m2 * (u2 * u2 - v2 * v2) / d2_2;
    const double der_im = -2.0 * g->m1 * u1 * v1 / d1_2 - 2.0 * g->m2 * u2 * v2 / d2_2;
    FL(34);
    return 1.0 - (der_re * der_re + der_im * der_im);
}

/* select_physical_roots (point_source_magnifier.cpp:225-279) */
static void select_physical(const geom* g, const cx* roots, int* physical)
{
    double res[5];
    int w1 = 0, w2 = 0, w3 = 0;
    for (int i = 0; i < 5; i++) {
        res[i] = residual_squared(g, roots[i]);
        if (i == 0) {
            w1 = i;
        } else if (i == 1) {
            if (res[i] > res[w1]) {
                w2 = w1;
                w1 = i;
            } else {
                w2 = i;
            }
        } else if (i == 2) {
            if (res[i] > res[w1]) {
                w3 = w2;
                w2 = w1;
                w1 = i;
            } else if (res[i] > res[w2]) {
                w3 = w2;
                w2 = i;
            } else {
                w3 = i;
            }
        } else if (res[i] > res[w1]) {
            w3 = w2;
            w2 = w1;
            w1 = i;
        } else if (res[i] > res[w2]) {
            w3 = w2;
            w2 = i;
        } else if (res[i] > res[w3]) {
            w3 = i;
        }
    }
    FL(2);
    const int three = res[w2] * 1.0e-8 > res[w3] + 1.0e-24;
    for (int i = 0; i < 5; i++) physical[i] = !(three && (i == w1 || i == w2));
}

static void solve_quintic(const geom* g, cx* roots)
{
    cx c[6];
    quintic_coefficients(g, c);
    sg_roots_gen(roots, c, 5, 1, 0);
}

/* point_source_result_from_roots (point_source_magnifier.cpp:281-297) */
static double ps_from_roots(const geom* g, const cx* roots, int* image_count)
{
    int physical[5];
    select_physical(g, roots, physical);
    double mag = 0.0;
    int n = 0;
    for (int i = 0; i < 5; i++) {
        if (!physical[i]) continue;
        mag += 1.0 / fabs(jacobian_det(g, roots[i]));
        FL(2);
        ++n;
    }
    if (image_count) *image_count = n;
    return mag;
}

/* binary_mag0 / binary_mag0_impl (point_source_magnifier.cpp:567-573, 652-693) */
double lcbo_binary_mag0(double s, double q, double x, double y, int* image_count)
{
    if (s == 0.0 || q <= 0.0) {
        if (image_count) *image_count = 0;
        return 0.0;
    }
    geom g = make_geom(s, q, x, y);
    cx roots[5];
    solve_quintic(&g, roots);
    return ps_from_roots(&g, roots, image_count);
}

void lcbo_binary_mag0_many(const double* s, const double* q, const double* x,
                           const double* y, long n, double* out, int* image_count)
{
    for (long i = 0; i < n; i++) {
        int k = 0;
        out[i] = lcbo_binary_mag0(s[i], q[i], x[i], y[i], &k);
        if (image_count) image_count[i] = k;
    }
}

/* derivative_error_indicator_from_roots (point_source_magnifier.cpp:313-359) */
static double derivative_indicator(const geom* g, const cx* roots, const int* physical)
{
    double indicator = 0.0;
    const cx a = C(g->a, 0.0), m1 = C(g->m1, 0.0), m2 = C(g->m2, 0.0);
    for (int i = 0; i < 5; i++) {
        if (!physical[i]) continue;
        const cx z = roots[i];
        const cx dza = csub(z, a);
        const cx za2 = cmul(dza, dza);
        const cx zb2 = cmul(z, z);
        const cx j1 = cadd(cdiv(m1, za2), cdiv(m2, zb2));
        const cx j1c = cconj(j1);
        const double det_j = cdsub(1.0, cmul(j1, j1c)).re;
        if (det_j == 0.0 || !isfinite(det_j)) return INFINITY;
        cx j2 = cscale(-2.0, cadd(cdiv(m1, cmul(za2, dza)), cdiv(m2, cmul(zb2, z))));
        cx j3 = cscale(6.0, cadd(cdiv(m1, cmul(za2, za2)), cdiv(m2, cmul(zb2, zb2))));
        const double det_j2 = det_j * det_j;
        const cx j1c2 = cmul(j1c, j1c);
        j3 = cmul(j3, j1c2);
        const double ob2 = (j2.re * j2.re + j2.im * j2.im) * (6.0 - 6.0 * det_j + det_j2);
        j2 = cmul(cmul(cmul(j2, j2), j1c2), j1c);
        const double denominator = fabs(det_j * det_j2 * det_j2);
        FL(10);
        if (denominator == 0.0 || !isfinite(denominator)) return INFINITY;
        const double cq =
            0.5 * (fabs(ob2 - 6.0 * j2.re - 2.0 * j3.re * det_j) + 3.0 * fabs(j2.im)) /
            denominator;
        FL(9);
        if (isfinite(cq))
            indicator += cq;
        else
            return INFINITY;
    }
    return indicator;
}

/* binary_mag0_with_derivatives (point_source_magnifier.cpp:583-614) */
double lcbo_binary_mag0_with_derivatives(double s, double q, double x, double y,
                                         int* image_count, double* indicator)
{
    if (s == 0.0 || q <= 0.0) {
        if (image_count) *image_count = 0;
        if (indicator) *indicator = 0.0;
        return 0.0;
    }
    geom g = make_geom(s, q, x, y);
    cx roots[5];
    solve_quintic(&g, roots);
    int physical[5];
    select_physical(&g, roots, physical);
    double mag = 0.0;
    int n = 0;
    for (int i = 0; i < 5; i++) {
        if (!physical[i]) continue;
        mag += 1.0 / fabs(jacobian_det(&g, roots[i]));
        ++n;
    }
    if (image_count) *image_count = n;
    if (indicator) *indicator = derivative_indicator(&g, roots, physical);
    return mag;
}

/* binary_image_candidates (point_source_magnifier.cpp:739-778) */
int lcbo_binary_image_candidates(double s, double q, double x, double y, double* out)
{
    if (s == 0.0 || q <= 0.0) return 0;
    geom g = make_geom(s, q, x, y);
    cx roots[5];
    solve_quintic(&g, roots);
    double rows[5][5];
    for (int i = 0; i < 5; i++) {
        rows[i][0] = roots[i].re;
        rows[i][1] = roots[i].im;
        rows[i][2] = jacobian_det(&g, roots[i]);
        rows[i][3] = sqrt(residual_squared(&g, roots[i]));
        rows[i][4] = 0.0;
    }
    /* std::sort on 5 elements is an insertion sort: stable for equal keys */
    for (int i = 1; i < 5; i++) {
        double tmp[5];
        memcpy(tmp, rows[i], sizeof(tmp));
        int k = i;
        while (k > 0 && tmp[3] < rows[k - 1][3]) {
            memcpy(rows[k], rows[k - 1], sizeof(tmp));
            --k;
        }
        memcpy(rows[k], tmp, sizeof(tmp));
    }
    int physical_count = 5;
    if (rows[3][3] * 1.0e-4 > rows[2][3] + 1.0e-12) physical_count = 3;
    for (int i = 0; i < 5; i++) {
        rows[i][4] = i < physical_count ? 1.0 : 0.0;
        memcpy(out + 5 * i, rows[i], sizeof(rows[i]));
    }
    return 5;
}

/* static VBM light curve, point source (lcbinint_pybind.cpp:1120-1139) */
void lcbo_light_curve_ps(const double* times, long n, double t0, double tE, double u0,
                         double alpha, double s, double q, double* out)
{
    const double cos_alpha = cos(alpha);
    const double sin_alpha = sin(alpha);
    const double inverse_tE = 1.0 / tE;
    const double effective_q = q != 0.0 ? 1.0 / q : q;
    for (long i = 0; i < n; i++) {
        const double tau = (times[i] - t0) * inverse_tE;
        out[i] = lcbo_binary_mag0(s, effective_q, tau * cos_alpha - u0 * sin_alpha,
                                  tau * sin_alpha + u0 * cos_alpha, NULL);
    }
}

/* ------------------------------------------------------ limb darkening */

/* source_surface_brightness (finite_source_magnifier.cpp:369-375) */
static double surface_brightness(double r2, double c, double d)
{
    double b = r2 < 0.0 ? 0.0 : (r2 > 1.0 ? 1.0 : r2);
    double t = 1.0 - b;
    const double mu = sqrt(t > 0.0 ? t : 0.0);
    return 1.0 - c * (1.0 - mu) - d * (1.0 - sqrt(mu));
}

/* ensure_limb_darkening_table (finite_source_magnifier.cpp:3979-3999) */
void lcbo_ld_table(double c, double d, double* table)
{
    for (int i = 0; i <= 20000; i++)
        table[i] = surface_brightness((double)i / (double)20000, c, d);
}

/* limb_darkening_table_brightness (finite_source_magnifier.cpp:4001-4012) */
double lcbo_ld_lookup(const double* table, double r2)
{
    if (r2 <= 0.0) return table[0];
    if (r2 >= 1.0) return table[20000];
    return table[(int)(r2 * 20000.0 + 0.5)];
}

/* ------------------------------------------------------ hexadecapole */

/* hexadecapole_binary (finite_source_magnifier.cpp:792-847), LD-corrected with
 * Gamma, Lambda of finite_source_magnifier.cpp:349-367. */
double lcbo_hexadecapole(double s, double q, double x, double y, double rho, double c,
                         double d, double a0, double* rel_err)
{
    const double h = 0.70710678118654752440;
    const double sineta[8] = {0.0, 1.0, 0.0, -1.0, h, h, -h, -h};
    const double coseta[8] = {1.0, 0.0, -1.0, 0.0, h, -h, -h, h};
    double mags[12];
    int k = 0;
    for (int i = 0; i < 4; i++) {
        mags[k++] = lcbo_binary_mag0(s, q, x + rho * coseta[i], y + rho * sineta[i], NULL);
        mags[k++] = lcbo_binary_mag0(s, q, x + 0.5 * rho * coseta[i], y + 0.5 * rho * sineta[i],
                                     NULL);
        mags[k++] = lcbo_binary_mag0(s, q, x + rho * coseta[i + 4], y + rho * sineta[i + 4], NULL);
    }
    double a1p = 0.0, a2p = 0.0, a1x = 0.0;
    k = 0;
    for (int i = 0; i < 4; i++) {
        a1p += mags[k++];
        a2p += mags[k++];
        a1x += mags[k++];
    }
    a1p = a1p / 4.0 - a0;
    a2p = a2p / 4.0 - a0;
    a1x = a1x / 4.0 - a0;
    const double a2rho2 = (16.0 * a2p - a1p) / 3.0;
    const double a4rho4 = (a1p + a1x) / 2.0 - a2rho2;
    const double den = 15.0 - 5.0 * c - 3.0 * d;
    const double gamma = den == 0.0 ? 0.0 : 10.0 * c / den;
    const double lambda = den == 0.0 ? 0.0 : 12.0 * d / den;
    const double quad = 0.5 * a2rho2 * (1.0 - 0.2 * gamma - lambda / 9.0);
    const double hex = a4rho4 / 3.0 * (1.0 - 11.0 * gamma / 35.0 - 7.0 * lambda / 39.0);
    const double mag = a0 + quad + hex;
    const double am = fabs(mag);
    if (rel_err) *rel_err = fabs(hex) / (am < 1.0e-10 ? 1.0e-10 : am); /* std::max */
    return mag;
}

/* ------------------------------------------------------ critical curve */

/* libgcc __divdc3 (GCC 12): Smith's method; the power-of-two prescaling of
 * the library is exact and omitted, the subnormal-ratio branch is kept. */
static cx std_cdiv(double a, double b, double c, double d)
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
    return C(x, y);
}

void lcbo_critical_curve(double s, double q, double cos_phi, double sin_phi, double* caustic,
                         double* critical)
{
    const double sa = fabs(s);
    const double q_input = fabs(q);
    const double qq = q_input < 1.0 ? q_input : 1.0 / q_input;
    const double a_re = q_input < 1.0 ? -sa : sa;
    const double m1r = 1.0 / (1.0 + qq);
    const double m2r = qq * m1r;
    /* std::complex products with the zero imaginary parts carried along */
    const cx a = C(a_re, 0.0), m1 = C(m1r, 0.0), m2 = C(m2r, qq * 0.0), ph = C(cos_phi, sin_phi);
    const cx a2 = cmul(a, a);
    cx c[5];
    c[0] = cmul(m2, a2);
    c[1] = cmul(cscale(-2.0, m2), a);
    c[2] = csub(cadd(m1, m2), cmul(ph, a2));
    c[3] = cmul(cscale(2.0, ph), a);
    c[4] = C(-ph.re, -ph.im);
    cx roots[4];
    sg_roots_gen(roots, c, 4, 1, 0);
    for (int r = 0; r < 4; r++) {
        const cx z = roots[r];
        /* binary_lens_equation: z - m1/(conj(z) - a) - m2/conj(z) - a*m1 */
        const cx zc = cconj(z);
        const cx t1 = std_cdiv(m1.re, m1.im, zc.re - a.re, zc.im - a.im);
        const cx t2 = std_cdiv(m2.re, m2.im, zc.re, zc.im);
        const cx sh = C((z.re - t1.re) - t2.re, (z.im - t1.im) - t2.im);
        const cx am1 = cmul(a, m1);
        caustic[2 * r] = sh.re - am1.re;
        caustic[2 * r + 1] = sh.im - am1.im;
        /* map_binary_lens_real */
        const double x = z.re, y = z.im;
        const double xa = x - a_re;
        const double den1 = xa * xa + y * y;
        const double den2 = x * x + y * y;
        critical[2 * r] = x - m1r * xa / den1 - m2r * x / den2 - a_re * m1r;
        critical[2 * r + 1] = y - m1r * y / den1 - m2r * y / den2;
    }
}
