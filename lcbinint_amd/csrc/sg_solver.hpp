// FP64 complex-polynomial root finder for gfx950: the Skowron & Gould (2012)
// three-regime iteration (Laguerre -> second-order general -> Newton) with the
// Adams stopping rule, deflation and a final Newton polish, as the reference
// runs it in cmplx_roots_gen (third_party/skowron_gould/SkowronGould.cpp:79-168).
//
// One GPU thread solves one polynomial.  The degree is a template parameter so
// that every coefficient and root lives in VGPRs: no private-array indexing is
// left after unrolling, hence no scratch traffic.  Operation order follows the
// reference expression by expression; the library is built with
// -ffp-contract=off so no multiply-add is fused, and double division and sqrt
// are the IEEE correctly rounded sequences, so each root is bitwise the
// reference's.
#pragma once

#include <hip/hip_runtime.h>

namespace lcbg {

struct cx {
    double re, im;
};

__device__ __forceinline__ cx mk(double re, double im) { return cx{re, im}; }
__device__ __forceinline__ cx operator+(cx a, cx b) { return cx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cx operator-(cx a, cx b) { return cx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cx operator*(cx a, cx b)
{
    return cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// SkowronGould.h:50-55: both parts divided by |p2|^2 (two IEEE divisions).
__device__ __forceinline__ cx operator/(cx p1, cx p2)
{
    const double md = p2.re * p2.re + p2.im * p2.im;
    return cx{(p1.re * p2.re + p1.im * p2.im) / md, (p1.im * p2.re - p1.re * p2.im) / md};
}
__device__ __forceinline__ cx operator*(double a, cx z) { return cx{a * z.re, a * z.im}; }
__device__ __forceinline__ cx operator+(double a, cx z) { return cx{z.re + a, z.im}; }
__device__ __forceinline__ cx operator-(double a, cx z) { return cx{a - z.re, -z.im}; }
__device__ __forceinline__ cx neg(cx z) { return cx{-z.re, -z.im}; }
__device__ __forceinline__ cx conj(cx z) { return cx{z.re, -z.im}; }
__device__ __forceinline__ bool eq(cx a, cx b) { return a.re == b.re && a.im == b.im; }
__device__ __forceinline__ bool is_zero(cx a) { return a.re == 0.0 && a.im == 0.0; }
__device__ __forceinline__ double abs2(cx z) { return z.re * z.re + z.im * z.im; }
__device__ __forceinline__ double cabs(cx z) { return __builtin_sqrt(z.re * z.re + z.im * z.im); }
// real(conj(a) * b) as the reference's operator chain evaluates it.
__device__ __forceinline__ double re_conj_mul(cx a, cx b)
{
    const cx c = conj(a);
    return c.re * b.re - c.im * b.im;
}
// SkowronGould.h:23-30
__device__ __forceinline__ cx sg_sqrt(cx z)
{
    const double md = __builtin_sqrt(z.re * z.re + z.im * z.im);
    if (md > 0.0) {
        return cx{__builtin_sqrt((md + z.re) / 2.0),
                  __builtin_sqrt((md - z.re) / 2.0) * ((z.im > 0.0) ? 1.0 : -1.0)};
    }
    return cx{0.0, 0.0};
}

// exp(i*2*pi*f_k) for the ten cycle-breaking jump fractions of the solver
// (SkowronGould.cpp:497-500), as glibc's cos/sin round them; both operand orders
// of the reference (f*2*pi and 2*pi*f) give these same doubles.
__device__ __forceinline__ cx jump_direction(int k)
{
    constexpr double c[10] = {-0x1.43a4ec067471ap-1, 0x1.b9f8520125e44p-1, -0x1.d9f09317910bbp-5,
                              -0x1.ffc29786c1b25p-1, 0x1.bde5c12533376p-1, 0x1.4ee0d0497a649p-1,
                              -0x1.61e482927f6dbp-2, -0x1.70abece2c21e4p-1, 0x1.ea9723c588b18p-1,
                              0x1.100ce2673d75fp-7};
    constexpr double s[10] = {-0x1.8cbc14adec55cp-1, -0x1.027835a80b001p-1, 0x1.ff24763655658p-1,
                              -0x1.f5776cee4e213p-6, 0x1.f7445f6e48663p-2, 0x1.834c92e8b2a3ep-1,
                              0x1.e073ae0e62aa1p-1, 0x1.63482dd7978b3p-1, 0x1.2504ed506e352p-2,
                              -0x1.fffb7b8d5ef39p-1};
    double cr = c[0], si = s[0];
#pragma unroll
    for (int j = 1; j < 10; ++j) {
        if (k == j) {
            cr = c[j];
            si = s[j];
        }
    }
    return cx{cr, si};
}

__device__ __forceinline__ double jump_fraction(int k)
{
    constexpr double f[10] = {0.64109297, 0.91577881, 0.25921289, 0.50487203, 0.08177045,
                              0.13653241, 0.306162,   0.37794326, 0.04618805, 0.75132137};
    double v = f[0];
#pragma unroll
    for (int j = 1; j < 10; ++j) {
        if (k == j) v = f[j];
    }
    return v;
}

constexpr int kMaxIt = 80;  // MAXIT = MT*MR (SkowronGould.h:5-7)
constexpr double kFracErr = 2.0e-15;

// Horner evaluation of p, p' and p''/2 plus the Adams bound (SkowronGould.cpp:883-895).
template <int N>
__device__ __forceinline__ void horner2(const cx* poly, cx root, cx& p, cx& dp, cx& d2p_half,
                                        double& ek, bool with_ek)
{
    p = poly[N];
    dp = cx{0.0, 0.0};
    d2p_half = cx{0.0, 0.0};
    const double absroot = with_ek ? cabs(root) : 0.0;
    if (with_ek) ek = cabs(poly[N]);
#pragma unroll
    for (int k = N; k >= 1; --k) {
        d2p_half = dp + d2p_half * root;
        dp = p + dp * root;
        p = poly[k - 1] + p * root;
        if (with_ek) ek = absroot * ek + cabs(p);
    }
}

template <int N>
__device__ __forceinline__ void horner1(const cx* poly, cx root, cx& p, cx& dp, double& ek,
                                        bool with_ek)
{
    p = poly[N];
    dp = cx{0.0, 0.0};
    const double absroot = with_ek ? cabs(root) : 0.0;
    if (with_ek) ek = cabs(poly[N]);
#pragma unroll
    for (int k = N; k >= 1; --k) {
        dp = p + dp * root;
        p = poly[k - 1] + p * root;
        if (with_ek) ek = absroot * ek + cabs(p);
    }
}

// cmplx_laguerre2newton with starting mode 2 (SkowronGould.cpp:764-1130).
template <int N>
__device__ bool laguerre2newton(const cx* poly, cx& root)
{
    const cx zero{0.0, 0.0};
    const cx c_one{1.0, 0.0};
    int iter = 0;
    double stopping_crit2 = 0.0;
    int j = 1;
    bool good_to_go = false;
    int mode = 2;
    cx p, dp, d2p_half, dx, newroot, fac_newton{0.0, 0.0};
    double ek = 0.0;
    for (;;) {
        if (mode >= 2) {
            const double one_nth = 1.0 / (N);
            const double n_1_nth = (N - 1) * one_nth;
            const double two_n_div_n_1 = 2.0 / n_1_nth;
            const cx c_one_nth{one_nth, 0.0};
            int i;
            for (i = 1; i <= kMaxIt; ++i) {
                horner2<N>(poly, root, p, dp, d2p_half, ek, true);
                const double abs2p = re_conj_mul(p, p);
                iter = iter + 1;
                if (abs2p == 0) return true;
                stopping_crit2 = (kFracErr * ek) * (kFracErr * ek);
                if (abs2p < stopping_crit2) {
                    if (abs2p < 0.01 * stopping_crit2) return true;
                    good_to_go = true;
                } else {
                    good_to_go = false;
                }
                cx denom = zero;
                if (!is_zero(dp)) {
                    fac_newton = p / dp;
                    const cx fac_extra = d2p_half / dp;
                    const cx F_half = fac_newton * fac_extra;
                    const double abs2_F_half = re_conj_mul(F_half, F_half);
                    if (abs2_F_half <= 0.0625) {
                        mode = (abs2_F_half <= 0.000625) ? 0 : 1;
                    }
                    const cx denom_sqrt = sg_sqrt(c_one - two_n_div_n_1 * F_half);
                    if (denom_sqrt.re > 0.0)
                        denom = c_one_nth + n_1_nth * denom_sqrt;
                    else
                        denom = c_one_nth - n_1_nth * denom_sqrt;
                }
                if (is_zero(denom)) {
                    // the reference adds (rather than scales by) the jump here
                    dx = (cabs(root) + 1.0) + jump_direction(i % 10);
                } else {
                    dx = fac_newton / denom;
                }
                newroot = root - dx;
                if (eq(newroot, root)) return true;
                if (good_to_go) {
                    root = newroot;
                    return true;
                }
                if (mode != 2) {
                    root = newroot;
                    j = i + 1;
                    break;
                }
                if ((i % 10) == 0) {
                    const double faq = jump_fraction((i / 10 - 1) % 10);
                    newroot = root - faq * dx;
                }
                root = newroot;
            }
            if (i >= kMaxIt) return false;
        }
        if (mode == 1) {
            int i;
            for (i = j; i <= kMaxIt; ++i) {
                const bool crit = ((i - j) % 10 == 0);
                horner2<N>(poly, root, p, dp, d2p_half, ek, crit);
                if (crit) stopping_crit2 = (kFracErr * ek) * (kFracErr * ek);
                const double abs2p = re_conj_mul(p, p);
                iter = iter + 1;
                if (abs2p == 0.0) return true;
                if (abs2p < stopping_crit2) {
                    if (is_zero(dp)) return true;
                    if (abs2p < 0.01 * stopping_crit2) return true;
                    good_to_go = true;
                } else {
                    good_to_go = false;
                }
                if (is_zero(dp)) {
                    dx = (cabs(root) + 1.0) * jump_direction(i % 10);
                } else {
                    fac_newton = p / dp;
                    const cx fac_extra = d2p_half / dp;
                    const cx F_half = fac_newton * fac_extra;
                    const double abs2_F_half = re_conj_mul(F_half, F_half);
                    if (abs2_F_half <= 0.000625) mode = 0;
                    dx = fac_newton * (c_one + F_half);
                }
                newroot = root - dx;
                if (eq(newroot, root)) return true;
                if (good_to_go) {
                    root = newroot;
                    return true;
                }
                if (mode != 1) {
                    root = newroot;
                    j = i + 1;
                    break;
                }
                if ((i % 10) == 0) {
                    const double faq = jump_fraction((i / 10 - 1) % 10);
                    newroot = root - faq * dx;
                }
                root = newroot;
            }
            if (i >= kMaxIt) return false;
        }
        if (mode == 0) {
            for (int i = j; i <= j + 10; ++i) {
                const bool crit = (i == j);
                horner1<N>(poly, root, p, dp, ek, crit);
                if (crit) stopping_crit2 = (kFracErr * ek) * (kFracErr * ek);
                const double abs2p = re_conj_mul(p, p);
                iter = iter + 1;
                if (abs2p == 0.0) return true;
                if (abs2p < stopping_crit2) {
                    if (is_zero(dp)) return true;
                    if (abs2p < 0.01 * stopping_crit2) return true;
                    good_to_go = true;
                } else {
                    good_to_go = false;
                }
                if (is_zero(dp)) {
                    dx = (cabs(root) + 1.0) * jump_direction(i % 10);
                } else {
                    dx = p / dp;
                }
                newroot = root - dx;
                if (eq(newroot, root)) return true;
                if (good_to_go) {
                    root = newroot;
                    return true;
                }
                root = newroot;
            }
            if (iter >= kMaxIt) return false;
            mode = 2;
        }
    }
}

// cmplx_laguerre, the fallback when laguerre2newton gives up
// (SkowronGould.cpp:472-624).  The leading coefficient of the lens polynomials
// is never zero, so the reference's degree-reduction preamble is a no-op here.
template <int N>
__device__ void laguerre(const cx* poly, cx& root)
{
    const cx zero{0.0, 0.0};
    const cx c_one{1.0, 0.0};
    const double one_nth = 1.0 / N;
    const double n_1_nth = (N - 1.0) * one_nth;
    const double two_n_div_n_1 = 2.0 / n_1_nth;
    const cx c_one_nth{one_nth, 0.0};
    bool good_to_go = false;
    cx fac_newton{0.0, 0.0};
    for (int i = 1; i <= kMaxIt; ++i) {
        cx p, dp, d2p_half;
        double ek;
        horner2<N>(poly, root, p, dp, d2p_half, ek, true);
        const double absroot = cabs(root);
        const double abs2p = re_conj_mul(p, p);
        if (abs2p == 0) return;
        const double stopping_crit2 = (kFracErr * ek) * (kFracErr * ek);
        if (abs2p < stopping_crit2) {
            if (abs2p < 0.01 * stopping_crit2) return;
            good_to_go = true;
        } else {
            good_to_go = false;
        }
        cx denom = zero;
        if (!is_zero(dp)) {
            fac_newton = p / dp;
            const cx fac_extra = d2p_half / dp;
            const cx F_half = fac_newton * fac_extra;
            const cx denom_sqrt = sg_sqrt(c_one - two_n_div_n_1 * F_half);
            if (denom_sqrt.re >= 0.0)
                denom = c_one_nth + n_1_nth * denom_sqrt;
            else
                denom = c_one_nth - n_1_nth * denom_sqrt;
        }
        cx dx;
        if (is_zero(denom))
            dx = (absroot + 1.0) * jump_direction(i % 10);
        else
            dx = fac_newton / denom;
        cx newroot = root - dx;
        if (eq(newroot, root)) return;
        if (good_to_go) {
            root = newroot;
            return;
        }
        if (i % 10 == 0) {
            const double faq = jump_fraction((i / 10 - 1) % 10);
            newroot = root - faq * dx;
        }
        root = newroot;
    }
}

// cmplx_newton_spec (SkowronGould.cpp:626-762): polish one root on the full polynomial.
template <int N>
__device__ void newton_spec(const cx* poly, cx& root)
{
    bool good_to_go = false;
    double stopping_crit2 = 0.0;
    for (int i = 1; i <= kMaxIt; ++i) {
        cx p, dp;
        double ek = 0.0;
        const bool crit = (i % 10 == 1);
        horner1<N>(poly, root, p, dp, ek, crit);
        if (crit) stopping_crit2 = (kFracErr * ek) * (kFracErr * ek);
        const double abs2p = re_conj_mul(p, p);
        if (abs2p == 0.0) return;
        if (abs2p < stopping_crit2) {
            if (is_zero(dp)) return;
            if (abs2p < 0.01 * stopping_crit2) return;
            good_to_go = true;
        } else {
            good_to_go = false;
        }
        cx dx;
        if (is_zero(dp))
            dx = (cabs(root) + 1.0) * jump_direction(i % 10);
        else
            dx = p / dp;
        cx newroot = root - dx;
        if (eq(newroot, root)) return;
        if (good_to_go) {
            root = newroot;
            return;
        }
        if (i % 10 == 0) {
            const double faq = jump_fraction((i / 10 - 1) % 10);
            newroot = root - faq * dx;
        }
        root = newroot;
    }
}

// solve_quadratic_eq (SkowronGould.cpp:1132-1154)
__device__ __forceinline__ void solve_quadratic(cx& x0, cx& x1, const cx* poly)
{
    const cx a = poly[2], b = poly[1], c = poly[0];
    const cx b2 = b * b;
    const cx four_a{4 * a.re, 4 * a.im};
    const cx delta = sg_sqrt(b2 - four_a * c);
    if (re_conj_mul(b, delta) >= 0)
        x0 = -0.5 * (b + delta);
    else
        x0 = -0.5 * (b - delta);
    if (is_zero(x0)) {
        x1 = cx{0.0, 0.0};
    } else {
        x1 = c / x0;
        x0 = x0 / a;
    }
}

// One deflation step of cmplx_roots_gen (SkowronGould.cpp:123-152) for the
// current degree n: pick the smallest-modulus remaining guess, find the root
// and divide it out of poly2.
template <int DEG, int n>
__device__ __forceinline__ void deflate_step(cx* roots, cx* poly2)
{
    int ismallest = n - 1;
    double abssmall = abs2(roots[n - 1]);
#pragma unroll
    for (int ii = 0; ii < n - 1; ++ii) {
        const double v = abs2(roots[ii]);
        if (v < abssmall) {
            ismallest = ii;
            abssmall = v;
        }
    }
    if (ismallest != n - 1) {
        cx coef = roots[n - 1];
#pragma unroll
        for (int ii = 0; ii < n - 1; ++ii) {
            if (ii == ismallest) {
                const cx t = roots[ii];
                roots[ii] = coef;
                coef = t;
            }
        }
        roots[n - 1] = coef;
    }
    cx r = roots[n - 1];
    if (!laguerre2newton<n>(poly2, r)) {
        r = cx{0.0, 0.0};
        laguerre<n>(poly2, r);
    }
    roots[n - 1] = r;
    cx coef = poly2[n];
#pragma unroll
    for (int i = n - 1; i >= 0; --i) {
        const cx prev = poly2[i];
        poly2[i] = coef;
        coef = prev + r * coef;
    }
}

template <int DEG, int n>
struct Deflate {
    __device__ __forceinline__ static void run(cx* roots, cx* poly2)
    {
        deflate_step<DEG, n>(roots, poly2);
        Deflate<DEG, n - 1>::run(roots, poly2);
    }
};
template <int DEG>
struct Deflate<DEG, 2> {
    __device__ __forceinline__ static void run(cx*, cx*) {}
};

// cmplx_roots_gen(roots, poly, DEG, polish=true, use_roots_as_starting_points)
template <int DEG>
__device__ __forceinline__ void roots_gen(cx* roots, const cx* poly, bool use_start)
{
    cx poly2[DEG + 1];
    if (!use_start) {
#pragma unroll
        for (int k = 0; k < DEG; ++k) roots[k] = cx{0.0, 0.0};
    }
#pragma unroll
    for (int k = 0; k <= DEG; ++k) poly2[k] = poly[k];
    Deflate<DEG, DEG>::run(roots, poly2);
    solve_quadratic(roots[1], roots[0], poly2);
#pragma unroll
    for (int k = 0; k < DEG - 1; ++k) newton_spec<DEG>(poly, roots[k]);
}

}  // namespace lcbg
