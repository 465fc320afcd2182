// Point-source magnification of the triple lens, one thread per epoch:
// PointSourceMagnifier::triple_mag0 (src/lcbinint/magnification/
// point_source_magnifier.cpp:861-882) on the trajectory of LensModel's static
// branch (src/lcbinint/model/lens_model.cpp:90-99, 126-140).
//
// Per epoch: the degree-10 lens polynomial built by the reference's sequence
// of complex polynomial products and sums (triple_polynomial_coefficients,
// :443-497), its roots by the Skowron-Gould solver at the polynomial's
// effective degree (math::PolynomialRootSolver, src/lcbinint/math/
// polynomial_roots.cpp:35-74), each root polished by 2D Newton on the real
// lens equation (:499-535), duplicates dropped, residuals and Jacobian
// determinants (:537-563), the candidates sorted by residual (a stable
// insertion sort, as std::sort runs on ten elements) and the physical ones
// selected by the absolute and relative residual rules (:929-994).  Complex
// products are the plain formula and quotients libgcc's, as std::complex
// evaluates them for finite values.
#include <cmath>
#include <vector>

#include "fs_math.hpp"
#include "lcbinint_amd.h"
#include "runtime.hpp"

namespace lcbg {

namespace {

// Per-curve record: trajectory and the geometry of
// make_triple_lens_geometry_vbm (src/lcbinint/model/triple_lens_geometry.cpp:36-50),
// cos/sin from the host libm.
struct TripleRec {
    double t0, tE, u0, cos_a, sin_a;
    double lx[3], ly[3], m[3];
};

constexpr int kTriDeg = 10;

__device__ __forceinline__ cx cneg(cx a) { return cx{-a.re, -a.im}; }

// out = lhs * rhs (coefficients constant first), as multiply_polynomial.
template <int NL, int NR>
__device__ __forceinline__ void pmul(const cx* lhs, const cx* rhs, cx* out)
{
#pragma unroll
    for (int k = 0; k < NL + NR - 1; ++k) out[k] = cx{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < NL; ++i)
#pragma unroll
        for (int j = 0; j < NR; ++j) out[i + j] = out[i + j] + lhs[i] * rhs[j];
}

__device__ void triple_coefficients(const TripleRec& R, double sx, double sy, cx* c)
{
    const cx y{sx, sy};
    const cx yc{sx, -sy};
    cx lens[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) lens[i] = cx{R.lx[i], R.ly[i]};
    // p = (z - l0)(z - l1)(z - l2), from {1}
    cx p[4];
    {
        cx a[2] = {cneg(lens[0]), cx{1.0, 0.0}};
        const cx one[1] = {cx{1.0, 0.0}};
        cx t1[2];
        pmul<1, 2>(one, a, t1);
        cx b[2] = {cneg(lens[1]), cx{1.0, 0.0}};
        cx t2[3];
        pmul<2, 2>(t1, b, t2);
        cx d[2] = {cneg(lens[2]), cx{1.0, 0.0}};
        pmul<3, 2>(t2, d, p);
    }
    // conjugate_denominators[i] = p (yc - conj(l_i)) + sum_j m_j prod_{k != j} (z - l_k)
    cx cd[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const cx s{yc.re - R.lx[i], yc.im - (-R.ly[i])};
#pragma unroll
        for (int k = 0; k < 4; ++k) cd[i][k] = p[k] * s;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // product_without_lens(j): from {1}, the two other factors in order
            cx w[3];
            {
                const cx one[1] = {cx{1.0, 0.0}};
                cx cur[3];
                int len = 1;
                cur[0] = one[0];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (k == j) continue;
                    const cx f[2] = {cneg(lens[k]), cx{1.0, 0.0}};
                    cx nxt[3];
                    if (len == 1) {
                        pmul<1, 2>(cur, f, nxt);
                    } else {
                        pmul<2, 2>(cur, f, nxt);
                    }
                    len += 1;
#pragma unroll
                    for (int t = 0; t < 3; ++t) cur[t] = nxt[t];
                }
#pragma unroll
                for (int t = 0; t < 3; ++t) w[t] = cur[t];
            }
            const cx mj{R.m[j], 0.0};
            // add_polynomial(cd, scale(w, m_j)): sizes 4 and 3
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                cx acc = cx{0.0, 0.0} + cd[i][k];
                if (k < 3) acc = acc + w[k] * mj;
                cd[i][k] = acc;
            }
        }
    }
    // all = cd0 cd1 cd2 (from {1})
    cx all[10];
    {
        const cx one[1] = {cx{1.0, 0.0}};
        cx a1[4];
        pmul<1, 4>(one, cd[0], a1);
        cx a2[7];
        pmul<4, 4>(a1, cd[1], a2);
        pmul<7, 4>(a2, cd[2], all);
    }
    // coefficients = (z - y) all
    const cx zy[2] = {cneg(y), cx{1.0, 0.0}};
    pmul<2, 10>(zy, all, c);
    // deflection_sum = sum_i m_i p prod_{j != i} cd_j, from {0}
    cx def[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) def[k] = cx{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j0 = i == 0 ? 1 : 0;
        const int j1 = i == 2 ? 1 : 2;
        cx t1[7];
        pmul<4, 4>(p, cd[j0], t1);
        cx term[10];
        pmul<7, 4>(t1, cd[j1], term);
        const cx mi{R.m[i], 0.0};
        // add_polynomial(def, scale(term, m_i)); the first sum starts from {0}
#pragma unroll
        for (int k = 0; k < 10; ++k) {
            cx acc = cx{0.0, 0.0} + def[k];
            acc = acc + term[k] * mi;
            def[k] = acc;
        }
    }
    // coefficients - deflection_sum
#pragma unroll
    for (int k = 0; k < 11; ++k) {
        cx acc = cx{0.0, 0.0} + c[k];
        if (k < 10) acc = acc - def[k];
        c[k] = acc;
    }
}

// polish_triple_image_root (:499-535)
__device__ __forceinline__ void polish_root(const TripleRec& R, double sx, double sy, double& zx, double& zy)
{
    for (int iter = 0; iter < 60; ++iter) {
        double ax = 0.0, ay = 0.0, dxx = 0.0, dxy = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double dx = zx - R.lx[i];
            const double dy = zy - R.ly[i];
            const double d2 = dx * dx + dy * dy;
            const double d4 = d2 * d2;
            ax += R.m[i] * dx / d2;
            ay += R.m[i] * dy / d2;
            dxx += R.m[i] * (dy * dy - dx * dx) / d4;
            dxy += R.m[i] * 2.0 * dx * dy / d4;
        }
        const double fx = zx - ax - sx;
        const double fy = zy - ay - sy;
        if (fabs(fx) + fabs(fy) < 1.0e-11) break;
        const double j00 = 1.0 - dxx;
        const double j01 = dxy;
        const double j11 = 1.0 + dxx;
        const double det = j00 * j11 - j01 * j01;
        if (fabs(det) < 1.0e-25) break;
        zx -= (j11 * fx - j01 * fy) / det;
        zy -= (j00 * fy - j01 * fx) / det;
    }
}

// triple_residual (:537-548) through triple_lens_equation
// (triple_lens_geometry.cpp:52-65)
__device__ __forceinline__ double triple_residual(const TripleRec& R, double sx, double sy, double zx, double zy)
{
    cx src{zx, zy};
    const cx zc{zx, -zy};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const cx d{zc.re - R.lx[i], zc.im - (-R.ly[i])};
        const cx q = std_cdiv(R.m[i], 0.0, d.re, d.im);
        src = cx{src.re - q.re, src.im - q.im};
    }
    const double dx = src.re - sx;
    const double dy = src.im - sy;
    return __builtin_sqrt(dx * dx + dy * dy);
}

// triple_jacobian_determinant (:550-563)
__device__ __forceinline__ double triple_jacdet(const TripleRec& R, double zx, double zy)
{
    cx der{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const cx dz{zx - R.lx[i], zy - R.ly[i]};
        const cx d2 = dz * dz;
        const cx q = std_cdiv(R.m[i], 0.0, d2.re, d2.im);
        der = der + q;
    }
    return 1.0 - (der.re * der.re + der.im * der.im);
}

template <int D>
__device__ __forceinline__ void roots_at(cx* roots, const cx* c)
{
    roots_gen<D>(roots, c, false);
}

// triple_mag0: magnification and physical image count; false if the
// polynomial has fewer than three effective degrees (outside this kernel).
__device__ bool triple_mag0(const TripleRec& R, double sx, double sy, double* mag, int* count)
{
    cx c[kTriDeg + 1];
    triple_coefficients(R, sx, sy, c);
    int deg = -1;
    for (int k = kTriDeg; k >= 0; --k) {
        if (!(c[k].re == 0.0 && c[k].im == 0.0)) {
            deg = k;
            break;
        }
    }
    cx roots[kTriDeg];
    for (int k = 0; k < kTriDeg; ++k) roots[k] = cx{0.0, 0.0};
    switch (deg) {
    case 10: roots_at<10>(roots, c); break;
    case 9: roots_at<9>(roots, c); break;
    case 8: roots_at<8>(roots, c); break;
    case 7: roots_at<7>(roots, c); break;
    case 6: roots_at<6>(roots, c); break;
    case 5: roots_at<5>(roots, c); break;
    case 4: roots_at<4>(roots, c); break;
    case 3: roots_at<3>(roots, c); break;
    default: return false;
    }
    double px[kTriDeg], py[kTriDeg], res[kTriDeg], det[kTriDeg];
    int n = 0;
    for (int k = 0; k < deg; ++k) {
        double zx = roots[k].re, zy = roots[k].im;
        if (!isfinite(zx) || !isfinite(zy)) continue;
        polish_root(R, sx, sy, zx, zy);
        bool dup = false;
        for (int e = 0; e < n; ++e) {
            const double ddx = zx - px[e];
            const double ddy = zy - py[e];
            if (ddx * ddx + ddy * ddy < 1.0e-16) {
                dup = true;
                break;
            }
        }
        if (dup) continue;
        const double r = triple_residual(R, sx, sy, zx, zy);
        const double dj = triple_jacdet(R, zx, zy);
        // insertion by residual after every equal one (std::sort's insertion
        // sort on ten elements keeps equal residuals in order)
        int at = n;
        while (at > 0 && r < res[at - 1]) {
            px[at] = px[at - 1];
            py[at] = py[at - 1];
            res[at] = res[at - 1];
            det[at] = det[at - 1];
            --at;
        }
        px[at] = zx;
        py[at] = zy;
        res[at] = r;
        det[at] = dj;
        ++n;
    }
    int physical = 0;
    for (int i = 0; i < n; ++i)
        if (res[i] <= 1.0e-7) ++physical;
    if (physical == 0 && n > 0) physical = 1;
    for (int i = 1; i < n; ++i) {
        if (res[i] * 1.0e-4 > res[i - 1] + 1.0e-7) {
            physical = physical > i ? physical : i;
            break;
        }
    }
    double a = 0.0;
    for (int i = 0; i < physical; ++i) a += 1.0 / fabs(det[i]);
    *mag = a;
    *count = physical;
    return true;
}

__global__ void __launch_bounds__(256) k_triple_ps(const TripleRec* __restrict__ curves, const int64_t* offsets,
                                                   const double* times, double* mag, int32_t* images,
                                                   unsigned* bad)
{
    const int c = blockIdx.x;
    const TripleRec R = curves[c];
    const int64_t begin = offsets[c];
    const int64_t end = offsets[c + 1];
    for (int64_t e = begin + threadIdx.x; e < end; e += blockDim.x) {
        // LensModel's (time - t0) / tE (the static binary loop multiplies by 1 / tE)
        const double tn = (times[e] - R.t0) / R.tE;
        const double sx = tn * R.cos_a - R.u0 * R.sin_a;
        const double sy = tn * R.sin_a + R.u0 * R.cos_a;
        double a = __builtin_nan("");
        int n = 0;
        if (!triple_mag0(R, sx, sy, &a, &n)) atomicOr(bad, 1u);
        mag[e] = a;
        if (images) images[e] = n;
    }
}

TripleRec make_triple_rec(const lcbg_triple& t)
{
    TripleRec r;
    r.t0 = t.t0;
    r.tE = t.tE;
    r.u0 = t.u0;
    r.cos_a = std::cos(t.alpha);
    r.sin_a = std::sin(t.alpha);
    const double total = 1.0 + t.q + t.q2;
    const double z1x = t.q * t.s / (1.0 + t.q);
    const double z2x = -t.s / (1.0 + t.q);
    r.lx[0] = z1x;
    r.ly[0] = 0.0;
    r.lx[1] = z2x;
    r.ly[1] = 0.0;
    r.lx[2] = z1x - t.sep2 * std::cos(t.ang);
    r.ly[2] = -t.sep2 * std::sin(t.ang);
    r.m[0] = 1.0 / total;
    r.m[1] = t.q / total;
    r.m[2] = t.q2 / total;
    return r;
}

}  // namespace

}  // namespace lcbg

using namespace lcbg;

extern "C" lcbg_status lcbg_triple_point_source_device(const lcbg_triple* curves, int32_t n_curves,
                                                       const int64_t* offsets, const double* times, double* mag,
                                                       int32_t* image_count, void* stream)
{
    if (n_curves < 0 || (n_curves > 0 && (!curves || !offsets || !times || !mag))) return LCBG_INVALID_ARGUMENT;
    if (n_curves == 0) return LCBG_OK;
    for (int32_t c = 0; c < n_curves; ++c) {
        const lcbg_triple& t = curves[c];
        if (!std::isfinite(t.tE) || t.tE == 0.0 || !(t.q2 > 0.0)) return LCBG_INVALID_ARGUMENT;
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    std::vector<TripleRec> recs(static_cast<std::size_t>(n_curves));
    for (int32_t c = 0; c < n_curves; ++c) recs[static_cast<std::size_t>(c)] = make_triple_rec(curves[c]);
    void* mem = nullptr;
    const std::size_t bytes = recs.size() * sizeof(TripleRec);
    lcbg_status s = map_hip(hipMallocAsync(&mem, bytes + 256, st));
    if (s != LCBG_OK) return s;
    unsigned* bad = reinterpret_cast<unsigned*>(static_cast<char*>(mem) + bytes);
    s = map_hip(hipMemcpyAsync(mem, recs.data(), bytes, hipMemcpyHostToDevice, st));
    if (s == LCBG_OK) s = map_hip(hipMemsetAsync(bad, 0, sizeof(unsigned), st));
    if (s == LCBG_OK) {
        hipLaunchKernelGGL(k_triple_ps, dim3(static_cast<unsigned>(n_curves)), dim3(256), 0, st,
                           static_cast<const TripleRec*>(mem), offsets, times, mag, image_count, bad);
        s = check_launch();
    }
    unsigned h_bad = 0;
    if (s == LCBG_OK) s = map_hip(hipMemcpyAsync(&h_bad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    if (s == LCBG_OK) s = map_hip(hipStreamSynchronize(st));
    (void)hipFreeAsync(mem, st);
    if (s != LCBG_OK) return s;
    return h_bad ? LCBG_UNSUPPORTED : LCBG_OK;
}

extern "C" lcbg_status lcbg_triple_light_curve_point_source(const lcbg_triple* curve, const double* times,
                                                            int64_t n, double* mag, int32_t* image_count)
{
    if (!curve || n < 0 || (n > 0 && (!times || !mag))) return LCBG_INVALID_ARGUMENT;
    if (n == 0) return LCBG_OK;
    void* mem = nullptr;
    const std::size_t tb = static_cast<std::size_t>(n) * sizeof(double);
    const std::size_t ib = static_cast<std::size_t>(n) * sizeof(int32_t);
    lcbg_status s = map_hip(hipMalloc(&mem, 2 * tb + ib + 1024));
    if (s != LCBG_OK) return s;
    char* p = static_cast<char*>(mem);
    double* d_t = reinterpret_cast<double*>(p);
    double* d_m = reinterpret_cast<double*>(p + tb);
    int64_t* d_off = reinterpret_cast<int64_t*>(p + 2 * tb);
    int32_t* d_i = reinterpret_cast<int32_t*>(p + 2 * tb + 256);
    const int64_t off[2] = {0, n};
    s = map_hip(hipMemcpy(d_t, times, tb, hipMemcpyHostToDevice));
    if (s == LCBG_OK) s = map_hip(hipMemcpy(d_off, off, sizeof(off), hipMemcpyHostToDevice));
    if (s == LCBG_OK) s = lcbg_triple_point_source_device(curve, 1, d_off, d_t, d_m, d_i, nullptr);
    if (s == LCBG_OK || s == LCBG_UNSUPPORTED) {
        const lcbg_status c1 = map_hip(hipMemcpy(mag, d_m, tb, hipMemcpyDeviceToHost));
        lcbg_status c2 = LCBG_OK;
        if (image_count) c2 = map_hip(hipMemcpy(image_count, d_i, ib, hipMemcpyDeviceToHost));
        if (s == LCBG_OK) s = c1 != LCBG_OK ? c1 : c2;
    }
    (void)hipFree(mem);
    if (s == LCBG_OK)
        for (int64_t i = 0; i < n; ++i)
            if (!std::isfinite(mag[i])) return LCBG_NUMERICAL_ERROR;
    return s;
}
