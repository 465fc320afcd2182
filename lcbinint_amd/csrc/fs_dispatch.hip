// Finite-source dispatcher, FiniteSourceMagnifier::binary_mag
// (finite_source_magnifier.cpp:4633-4843), split into two kernels.
//
// Stage 1, one 256-thread workgroup per curve: source position, point-source
// solve and the caustic bounding-box exit (:4676-4698); every other epoch is
// appended to the stage-2 queue.  Most epochs of a light curve end here.
//
// Stage 2, resident 64-lane waves pulling epochs from the queue: the sampled
// and segment caustic distances (:4183-4246, :4107-4181) as lane-parallel
// minima over the curve's caustic points, the point-source derivative test
// (:4719-4763) and the hexadecapole test (:4765-4803, twelve solves on twelve
// lanes).  Epochs that fail both are appended to the inverse-ray queue with
// their method (Cartesian or polar, :4805-4842).
#include "fs_launch.hpp"
#include "fs_math.hpp"
#include "fs_wave.hpp"

namespace lcbg {

constexpr int kStage1Block = 256;

__global__ void __launch_bounds__(kStage1Block) k_fs_stage1(const CurveRec* __restrict__ curves,
                                                            FsCache cache, FsOptions opt,
                                                            FsEpochs ep)
{
    const int c = blockIdx.x;
    const CurveRec cv = curves[c];
    const int64_t begin = ep.offsets[c];
    const int64_t end = ep.offsets[c + 1];
    const CausticScalars sc = cache.sc[c];
    for (int64_t e = begin + threadIdx.x; e < end; e += kStage1Block) {
        const double tau = (ep.times[e] - cv.t0) * cv.inverse_tE;
        const Pt src{tau * cv.cos_alpha - cv.u0 * cv.sin_alpha,
                     tau * cv.sin_alpha + cv.u0 * cv.cos_alpha};
        ep.curve_of[e] = c;
        ep.src[e] = src;
        if (cv.s == 0.0 || cv.q_eff <= 0.0) {
            ep.mag[e] = 0.0;
            ep.method[e] = LCBG_METHOD_POINT_SOURCE;
            ep.a_ps[e] = 0.0;
            continue;
        }
        const Geom g = make_geom(cv.s, cv.q_eff, src.x, src.y);
        cx roots[5];
        solve_quintic(g, roots);
        const unsigned mask = select_physical(g, roots);
        // LensModel path with rho > 0 (lens_model.cpp:199-212): the sum over
        // binary_images in their residual order; rho == 0 keeps binary_mag0
        const double a_ps = (opt.ps_images && cv.rho != 0.0) ? ps_from_images(g, roots)
                                                              : ps_from_roots(g, roots, mask).mag;
        ep.a_ps[e] = a_ps;
        double* rr = ep.roots + 10 * e;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            rr[2 * k] = roots[k].re;
            rr[2 * k + 1] = roots[k].im;
        }
        // rho == 0: binary_mag0 only (python/lcbinint_pybind.cpp:1129-1139);
        // finite_mode <= 0: point-source magnification (:4671-4674)
        const double rho = cv.rho;
        if (!(rho > 0.0) || opt.mode <= 0 || sc.n == 0) {
            ep.mag[e] = a_ps;
            ep.method[e] = LCBG_METHOD_POINT_SOURCE;
            continue;
        }
        const double margin = opt.kinji * rho;
        if (src.x < sc.min_x - margin || src.x > sc.max_x + margin || src.y < sc.min_y - margin ||
            src.y > sc.max_y + margin) {
            ep.mag[e] = a_ps;
            ep.method[e] = LCBG_METHOD_POINT_SOURCE;
            continue;
        }
        const unsigned slot = atomicAdd(&ep.counters[0], 1u);
        ep.near_list[slot] = e;
    }
}

__device__ __forceinline__ int cell_coord(double v, double lo, double step)
{
    return iclamp(static_cast<int>((v - lo) / step), 0, kGrid - 1);
}

__global__ void __launch_bounds__(kWave) k_fs_stage2(const CurveRec* __restrict__ curves,
                                                     FsCache cache, FsOptions opt, FsEpochs ep)
{
    const int lane = lane_id();
    const int count = uni(static_cast<int>(ep.counters[0]));
    for (;;) {
        const int item = next_item(&ep.counters[1]);
        if (item >= count) break;
        const int64_t e = ep.near_list[item];
        const int c = ep.curve_of[e];
        const CurveRec cv = curves[c];
        const CausticScalars sc = cache.sc[c];
        const int bins = cache.bins;
        const Pt* pts = cache.pts + static_cast<size_t>(c) * bins * 4;
        const Pt* br = cache.br + static_cast<size_t>(c) * bins * 4;
        const int* pt_cell = cache.pt_cell + static_cast<size_t>(c) * bins * 4;
        const int* br_cell = cache.br_cell + static_cast<size_t>(c) * bins * 4;
        const Pt src = ep.src[e];
        const double rho = cv.rho;
        const double a_ps = ep.a_ps[e];

        // binary_sampled_caustic_distance (:4183-4246) with the bbox margin
        const double margin = opt.kinji * rho;
        const int ix0 = cell_coord(src.x - margin, sc.min_x, sc.step_x);
        const int ix1 = cell_coord(src.x + margin, sc.min_x, sc.step_x);
        const int iy0 = cell_coord(src.y - margin, sc.min_y, sc.step_y);
        const int iy1 = cell_coord(src.y + margin, sc.min_y, sc.step_y);
        double d2 = __builtin_inf();
        double d2all = __builtin_inf();
        for (int k = lane; k < 4 * bins; k += kWave) {
            const int cell = pt_cell[k];
            const int iy = cell / kGrid;
            const int ix = cell - iy * kGrid;
            const Pt p = pts[k];
            const double dx = src.x - p.x;
            const double dy = src.y - p.y;
            const double cand = dx * dx + dy * dy;
            if (ix >= ix0 && ix <= ix1 && iy >= iy0 && iy <= iy1 && cand < d2) d2 = cand;
            if (cand < d2all) d2all = cand;
        }
        d2 = uni(wave_min(d2));
        const double sampled = (d2 == __builtin_inf()) ? d2 : __builtin_sqrt(d2);

        // binary_caustic_distance (:4107-4181): segments next to branch points
        // in the cells within nearest + max_seg_len
        double nearest = sampled;
        if (!isfinite(nearest)) nearest = uni(__builtin_sqrt(wave_min(d2all)));
        const double seg_radius = nearest + sc.max_seg_len;
        const int jx0 = cell_coord(src.x - seg_radius, sc.min_x, sc.step_x);
        const int jx1 = cell_coord(src.x + seg_radius, sc.min_x, sc.step_x);
        const int jy0 = cell_coord(src.y - seg_radius, sc.min_y, sc.step_y);
        const int jy1 = cell_coord(src.y + seg_radius, sc.min_y, sc.step_y);
        double dseg = nearest;
        const int n = sc.n;
        if (n >= 2) {
            for (int k = lane; k < 4 * bins; k += kWave) {
                const int cell = br_cell[k];
                const int iy = cell / kGrid;
                const int ix = cell - iy * kGrid;
                if (ix < jx0 || ix > jx1 || iy < jy0 || iy > jy1) continue;
                const int b = k / bins;
                const int pos = k - b * bins;
                const int prev = pos > 0 ? pos - 1 : n - 1;
                const int next = (pos + 1) % n;
                const Pt* bb = br + static_cast<size_t>(b) * bins;
                dseg = dmin(dseg, point_segment_distance(src, bb[prev], bb[pos]));
                dseg = dmin(dseg, point_segment_distance(src, bb[pos], bb[next]));
            }
        }
        const double refined = uni(wave_min(dseg));
        const bool near_caustic = refined < opt.hex_threshold * rho;

        int method = -1;
        double mag = 0.0;
        if (!near_caustic) {
            double safety = 1.0;
            if (rho >= 1.0e-3 && isfinite(refined) && rho > 0.0) {
                const double dist_ratio = refined / rho;
                const double t = opt.hex_threshold / dmax(dist_ratio, opt.hex_threshold);
                safety = dmax(1.0, 30.0 * t * t * t);
            }
            const double requested = opt.adaptive_hex > 0.0 ? opt.adaptive_hex : 1.0e-3;
            const Geom g = make_geom(cv.s, cv.q_eff, src.x, src.y);
            cx roots[5];
            const double* rr = ep.roots + 10 * e;
#pragma unroll
            for (int k = 0; k < 5; ++k) roots[k] = cx{rr[2 * k], rr[2 * k + 1]};
            if (opt.hex_threshold > 0.0) {
                // binary_mag0_with_derivatives on the same quintic as the
                // point-source solve: identical roots
                const unsigned mask = select_physical(g, roots);
                const double a_d = ps_from_roots(g, roots, mask).mag;
                const double ind = derivative_indicator(g, roots, mask);
                const double dre = ind * rho * rho / dmax(fabs(a_d), 1.0e-10);
                const double thr = requested / safety;
                if (isfinite(dre) && dre <= thr) {
                    method = LCBG_METHOD_POINT_SOURCE;
                    mag = a_d;
                }
            }
            if (method < 0) {
                // hexadecapole_binary (:792-847): 12 samples on lanes 0..11
                const double h = 0.70710678118654752440;
                const double sineta[8] = {0.0, 1.0, 0.0, -1.0, h, h, -h, -h};
                const double coseta[8] = {1.0, 0.0, -1.0, 0.0, h, -h, -h, h};
                double my = 0.0;
                if (lane < 12) {
                    const int i = lane / 3;
                    const int kind = lane - 3 * i;
                    double sx, sy;
                    if (kind == 0) {
                        sx = src.x + rho * coseta[i];
                        sy = src.y + rho * sineta[i];
                    } else if (kind == 1) {
                        sx = src.x + 0.5 * rho * coseta[i];
                        sy = src.y + 0.5 * rho * sineta[i];
                    } else {
                        sx = src.x + rho * coseta[i + 4];
                        sy = src.y + rho * sineta[i + 4];
                    }
                    my = binary_mag0(cv.s, cv.q_eff, sx, sy).mag;
                }
                double a1p = 0.0, a2p = 0.0, a1x = 0.0;
                for (int i = 0; i < 4; ++i) {
                    a1p += uni(shfl(my, 3 * i));
                    a2p += uni(shfl(my, 3 * i + 1));
                    a1x += uni(shfl(my, 3 * i + 2));
                }
                a1p = a1p / 4.0 - a_ps;
                a2p = a2p / 4.0 - a_ps;
                a1x = a1x / 4.0 - a_ps;
                const double a2rho2 = (16.0 * a2p - a1p) / 3.0;
                const double a4rho4 = (a1p + a1x) / 2.0 - a2rho2;
                const double den = 15.0 - 5.0 * cv.ld_c - 3.0 * cv.ld_d;
                const double gamma = den == 0.0 ? 0.0 : 10.0 * cv.ld_c / den;
                const double lambda = den == 0.0 ? 0.0 : 12.0 * cv.ld_d / den;
                const double quad = 0.5 * a2rho2 * (1.0 - 0.2 * gamma - lambda / 9.0);
                const double hex = a4rho4 / 3.0 * (1.0 - 11.0 * gamma / 35.0 - 7.0 * lambda / 39.0);
                const double hmag = a_ps + quad + hex;
                const double rel = fabs(hex) / dmax(fabs(hmag), 1.0e-10);
                if (rel <= opt.adaptive_hex) {
                    method = LCBG_METHOD_HEXADECAPOLE;
                    mag = hmag;
                }
            }
        }
        if (method >= 0) {
            wave_store(&ep.mag[e], mag);
            wave_store(&ep.method[e], method);
        } else {
            const bool auto_polar = opt.mode == 4 && fabs(a_ps) >= 100.0;
            if (opt.mode == 3) {
                wave_store(&ep.mag[e], __builtin_nan(""));
                wave_store(&ep.method[e], -1);
                atomicOr(&ep.counters[4], kErrUnsupported);
            } else {
                wave_store(&ep.method[e], (opt.mode == 2 || auto_polar) ? LCBG_METHOD_INVERSE_RAY_POLAR
                                                                       : LCBG_METHOD_INVERSE_RAY_CARTESIAN);
                wave_store(&ep.dist[e], refined);
                const int slot = wave_fetch_inc(&ep.counters[2]);
                wave_store(&ep.ir_list[slot], e);
            }
        }
    }
}

__global__ void k_reset_counters(unsigned* counters, int n)
{
    for (int i = static_cast<int>(threadIdx.x); i < n; i += static_cast<int>(blockDim.x)) counters[i] = 0u;
}

lcbg_status launch_reset_counters(unsigned* counters, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, counters, n);
    return check_launch();
}

lcbg_status launch_fs_stage1(const CurveRec* d_curves, int n_curves, const FsCache& cache,
                             const FsOptions& opt, const FsEpochs& ep, hipStream_t st)
{
    if (n_curves <= 0) return LCBG_OK;
    hipLaunchKernelGGL(k_fs_stage1, dim3(static_cast<unsigned>(n_curves)), dim3(kStage1Block), 0,
                       st, d_curves, cache, opt, ep);
    return check_launch();
}

lcbg_status launch_fs_stage2(const CurveRec* d_curves, const FsCache& cache, const FsOptions& opt,
                             const FsEpochs& ep, int n_waves, hipStream_t st)
{
    hipLaunchKernelGGL(k_fs_stage2, dim3(static_cast<unsigned>(n_waves)), dim3(kWave), 0, st,
                       d_curves, cache, opt, ep);
    return check_launch();
}

}  // namespace lcbg
