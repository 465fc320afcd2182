// Point-source kernels and their C-ABI entry points.
//
// Work decomposition: one thread = one source position (one quintic solve).
// Light curves: one workgroup per curve walks that curve's epochs, so the
// curve's record is workgroup-uniform (scalar loads into SGPRs) and the
// trajectory (python/lcbinint_pybind.cpp:1144-1148) is fused into the solve.
#include <vector>

#include "binary_lens.hpp"
#include "lcbinint_amd.h"
#include "runtime.hpp"

namespace lcbg {

constexpr int kPsBlock = 256;

__global__ void __launch_bounds__(kPsBlock) k_binary_mag0(const double* __restrict__ s,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ y,
                                                          int64_t n, double* __restrict__ mag,
                                                          int32_t* __restrict__ image_count)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kPsBlock + threadIdx.x;
    if (i >= n) return;
    const PointMag r = binary_mag0(s[i], q[i], x[i], y[i]);
    mag[i] = r.mag;
    if (image_count) image_count[i] = r.images;
}

__global__ void __launch_bounds__(kPsBlock) k_light_curves_ps(const CurveRec* __restrict__ curves,
                                                              const int64_t* __restrict__ offsets,
                                                              const double* __restrict__ times,
                                                              double* __restrict__ mag)
{
    const int c = blockIdx.x;
    const CurveRec cv = curves[c];
    const int64_t begin = offsets[c];
    const int64_t end = offsets[c + 1];
    for (int64_t i = begin + threadIdx.x; i < end; i += kPsBlock) {
        const double tau = (times[i] - cv.t0) * cv.inverse_tE;
        const double sx = tau * cv.cos_alpha - cv.u0 * cv.sin_alpha;
        const double sy = tau * cv.sin_alpha + cv.u0 * cv.cos_alpha;
        mag[i] = binary_mag0(cv.s, cv.q_eff, sx, sy).mag;
    }
}

lcbg_status launch_light_curves_ps(const CurveRec* d_curves, int32_t n_curves,
                                   const int64_t* offsets, const double* times, double* mag,
                                   hipStream_t st)
{
    KernelTimer timer(st);
    hipLaunchKernelGGL(k_light_curves_ps, dim3(static_cast<unsigned>(n_curves)), dim3(kPsBlock), 0,
                       st, d_curves, offsets, times, mag);
    timer.stop();
    return check_launch();
}

}  // namespace lcbg

using namespace lcbg;

extern "C" lcbg_status lcbg_binary_mag0_device(const double* s, const double* q, const double* x,
                                               const double* y, int64_t n, double* mag,
                                               int32_t* image_count, void* stream)
{
    if (n < 0 || (n > 0 && (!s || !q || !x || !y || !mag))) return LCBG_INVALID_ARGUMENT;
    if (n == 0) return LCBG_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t blocks = (n + kPsBlock - 1) / kPsBlock;
    if (blocks > 0x7fffffffLL) return LCBG_INVALID_ARGUMENT;
    KernelTimer timer(st);
    hipLaunchKernelGGL(k_binary_mag0, dim3(static_cast<unsigned>(blocks)), dim3(kPsBlock), 0, st,
                       s, q, x, y, n, mag, image_count);
    timer.stop();
    return check_launch();
}

extern "C" lcbg_status lcbg_batch_point_source_device(lcbg_batch* batch, const int64_t* offsets,
                                                      const double* times, double* mag,
                                                      void* stream)
{
    if (!batch) return LCBG_INVALID_ARGUMENT;
    if (batch->n_curves == 0) return LCBG_OK;
    if (!offsets || !times || !mag) return LCBG_INVALID_ARGUMENT;
    return launch_light_curves_ps(batch->d_curves, batch->n_curves, offsets, times, mag,
                                  static_cast<hipStream_t>(stream));
}

extern "C" lcbg_status lcbg_light_curve_point_source(const lcbg_curve* curve, const double* times,
                                                     int64_t n, double* mag)
{
    if (!curve || n < 0 || (n > 0 && (!times || !mag))) return LCBG_INVALID_ARGUMENT;
    if (n == 0) return LCBG_OK;
    lcbg_batch* b = nullptr;
    lcbg_status st = lcbg_batch_create(curve, 1, nullptr, &b);
    if (st != LCBG_OK) return st;
    double* d_times = nullptr;
    double* d_mag = nullptr;
    int64_t* d_off = nullptr;
    const int64_t off[2] = {0, n};
    const size_t bytes = static_cast<size_t>(n) * sizeof(double);
    st = map_hip(hipMalloc(reinterpret_cast<void**>(&d_times), bytes));
    if (st == LCBG_OK) st = map_hip(hipMalloc(reinterpret_cast<void**>(&d_mag), bytes));
    if (st == LCBG_OK) st = map_hip(hipMalloc(reinterpret_cast<void**>(&d_off), sizeof(off)));
    if (st == LCBG_OK) st = map_hip(hipMemcpy(d_times, times, bytes, hipMemcpyHostToDevice));
    if (st == LCBG_OK) st = map_hip(hipMemcpy(d_off, off, sizeof(off), hipMemcpyHostToDevice));
    if (st == LCBG_OK) st = lcbg_batch_point_source_device(b, d_off, d_times, d_mag, nullptr);
    if (st == LCBG_OK) st = map_hip(hipMemcpy(mag, d_mag, bytes, hipMemcpyDeviceToHost));
    if (d_times) (void)hipFree(d_times);
    if (d_mag) (void)hipFree(d_mag);
    if (d_off) (void)hipFree(d_off);
    lcbg_batch_destroy(b);
    return st;
}
