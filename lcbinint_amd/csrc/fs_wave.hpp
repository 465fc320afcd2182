// Wave64 helpers for the finite-source kernels.  These kernels run one
// 64-lane workgroup (= one wavefront) per work item: the item's control program
// is executed uniformly by every lane, and data-parallel sections (root solves,
// probe batches, lattice cells of a row) are spread over the lanes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lcbg {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int first_set(uint64_t m) { return m ? __ffsll(static_cast<unsigned long long>(m)) - 1 : -1; }

__device__ __forceinline__ int popcount64(uint64_t m) { return __popcll(static_cast<unsigned long long>(m)); }

__device__ __forceinline__ double shfl(double v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ int shfl(int v, int src) { return __shfl(v, src, kWave); }

// Wave minimum with std::min semantics for the NaN-free values it is used on.
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = (o < v) ? o : v;
    }
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = (v < o) ? o : v;
    }
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Minimum value and the smallest index attaining it (ties -> lowest index;
// indices with a NaN value never win).
__device__ __forceinline__ void wave_argmin(double& v, int& idx)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(v, off, kWave);
        const int oi = __shfl_xor(idx, off, kWave);
        const bool take = (ov < v) || (ov == v && oi < idx) || (v != v && ov == ov);
        if (take) {
            v = ov;
            idx = oi;
        }
    }
}

__device__ __forceinline__ void wave_sync() { __syncthreads(); }

}  // namespace lcbg
