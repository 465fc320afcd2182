// Wave64 helpers for the finite-source kernels.  These kernels run one
// 64-lane workgroup (= one wavefront) per work item: the item's control program
// is executed uniformly by every lane, and data-parallel sections (root solves,
// probe batches, lattice cells of a row) are spread over the lanes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fs_types.hpp"

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

// A double moved across lanes by a DPP control (GFX9 encodings: quad_perm
// 0x00-0xff, wave_shr:1 0x138, row_mirror 0x140, row_half_mirror 0x141,
// row_bcast:15 0x142, row_bcast:31 0x143); lanes that receive nothing (rows
// outside kRowMask, lane 0 of a shift) get 0.  These moves stay in the VALU,
// where __shfl goes through the LDS crossbar.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ double dpp_move(double v)
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(u & 0xffffffffull), kCtrl, kRowMask, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(u >> 32), kCtrl, kRowMask, 0xf, false);
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                                       static_cast<unsigned>(lo)));
}

// Value of lane - 1 (lane 0: 0).
__device__ __forceinline__ double lane_prev(double v) { return dpp_move<0x138>(v); }

// Sum over the wave, as a scalar: pairs, quads, half rows, rows by DPP, then
// row 1 += row 0 and row 3 += row 2 (row_bcast:15), rows 2-3 += rows 0-1
// (row_bcast:31); lane 63 holds the total.
__device__ __forceinline__ double read_lane(double v, int l);
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_move<0xB1>(v);       // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);       // quad_perm [2,3,0,1]
    v += dpp_move<0x141>(v);      // row_half_mirror
    v += dpp_move<0x140>(v);      // row_mirror
    v += dpp_move<0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3
    return read_lane(v, kWave - 1);
}

// v of lane l (l uniform), as a scalar.
__device__ __forceinline__ double read_lane(double v, int l)
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u & 0xffffffffull), l));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), l));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
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

// The wave kernels run one wavefront per workgroup, and their control flow is
// scalar (below), so every memory operation is issued by the whole wave in
// program order; the hardware performs one wavefront's accesses in that order
// (the AMDGPU memory model needs no wait for wavefront-scope ordering).
// wave_sync() therefore only has to keep the compiler from moving memory
// operations across it: a wavefront-scope fence, which emits no wait for
// outstanding stores, plus a wave barrier for convergent operations.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Values of the wave's control program are the same on every lane.  uni()
// hands such a value to the compiler as a scalar (lane 0's copy, which equals
// every other lane's), so the branches that depend on it are scalar branches:
// the control program then stays converged as written, instead of being
// restructured into per-lane loops around the lane-0 stores it relies on.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned uni(unsigned v)
{
    return static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}
__device__ __forceinline__ bool uni(bool v) { return __builtin_amdgcn_readfirstlane(static_cast<int>(v)) != 0; }
__device__ __forceinline__ long long uni(long long v)
{
    const unsigned long long u = static_cast<unsigned long long>(v);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(u & 0xffffffffull)));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(u >> 32)));
    return static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo);
}
__device__ __forceinline__ long uni(long v) { return static_cast<long>(uni(static_cast<long long>(v))); }
__device__ __forceinline__ double uni(double v) { return __longlong_as_double(uni(__double_as_longlong(v))); }

__device__ __forceinline__ Pt2 uni(Pt2 p) { return Pt2{uni(p.x), uni(p.y)}; }

// Side effects of the control program carry no `lane == 0` branch.  Such a
// branch is loop-invariant, and the compiler may thread it across a loop's back
// edge, which splits the wave into a lane-0 loop and a loop of the other lanes
// that never meet again.  Instead every lane performs the operation: a store
// writes lane 0's value from every lane, and a counter is advanced by every lane
// with an increment that is zero except on lane 0.

// *p = v, v being uniform.
template <class T>
struct same_type {
    using type = T;
};
template <class T>
__device__ __forceinline__ void wave_store(T* p, typename same_type<T>::type v)
{
    *p = uni(v);
}

// atomicAdd(counter, 1) on behalf of the wave; every lane receives the old value.
__device__ __forceinline__ int wave_fetch_inc(unsigned* counter)
{
    const unsigned inc = lane_id() == 0 ? 1u : 0u;
    return __builtin_amdgcn_readfirstlane(static_cast<int>(atomicAdd(counter, inc)));
}

// atomicAdd(counter, n) on behalf of the wave (n uniform); every lane receives
// the old value.
__device__ __forceinline__ int wave_fetch_add(unsigned* counter, unsigned n)
{
    const unsigned inc = lane_id() == 0 ? n : 0u;
    return __builtin_amdgcn_readfirstlane(static_cast<int>(atomicAdd(counter, inc)));
}

// Next item of a work queue, as a scalar.
__device__ __forceinline__ int next_item(unsigned* cursor) { return wave_fetch_inc(cursor); }

}  // namespace lcbg
