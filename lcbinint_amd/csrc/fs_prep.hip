// Per-curve precomputation for the finite-source path: everything that depends
// only on the lens (s, q) and not on the epoch.
//
//  * the caustic cache of ensure_binary_caustic_cache
//    (finite_source_magnifier.cpp:4014-4105): critical-curve quartic at `bins`
//    phases, caustic points through the complex lens equation, branches tracked
//    by the minimum-cost permutation (append_tracked_caustic_points :569-611),
//    bounding box, 128x128 grid cell of every point, longest non-wrap segment;
//  * the critical sources of the 1400-phase seed scan (critical_sources_at_phase
//    :1622-1641) and the greedy branch matching of seed Phase 2 (:1813-1840),
//    which the reference recomputes for every inverse-ray epoch although they
//    depend on (s, q) only.
//
// One 256-thread workgroup per curve for the root solves (one thread each),
// then one wave per curve for the sequential branch tracking (the 24
// permutations across lanes), then the branch positions sorted by cell.
#include "fs_launch.hpp"
#include "fs_math.hpp"
#include "fs_wave.hpp"

namespace lcbg {

constexpr int kPrepBlock = 256;

__constant__ unsigned char c_perm[24][4] = {
    {0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 1, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {0, 3, 2, 1},
    {1, 0, 2, 3}, {1, 0, 3, 2}, {1, 2, 0, 3}, {1, 2, 3, 0}, {1, 3, 0, 2}, {1, 3, 2, 0},
    {2, 0, 1, 3}, {2, 0, 3, 1}, {2, 1, 0, 3}, {2, 1, 3, 0}, {2, 3, 0, 1}, {2, 3, 1, 0},
    {3, 0, 1, 2}, {3, 0, 2, 1}, {3, 1, 0, 2}, {3, 1, 2, 0}, {3, 2, 0, 1}, {3, 2, 1, 0}};

__device__ __forceinline__ int grid_index(double v, double lo, double step)
{
    return iclamp(static_cast<int>((v - lo) / step), 0, kGrid - 1);
}

// The two (cos, sin) pairs are the same bits.
__device__ __forceinline__ bool same_phase(Pt a, Pt b)
{
    return __double_as_longlong(a.x) == __double_as_longlong(b.x) &&
           __double_as_longlong(a.y) == __double_as_longlong(b.y);
}

// One of four values by the bits of k (0..3): selects between registers.  The
// four points are separate values, not an array, so that a per-lane k cannot
// become an indexed access (which would go through scratch memory).
__device__ __forceinline__ double pick4d(double a, double b, double c, double d, int k)
{
    const double ab = (k & 1) ? b : a;
    const double cd = (k & 1) ? d : c;
    return (k & 2) ? cd : ab;
}

__device__ __forceinline__ Pt pick4(Pt c0, Pt c1, Pt c2, Pt c3, int k)
{
    return Pt{pick4d(c0.x, c1.x, c2.x, c3.x, k), pick4d(c0.y, c1.y, c2.y, c3.y, k)};
}

// (cost, index) pairs ordered by cost, then index, NaN costs last: the order
// of wave_argmin, a total order, so any reduction tree gives its result.
__device__ __forceinline__ void argmin_take(double& v, int& idx, double ov, int oi)
{
    const bool take = (ov < v) || (ov == v && oi < idx) || (v != v && ov == ov);
    v = take ? ov : v;
    idx = take ? oi : idx;
}

template <int kCtrl>
__device__ __forceinline__ void argmin_dpp(double& v, int& idx)
{
    const double ov = dpp_move<kCtrl>(v);
    const int oi = __builtin_amdgcn_update_dpp(0, idx, kCtrl, 0xf, 0xf, false);
    argmin_take(v, idx, ov, oi);
}

// The smallest (cost, index) of lanes 0-31, as scalars: quads, half rows and
// rows by DPP (every lane of a row has a source), then rows 0 and 1.
__device__ __forceinline__ void argmin_rows01(double& v, int& idx)
{
    argmin_dpp<0xB1>(v, idx);   // quad_perm [1,0,3,2]
    argmin_dpp<0x4E>(v, idx);   // quad_perm [2,3,0,1]
    argmin_dpp<0x141>(v, idx);  // row_half_mirror
    argmin_dpp<0x140>(v, idx);  // row_mirror
    const double v1 = read_lane(v, 16);
    const int i1 = __builtin_amdgcn_readlane(idx, 16);
    v = read_lane(v, 0);
    idx = __builtin_amdgcn_readlane(idx, 0);
    argmin_take(v, idx, v1, i1);
}

__global__ void __launch_bounds__(kPrepBlock) k_curve_prep(const CurveRec* __restrict__ curves,
                                                           FsCache cache, FsTables tab)
{
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    const CurveRec cv = curves[c];
    const int bins = cache.bins;
    Pt* pts = cache.pts + static_cast<size_t>(c) * bins * 4;
    int* pt_cell = cache.pt_cell + static_cast<size_t>(c) * bins * 4;
    Pt* crit = cache.crit + static_cast<size_t>(c) * kSeedPhases * 4;
    uint8_t* assign = cache.assign + static_cast<size_t>(c) * kSeedPhases * 4;
    CausticScalars* sc = cache.sc + c;

    __shared__ double red[4][kPrepBlock];

    if (!(cv.rho > 0.0) || cv.s == 0.0 || !(cv.q_eff > 0.0)) {
        if (tid == 0) sc->n = 0;
        return;
    }
    const Lens L = make_lens(cv.s, cv.q_eff);

    // A seed-scan phase whose cos and sin round to the caustic sample's (752 of
    // 1400 at 1400 bins) has the same quartic, hence the same roots: one solve
    // gives both.
    for (int i = tid; i < bins; i += kPrepBlock) {
        const Pt ph = tab.caus_trig[i];
        cx roots[4];
        critical_roots(L, cx{ph.x, ph.y}, roots);
#pragma unroll
        for (int r = 0; r < 4; ++r) pts[4 * i + r] = lens_equation_complex(L, roots[r]);
        if (i < kSeedPhases && same_phase(tab.seed_trig[i], ph)) {
#pragma unroll
            for (int r = 0; r < 4; ++r) crit[4 * i + r] = map_real(L, roots[r].re, roots[r].im);
        }
    }
    for (int k = tid; k < kSeedPhases; k += kPrepBlock) {
        const Pt ph = tab.seed_trig[k];
        if (k < bins && same_phase(tab.caus_trig[k], ph)) continue;
        cx roots[4];
        critical_roots(L, cx{ph.x, ph.y}, roots);
#pragma unroll
        for (int r = 0; r < 4; ++r) crit[4 * k + r] = map_real(L, roots[r].re, roots[r].im);
    }
    __syncthreads();

    // bounding box (finite_source_magnifier.cpp:4034-4041); min/max are order-free
    double mnx = __builtin_inf(), mxx = -__builtin_inf(), mny = __builtin_inf(), mxy = -__builtin_inf();
    for (int i = tid; i < bins * 4; i += kPrepBlock) {
        const Pt p = pts[i];
        mnx = dmin(mnx, p.x);
        mxx = dmax(mxx, p.x);
        mny = dmin(mny, p.y);
        mxy = dmax(mxy, p.y);
    }
    red[0][tid] = mnx;
    red[1][tid] = mxx;
    red[2][tid] = mny;
    red[3][tid] = mxy;
    __syncthreads();
    for (int off = kPrepBlock / 2; off >= 1; off >>= 1) {
        if (tid < off) {
            red[0][tid] = dmin(red[0][tid], red[0][tid + off]);
            red[1][tid] = dmax(red[1][tid], red[1][tid + off]);
            red[2][tid] = dmin(red[2][tid], red[2][tid + off]);
            red[3][tid] = dmax(red[3][tid], red[3][tid + off]);
        }
        __syncthreads();
    }
    const double min_x = red[0][0], max_x = red[1][0], min_y = red[2][0], max_y = red[3][0];
    const double width = dmax(max_x - min_x, 1.0e-12);
    const double height = dmax(max_y - min_y, 1.0e-12);
    const double step_x = width / static_cast<double>(kGrid);
    const double step_y = height / static_cast<double>(kGrid);
    for (int i = tid; i < bins * 4; i += kPrepBlock) {
        const Pt p = pts[i];
        pt_cell[i] = grid_index(p.y, min_y, step_y) * kGrid + grid_index(p.x, min_x, step_x);
    }
    __syncthreads();

    if (tid == 0) {
        sc->min_x = min_x;
        sc->max_x = max_x;
        sc->min_y = min_y;
        sc->max_y = max_y;
        sc->step_x = step_x;
        sc->step_y = step_y;
        sc->n = bins;
    }
    // seed Phase-2 greedy nearest-neighbour matching (:1813-1840): the
    // reference starts every branch at (1e30, 1e30) with a best distance of
    // 1e60, and each squared distance to that start is about 2e60 (or NaN), so
    // no branch is ever matched and every start stays put: all 255
    for (int k = tid; k < kSeedPhases * 4; k += kPrepBlock) assign[k] = 255;
}

// The four points of each of `bins` phases (pts [phase][4]) tracked into four
// branches (br [branch][bins]) by append_tracked_caustic_points
// (finite_source_magnifier.cpp:569-611): the first phase ordered by atan2,
// every later one by the permutation of least total squared step, the first
// such permutation in std::next_permutation order, the identity when no cost
// is below +inf.  One wave: 64 phases per coalesced load held in lanes, each
// phase's points taken by readlane and picked by register selects, the 24
// permutations priced on lanes 0-23, argmin by DPP.
__device__ __forceinline__ void track_branches(const Pt* pts, int bins, Pt* br, int lane)
{
    Pt prev[4];
    {
        // first phase: order by atan2 (std::sort of 4 elements = insertion sort)
        Pt p[4];
        double key[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            p[r] = pts[r];
            key[r] = atan2(p[r].y, p[r].x);
        }
        for (int i = 1; i < 4; ++i) {
            const Pt vp = p[i];
            const double vk = key[i];
            int j = i;
            while (j > 0 && vk < key[j - 1]) {
                p[j] = p[j - 1];
                key[j] = key[j - 1];
                --j;
            }
            p[j] = vp;
            key[j] = vk;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) prev[r] = p[r];
#pragma unroll
        for (int r = 0; r < 4; ++r) wave_store(&br[r * bins], prev[r]);
    }
    const int pl = lane < 24 ? lane : 0;
    const int p0 = c_perm[pl][0], p1 = c_perm[pl][1], p2 = c_perm[pl][2], p3 = c_perm[pl][3];
    for (int base = 0; base < bins; base += kWave) {
        Pt ch[4], keep[4];
        const int ph = base + lane;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            ch[r] = ph < bins ? pts[4 * ph + r] : Pt{0.0, 0.0};
            keep[r] = Pt{0.0, 0.0};
        }
        const int i0 = base > 0 ? base : 1;
        const int i1 = base + kWave < bins ? base + kWave : bins;
        for (int i = i0; i < i1; ++i) {
            const int j = i - base;
            const Pt c0{read_lane(ch[0].x, j), read_lane(ch[0].y, j)};
            const Pt c1{read_lane(ch[1].x, j), read_lane(ch[1].y, j)};
            const Pt c2{read_lane(ch[2].x, j), read_lane(ch[2].y, j)};
            const Pt c3{read_lane(ch[3].x, j), read_lane(ch[3].y, j)};
            double cost = 0.0;
            cost += dist2(prev[0], pick4(c0, c1, c2, c3, p0));
            cost += dist2(prev[1], pick4(c0, c1, c2, c3, p1));
            cost += dist2(prev[2], pick4(c0, c1, c2, c3, p2));
            cost += dist2(prev[3], pick4(c0, c1, c2, c3, p3));
            if (lane >= 24) cost = __builtin_inf();
            int idx = lane;
            argmin_rows01(cost, idx);
            // the reference keeps the identity when no cost is below +inf
            if (!(cost < __builtin_inf())) idx = 0;
            const int b0 = __builtin_amdgcn_readlane(p0, idx), b1 = __builtin_amdgcn_readlane(p1, idx);
            const int b2 = __builtin_amdgcn_readlane(p2, idx), b3 = __builtin_amdgcn_readlane(p3, idx);
            prev[0] = pick4(c0, c1, c2, c3, b0);
            prev[1] = pick4(c0, c1, c2, c3, b1);
            prev[2] = pick4(c0, c1, c2, c3, b2);
            prev[3] = pick4(c0, c1, c2, c3, b3);
            const bool mine = lane == j;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                keep[r].x = mine ? prev[r].x : keep[r].x;
                keep[r].y = mine ? prev[r].y : keep[r].y;
            }
        }
        if (ph >= i0 && ph < i1) {
#pragma unroll
            for (int r = 0; r < 4; ++r) br[r * bins + ph] = keep[r];
        }
    }
}

// Branch tracking of the caustic cache, one wave per curve so that every
// curve's sequential chain of phases runs at once (track_branches).  Then the
// branch grid cells and the longest non-wrap segment (:4065-4095).  Stage 1
// needs only what k_curve_prep writes (bounding box, n), so it may run beside
// this kernel.
__global__ void __launch_bounds__(kWave) k_branch_track(const CurveRec* __restrict__ curves, FsCache cache,
                                                        bool only_near)
{
    const int c = blockIdx.x;
    if (only_near && cache.near_n[c] == 0) return;
    const CurveRec cv = curves[c];
    if (!(cv.rho > 0.0) || cv.s == 0.0 || !(cv.q_eff > 0.0)) return;  // no caustic cache
    const int lane = lane_id();
    const int bins = cache.bins;
    const Pt* pts = cache.pts + static_cast<size_t>(c) * bins * 4;
    Pt* br = cache.br + static_cast<size_t>(c) * bins * 4;
    int* br_cell = cache.br_cell + static_cast<size_t>(c) * bins * 4;
    CausticScalars* sc = cache.sc + c;
    track_branches(pts, bins, br, lane);
    __syncthreads();
    const double min_x = sc->min_x, min_y = sc->min_y, step_x = sc->step_x, step_y = sc->step_y;
    double seg2 = 0.0;
    for (int k = lane; k < 4 * bins; k += kWave) {
        const int b = k / bins;
        const int j = k - b * bins;
        const Pt pt = br[k];
        if (j < bins - 1) seg2 = dmax(seg2, dist2(pt, br[k + 1]));
        br_cell[k] = grid_index(pt.y, min_y, step_y) * kGrid + grid_index(pt.x, min_x, step_x);
    }
    seg2 = wave_max(seg2);
    wave_store(&sc->max_seg_len, __builtin_sqrt(seg2));
}

// The branch positions of each curve in (grid cell, position) order, the
// order in which augment_seeds_from_caustic_branches (:4845-4926) visits its
// candidate segments, and where each grid row starts in that order: a counting sort over the 128x128 cells in LDS, scattered
// 256 positions at a time with each position's rank among the earlier ones of
// its cell, which keeps positions of one cell in index order.
constexpr int kOrderBlock = 256;
constexpr int kCells = kGrid * kGrid;

__global__ void __launch_bounds__(kOrderBlock) k_branch_order(const CurveRec* __restrict__ curves,
                                                              FsCache cache, bool only_near)
{
    const int c = blockIdx.x;
    if (only_near && cache.near_n[c] == 0) return;
    const int tid = threadIdx.x;
    const CurveRec cv = curves[c];
    if (!(cv.rho > 0.0) || cv.s == 0.0 || !(cv.q_eff > 0.0)) return;  // no caustic cache
    const int n = 4 * cache.bins;
    const int* cell = cache.br_cell + static_cast<size_t>(c) * n;
    int* order = cache.br_order + static_cast<size_t>(c) * n;
    __shared__ int hist[kCells];
    __shared__ int part[kOrderBlock];
    __shared__ int chunk[kOrderBlock];
    for (int i = tid; i < kCells; i += kOrderBlock) hist[i] = 0;
    __syncthreads();
    for (int k = tid; k < n; k += kOrderBlock) atomicAdd(&hist[iclamp(cell[k], 0, kCells - 1)], 1);
    __syncthreads();
    // exclusive prefix sum: thread t owns cells [64 t, 64 t + 64)
    constexpr int kPer = kCells / kOrderBlock;
    int sum = 0;
    for (int i = 0; i < kPer; ++i) sum += hist[tid * kPer + i];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < kOrderBlock; off <<= 1) {
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - sum;
    for (int i = 0; i < kPer; ++i) {
        const int h = hist[tid * kPer + i];
        hist[tid * kPer + i] = run;
        run += h;
    }
    __syncthreads();
    // first sorted position of each grid row (read before the scatter below
    // advances the cell starts)
    int* row = cache.br_row + static_cast<size_t>(c) * (kGrid + 1);
    for (int iy = tid; iy <= kGrid; iy += kOrderBlock) row[iy] = iy < kGrid ? hist[iy * kGrid] : n;
    for (int base = 0; base < n; base += kOrderBlock) {
        const int k = base + tid;
        const int my = k < n ? iclamp(cell[k], 0, kCells - 1) : -1;
        chunk[tid] = my;
        __syncthreads();
        int rank = 0;
        for (int j = 0; j < tid; ++j) rank += chunk[j] == my ? 1 : 0;
        const int pos = k < n ? hist[my] + rank : 0;
        __syncthreads();
        if (k < n) {
            order[pos] = k;
            atomicAdd(&hist[my], 1);
        }
        __syncthreads();
    }
}

lcbg_status launch_curve_prep(const CurveRec* d_curves, int n_curves, const FsCache& cache,
                              const FsTables& tab, hipStream_t st)
{
    if (n_curves <= 0) return LCBG_OK;
    hipLaunchKernelGGL(k_curve_prep, dim3(static_cast<unsigned>(n_curves)), dim3(kPrepBlock), 0,
                       st, d_curves, cache, tab);
    return check_launch();
}

lcbg_status launch_branch_cache(const CurveRec* d_curves, int n_curves, const FsCache& cache, hipStream_t st,
                                bool only_near)
{
    if (n_curves <= 0) return LCBG_OK;
    hipLaunchKernelGGL(k_branch_track, dim3(static_cast<unsigned>(n_curves)), dim3(kWave), 0, st, d_curves,
                       cache, only_near);
    hipLaunchKernelGGL(k_branch_order, dim3(static_cast<unsigned>(n_curves)), dim3(kOrderBlock), 0,
                       st, d_curves, cache, only_near);
    return check_launch();
}

// The geometry helpers' critical curves and caustics (geometry_branches,
// python/lcbinint_pybind.cpp:1882-1926): the critical-curve quartic at n
// phases, one thread each (the same polynomial and solve as the caustic
// cache), its roots or their images through the lens equation; then the four
// branches tracked by one wave as the caustic cache tracks them.
__global__ void __launch_bounds__(kPrepBlock) k_geometry_points(double s, double q_eff, const Pt* __restrict__ trig,
                                                                int n, bool to_source, Pt* pts)
{
    const int i = static_cast<int>(blockIdx.x) * kPrepBlock + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    const Lens L = make_lens(s, q_eff);
    const Pt ph = trig[i];
    cx roots[4];
    critical_roots(L, cx{ph.x, ph.y}, roots);
#pragma unroll
    for (int r = 0; r < 4; ++r) pts[4 * i + r] = to_source ? lens_equation_complex(L, roots[r]) : Pt{roots[r].re, roots[r].im};
}

__global__ void __launch_bounds__(kWave) k_geometry_track(const Pt* __restrict__ pts, int n, Pt* br)
{
    track_branches(pts, n, br, lane_id());
}

lcbg_status launch_geometry_branches(double s, double q_eff, const Pt2* trig, int n, bool to_source, Pt2* pts,
                                     Pt2* br, hipStream_t st)
{
    if (n <= 0) return LCBG_OK;
    const unsigned blocks = static_cast<unsigned>((n + kPrepBlock - 1) / kPrepBlock);
    hipLaunchKernelGGL(k_geometry_points, dim3(blocks), dim3(kPrepBlock), 0, st, s, q_eff, trig, n, to_source, pts);
    hipLaunchKernelGGL(k_geometry_track, dim3(1), dim3(kWave), 0, st, pts, n, br);
    return check_launch();
}

}  // namespace lcbg
