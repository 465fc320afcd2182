// Finite-source dispatcher, FiniteSourceMagnifier::binary_mag
// (finite_source_magnifier.cpp:4633-4843), split into four kernels.
//
// Stage 1, one 256-thread workgroup per curve: source position, point-source
// solve and the caustic bounding-box exit (:4676-4698); every other epoch is
// appended to the stage-2 queue.  Most epochs of a light curve end here.
//
// Stage 2a, one 64-lane wave per queued epoch: the sampled and segment caustic
// distances (:4183-4246, :4107-4181) as lane-parallel minima over the curve's
// caustic points.  Stage 2b, one thread per queued epoch: the caustic-distance
// test and the point-source derivative test (:4699-4763).  Stage 2c: the
// hexadecapole test (:4765-4803), twelve solves of an epoch on twelve lanes and
// five epochs per wave.  Epochs that fail both tests are appended to the
// inverse-ray queue with their method (Cartesian or polar, :4805-4842).
#include "fs_launch.hpp"
#include "fs_math.hpp"
#include "fs_wave.hpp"

namespace lcbg {

constexpr int kStage1Block = 256;

// Stage 1 of one epoch e of curve c: source position, quintic roots and
// point-source magnification; epochs outside the caustic's bounding box (plus
// kinji rho) settle as point source.  Returns true if the epoch goes on to the
// caustic-distance test.
__device__ __forceinline__ bool stage1_epoch(const CurveRec& cv, const CausticScalars& sc, const FsOptions& opt,
                                             const FsEpochs& ep, int c, int64_t e)
{
    const double tau = (ep.times[e] - cv.t0) * cv.inverse_tE;
    const double u0 = ep.src_y ? ep.src_y[e] : cv.u0;
    const Pt src{tau * cv.cos_alpha - u0 * cv.sin_alpha,
                 tau * cv.sin_alpha + u0 * cv.cos_alpha};
    ep.curve_of[e] = c;
    ep.src[e] = src;
    ep.err[e] = 0.0;  // point source: no error estimate (:4668-4698)
    if (cv.s == 0.0 || cv.q_eff <= 0.0) {
        ep.mag[e] = 0.0;
        ep.method[e] = LCBG_METHOD_POINT_SOURCE;
        ep.a_ps[e] = 0.0;
        return false;
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
    const double margin = bbox_margin(opt, rho);
    if (!(rho > 0.0) || opt.mode <= 0 || sc.n == 0 || src.x < sc.min_x - margin || src.x > sc.max_x + margin ||
        src.y < sc.min_y - margin || src.y > sc.max_y + margin) {
        ep.mag[e] = a_ps;
        ep.method[e] = LCBG_METHOD_POINT_SOURCE;
        return false;
    }
    return true;
}

__global__ void __launch_bounds__(kStage1Block) k_fs_stage1(const CurveRec* __restrict__ curves,
                                                            FsCache cache, FsOptions opt,
                                                            FsEpochs ep)
{
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const CurveRec cv = curves[c];
    const int64_t begin = ep.offsets[c];
    const int64_t end = ep.offsets[c + 1];
    const CausticScalars sc = cache.sc[c];
    __shared__ int wave_near[kStage1Block / kWave];
    __shared__ int chunk_slot;
    int run = 0;  // queued epochs of this curve so far
    for (int64_t first = begin; first < end; first += kStage1Block) {
        const int64_t e = first + tid;
        const bool near = e < end && stage1_epoch(cv, sc, opt, ep, c, e);
        // queue this chunk's remaining epochs in order: after the curve's
        // earlier ones in near_curve, and as one block in near_list
        const uint64_t m = ballot(near);
        wave_near[wave] = popcount64(m);
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < kStage1Block / kWave; ++w) {
            before += w < wave ? wave_near[w] : 0;
            total += wave_near[w];
        }
        if (wave == 0) chunk_slot = wave_fetch_add(&ep.counters[kQNearCount], static_cast<unsigned>(total));
        __syncthreads();
        const int rank = before + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                             static_cast<unsigned>(m >> 32),
                             __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(m), 0u)));
        if (near) {
            ep.near_list[chunk_slot + rank] = e;
            ep.near_curve[begin + run + rank] = e;
        }
        run += total;
        __syncthreads();
    }
    // the curve's work items for the stage-2 distances, kNearChunk epochs each
    const int n_chunks = (run + kNearChunk - 1) / kNearChunk;
    if (wave == 0) {
        const int first = wave_fetch_add(&ep.counters[kQNearChunkCount], static_cast<unsigned>(n_chunks));
        for (int i = lane_id(); i < n_chunks; i += kWave)
            ep.near_chunks[first + i] = (static_cast<int64_t>(c) << 32) | static_cast<int64_t>(i * kNearChunk);
        wave_store(&cache.near_n[c], run);
    }
}

__device__ __forceinline__ int cell_coord(double v, double lo, double step)
{
    return iclamp(static_cast<int>((v - lo) / step), 0, kGrid - 1);
}

// Slot in a list for every lane with `want`, one atomic per wave; -1 elsewhere.
__device__ __forceinline__ int wave_append(unsigned* count, bool want)
{
    const uint64_t m = ballot(want);
    if (m == 0) return -1;
    const int leader = first_set(m);
    const int below = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
        static_cast<unsigned>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(m), 0u)));
    int base = 0;
    if (lane_id() == leader) base = static_cast<int>(atomicAdd(count, static_cast<unsigned>(popcount64(m))));
    base = shfl(base, leader);
    return want ? base + below : -1;
}

// Stage 1 for batches whose curves hold at most one epoch each (the per-epoch
// pseudo-curves of orbital motion): one lane per curve, the queues as
// k_fs_stage1 leaves them for one-epoch curves (their order aside).
__global__ void __launch_bounds__(kStage1Block) k_fs_stage1_flat(const CurveRec* __restrict__ curves,
                                                                 int n_curves, FsCache cache, FsOptions opt,
                                                                 FsEpochs ep)
{
    const int c = static_cast<int>(blockIdx.x) * kStage1Block + static_cast<int>(threadIdx.x);
    bool near = false;
    int64_t e = -1;
    if (c < n_curves) {
        const int64_t begin = ep.offsets[c];
        if (ep.offsets[c + 1] > begin) {
            e = begin;
            near = stage1_epoch(curves[c], cache.sc[c], opt, ep, c, e);
        }
        cache.near_n[c] = near ? 1 : 0;
    }
    const int slot = wave_append(&ep.counters[kQNearCount], near);
    const int chunk = wave_append(&ep.counters[kQNearChunkCount], near);
    if (near) {
        ep.near_list[slot] = e;
        ep.near_curve[e] = e;  // the curve's first epoch is e itself
        ep.near_chunks[chunk] = static_cast<int64_t>(c) << 32;
    }
}

// True if neither branch segment at position pos can come closer to p than
// `nearest`, the value the minimum starts at.  reach2 is
// ((nearest + max_seg_len)(1 + 1e-9))^2 and reach2_wrap the same with the
// larger of max_seg_len and the longest wrap-around segment of the curve
// (positions n - 1 -> 0), which max_seg_len does not bound: a segment of length l from a lies at least
// |p - a| - l from p, so beyond these radii it exceeds nearest by far more
// than the rounding of either value and leaves the minimum unchanged.
__device__ __forceinline__ bool segments_beyond(Pt p, Pt a, int pos, int n, double reach2, double reach2_wrap)
{
    const double dx = p.x - a.x;
    const double dy = p.y - a.y;
    const double d2 = dx * dx + dy * dy;
    return d2 > ((pos == 0 || pos == n - 1) ? reach2_wrap : reach2);
}

// Longest wrap-around segment (position n - 1 back to 0) of the four branches.
__device__ __forceinline__ double wrap_segment_len(const Pt* br, int bins, int n)
{
    double l2 = 0.0;
    for (int b = 0; b < 4; ++b) {
        const Pt* bb = br + static_cast<size_t>(b) * bins;
        l2 = dmax(l2, dist2(bb[n - 1], bb[0]));
    }
    return __builtin_sqrt(l2);
}

// Stage 2a, one wave per queued epoch (a static stride over the queue): the
// sampled caustic distance (binary_sampled_caustic_distance, :4183-4246) with
// the bounding-box margin, then the segment distance (binary_caustic_distance,
// :4107-4181) over the branch segments next to points in the cells within
// nearest + max_seg_len.  Both visit only the grid rows of their box, as runs
// of the curve's branch points sorted by cell; the branch points are the
// caustic points reordered (the same values and cells), and a minimum is exact
// in any order.  With no point inside the margin box, the nearest point is a
// minimum over all of them, as in the reference.
__global__ void __launch_bounds__(kWave) k_fs_near_dist(const CurveRec* __restrict__ curves,
                                                        FsCache cache, FsOptions opt, FsEpochs ep)
{
    const int lane = lane_id();
    const int count = uni(static_cast<int>(ep.counters[kQNearCount]));
    for (int item = static_cast<int>(blockIdx.x); item < count; item += static_cast<int>(gridDim.x)) {
        const int64_t e = uni(static_cast<long long>(ep.near_list[item]));
        const int c = uni(ep.curve_of[e]);
        const CausticScalars sc = cache.sc[c];
        const int bins = cache.bins;
        const size_t base = static_cast<size_t>(c) * bins * 4;
        const Pt* pts = cache.pts + base;
        const Pt* br = cache.br + base;
        const int* br_cell = cache.br_cell + base;
        const int* order = cache.br_order + base;
        const int* row = cache.br_row + static_cast<size_t>(c) * (kGrid + 1);
        const Pt src = ep.src[e];
        const double rho = curves[c].rho;

        const double margin = bbox_margin(opt, rho);
        const int ix0 = cell_coord(src.x - margin, sc.min_x, sc.step_x);
        const int ix1 = cell_coord(src.x + margin, sc.min_x, sc.step_x);
        const int iy0 = cell_coord(src.y - margin, sc.min_y, sc.step_y);
        const int iy1 = cell_coord(src.y + margin, sc.min_y, sc.step_y);
        double d2 = __builtin_inf();
        {
            const int lo = row[iy0], hi = row[iy1 + 1];
            for (int k = lo + lane; k < hi; k += kWave) {
                const int pos = order[k];
                const int cell = br_cell[pos];
                const int ix = cell - (cell / kGrid) * kGrid;
                if (ix < ix0 || ix > ix1) continue;
                const Pt p = br[pos];
                const double dx = src.x - p.x;
                const double dy = src.y - p.y;
                const double cand = dx * dx + dy * dy;
                if (cand < d2) d2 = cand;
            }
        }
        d2 = uni(wave_min(d2));
        double nearest;
        if (d2 != __builtin_inf()) {
            nearest = __builtin_sqrt(d2);
        } else {
            double d2all = __builtin_inf();
#pragma unroll 4
            for (int k = lane; k < 4 * bins; k += kWave) {
                const Pt p = pts[k];
                const double dx = src.x - p.x;
                const double dy = src.y - p.y;
                const double cand = dx * dx + dy * dy;
                if (cand < d2all) d2all = cand;
            }
            nearest = uni(__builtin_sqrt(wave_min(d2all)));
        }
        const double seg_radius = nearest + sc.max_seg_len;
        const int jx0 = cell_coord(src.x - seg_radius, sc.min_x, sc.step_x);
        const int jx1 = cell_coord(src.x + seg_radius, sc.min_x, sc.step_x);
        const int jy0 = cell_coord(src.y - seg_radius, sc.min_y, sc.step_y);
        const int jy1 = cell_coord(src.y + seg_radius, sc.min_y, sc.step_y);
        const int n = sc.n;
        const double reach = seg_radius * (1.0 + 1.0e-9);
        const double reach2 = reach * reach;
        const double reach_wrap =
            (nearest + dmax(sc.max_seg_len, n >= 2 ? wrap_segment_len(br, bins, n) : 0.0)) * (1.0 + 1.0e-9);
        const double reach2_wrap = reach_wrap * reach_wrap;
        double dseg = nearest;
        if (n >= 2) {
            const int lo = row[jy0], hi = row[jy1 + 1];
            for (int k = lo + lane; k < hi; k += kWave) {
                const int at = order[k];
                const int cell = br_cell[at];
                const int ix = cell - (cell / kGrid) * kGrid;
                if (ix < jx0 || ix > jx1) continue;
                const int b = at / bins;
                const int pos = at - b * bins;
                const Pt* bb = br + static_cast<size_t>(b) * bins;
                if (segments_beyond(src, bb[pos], pos, n, reach2, reach2_wrap)) continue;
                const int prev = pos > 0 ? pos - 1 : n - 1;
                const int next = (pos + 1) % n;
                dseg = dmin(dseg, point_segment_distance(src, bb[prev], bb[pos]));
                dseg = dmin(dseg, point_segment_distance(src, bb[pos], bb[next]));
            }
        }
        wave_store(&ep.dist[e], wave_min(dseg));
    }
}

// Stage 2a with the curve's caustic in LDS: one resident 1024-thread
// workgroup per CU takes work items (a curve and up to kNearChunk of its
// queued epochs) from the list stage 1 wrote; it loads the curve's branch
// points, their cells, their order by cell and the grid-row starts into LDS
// when the curve changes (the caustic points reordered: the same values and
// cells), and its waves take the item's epochs in turn, each visiting only the
// grid rows of its boxes.  Both distances are minima over LDS, exact in any order.  Used when
// 4 * bins points fit kLdsBins.
constexpr int kLdsBins = 1400;
constexpr int kNearCurveBlock = 1024;

__global__ void __launch_bounds__(kNearCurveBlock) k_fs_near_dist_curve(const CurveRec* __restrict__ curves,
                                                                        FsCache cache, FsOptions opt,
                                                                        FsEpochs ep)
{
    __shared__ Pt lds_pt[4 * kLdsBins];
    __shared__ int lds_cell[4 * kLdsBins];
    __shared__ float lds_seg[4 * kLdsBins];  // longer segment at each point, rounded up
    __shared__ int lds_order[4 * kLdsBins];  // positions sorted by (cell, position)
    __shared__ int lds_row[kGrid + 1];       // first entry of each grid row in lds_order
    __shared__ int s_item;
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = lane_id();
    const int wave = tid / kWave;
    const int bins = cache.bins;
    const int nk = 4 * bins;
    const int n_items = static_cast<int>(ep.counters[kQNearChunkCount]);
    int loaded = -1;
    for (;;) {
        if (wave == 0) wave_store(&s_item, wave_fetch_inc(&ep.counters[kQNearChunkCursor]));
        __syncthreads();
        const int item = s_item;
        if (item >= n_items) break;
        const int64_t desc = ep.near_chunks[item];
        const int c = static_cast<int>(desc >> 32);
        const int chunk = static_cast<int>(desc & 0xffffffff);
        const CausticScalars sc = cache.sc[c];
        const int n = sc.n;
        if (c != loaded) {
            const size_t base = static_cast<size_t>(c) * nk;
            for (int k = tid; k < nk; k += kNearCurveBlock) {
                lds_pt[k] = cache.br[base + k];
                lds_cell[k] = cache.br_cell[base + k];
                lds_order[k] = cache.br_order[base + k];
            }
            for (int r = tid; r <= kGrid; r += kNearCurveBlock)
                lds_row[r] = cache.br_row[static_cast<size_t>(c) * (kGrid + 1) + r];
            loaded = c;
            __syncthreads();
            if (n >= 2) {
                for (int k = tid; k < nk; k += kNearCurveBlock) {
                    const int b = k / bins;
                    const int pos = k - b * bins;
                    const Pt* bb = lds_pt + b * bins;
                    const Pt a = bb[pos];
                    const double l2 = dmax(dist2(a, bb[pos > 0 ? pos - 1 : n - 1]), dist2(a, bb[pos + 1 < n ? pos + 1 : 0]));
                    lds_seg[k] = __double2float_ru(__builtin_sqrt(l2));
                }
            }
            __syncthreads();
        }
        const int64_t begin = ep.offsets[c];
        const int count = cache.near_n[c];
        const double rho = curves[c].rho;
        const double margin = bbox_margin(opt, rho);
        // The near test reads the distance only through d < hex_threshold rho
        // and, for rho >= 1e-3, the safety factor max(1, 30 (hex_threshold /
        // max(d / rho, hex_threshold))^3); both are the same for every
        // d >= 3.2 hex_threshold rho, as 30 / 3.2^3 < 1.  Past `settle` (that
        // bound, slightly raised) the search stops refining: it returns some
        // value at or above it, which changes no decision.
        const double settle = opt.hex_threshold > 0.0 ? 3.2 * opt.hex_threshold * rho * (1.0 + 1.0e-10)
                                                      : __builtin_inf();
        const double wrap_len = n >= 2 ? wrap_segment_len(lds_pt, bins, n) : 0.0;
        const double any_len = dmax(sc.max_seg_len, wrap_len);
        const int stop = chunk + kNearChunk < count ? chunk + kNearChunk : count;
        for (int j = chunk + wave; j < stop; j += kNearCurveBlock / kWave) {
            const int64_t e = uni(static_cast<long long>(ep.near_curve[begin + j]));
            const Pt src = ep.src[e];
            const int ix0 = cell_coord(src.x - margin, sc.min_x, sc.step_x);
            const int ix1 = cell_coord(src.x + margin, sc.min_x, sc.step_x);
            const int iy0 = cell_coord(src.y - margin, sc.min_y, sc.step_y);
            const int iy1 = cell_coord(src.y + margin, sc.min_y, sc.step_y);
            double d2 = __builtin_inf();
            // the points of the box's grid rows, in (cell, position) order
            for (int t = lds_row[iy0] + lane; t < lds_row[iy1 + 1]; t += kWave) {
                const int k = lds_order[t];
                const int cell = lds_cell[k];
                const int ix = cell - (cell / kGrid) * kGrid;
                if (ix < ix0 || ix > ix1) continue;
                const Pt p = lds_pt[k];
                const double dx = src.x - p.x;
                const double dy = src.y - p.y;
                const double cand = dx * dx + dy * dy;
                if (cand < d2) d2 = cand;
            }
            d2 = uni(wave_min(d2));
            // With no caustic point in the margin box every point lies beyond
            // margin.  If margin is past settle and no wrap-around segment (which
            // max_seg_len does not bound) comes within settle, the nearest
            // point's exact distance cannot matter: only segments closer than
            // settle can, and both ends of such a segment lie within settle +
            // max_seg_len < nearest + max_seg_len, inside the reference's box.
            bool exact = true;
            double nearest = 0.0;
            if (d2 != __builtin_inf()) {
                nearest = __builtin_sqrt(d2);
            } else {
                // (the slack covers the rounding of src -/+ margin in the box)
                if (n >= 2 && settle + 1.0e-12 * (1.0 + fabs(src.x) + fabs(src.y)) < margin * (1.0 - 1.0e-9)) {
                    double wrap_min = __builtin_inf();
                    for (int b = 0; b < 4; ++b) {
                        const Pt* bb = lds_pt + b * bins;
                        wrap_min = dmin(wrap_min, point_segment_distance(src, bb[n - 1], bb[0]));
                    }
                    exact = !(wrap_min >= settle);
                }
                if (exact) {
                    double d2all = __builtin_inf();
#pragma unroll 2
                    for (int k = lane; k < nk; k += kWave) {
                        const Pt p = lds_pt[k];
                        const double dx = src.x - p.x;
                        const double dy = src.y - p.y;
                        const double cand = dx * dx + dy * dy;
                        if (cand < d2all) d2all = cand;
                    }
                    nearest = uni(__builtin_sqrt(wave_min(d2all)));
                } else {
                    nearest = settle;
                }
            }
            // the reference's box (nearest + max_seg_len), or the box that holds
            // every point within settle + the longest segment
            const double seg_radius = exact ? nearest + sc.max_seg_len : (settle + any_len) * (1.0 + 1.0e-8);
            const int jx0 = cell_coord(src.x - seg_radius, sc.min_x, sc.step_x);
            const int jx1 = cell_coord(src.x + seg_radius, sc.min_x, sc.step_x);
            const int jy0 = cell_coord(src.y - seg_radius, sc.min_y, sc.step_y);
            const int jy1 = cell_coord(src.y + seg_radius, sc.min_y, sc.step_y);
            // A segment of length l from point a lies at least |src - a| - l
            // from src; beyond (best + l)(1 + 1e-9), best being this lane's
            // minimum so far capped at settle, both segments at a exceed best
            // by far more than the rounding of either value and cannot change
            // the minimum below settle.
            double dseg = nearest;
            if (n >= 2) {
                for (int t = lds_row[jy0] + lane; t < lds_row[jy1 + 1]; t += kWave) {
                    const int k = lds_order[t];
                    const int cell = lds_cell[k];
                    const int ix = cell - (cell / kGrid) * kGrid;
                    if (ix < jx0 || ix > jx1) continue;
                    const int b = k >= 2 * bins ? (k >= 3 * bins ? 3 : 2) : (k >= bins ? 1 : 0);
                    const int pos = k - b * bins;
                    const Pt* bb = lds_pt + b * bins;
                    const double reach = (dmin(dseg, settle) + static_cast<double>(lds_seg[k])) * (1.0 + 1.0e-9);
                    if (dist2(src, bb[pos]) > reach * reach) continue;
                    const int prev = pos > 0 ? pos - 1 : n - 1;
                    const int next = pos + 1 < n ? pos + 1 : 0;
                    dseg = dmin(dseg, point_segment_distance(src, bb[prev], bb[pos]));
                    dseg = dmin(dseg, point_segment_distance(src, bb[pos], bb[next]));
                }
            }
            wave_store(&ep.dist[e], wave_min(dseg));
        }
        // every wave is done with this item (and the LDS curve) before the next
        __syncthreads();
    }
}

// An epoch that neither the point-source test nor the hexadecapole accepts:
// inverse-ray shooting, Cartesian or polar (:4805-4842); finite_mode 3
// (adaptive refinement) is outside this library.  Returns true if the epoch
// joins the inverse-ray queue.
__device__ __forceinline__ bool mark_inverse_ray(const FsOptions& opt, const FsEpochs& ep, int64_t e,
                                                 double a_ps)
{
    if (opt.mode == 3) {
        ep.mag[e] = __builtin_nan("");
        ep.method[e] = -1;
        atomicOr(&ep.counters[kQFlags], kErrUnsupported);
        return false;
    }
    const bool auto_polar = opt.mode == 4 && fabs(a_ps) >= 100.0;
    ep.method[e] = (opt.mode == 2 || auto_polar) ? LCBG_METHOD_INVERSE_RAY_POLAR
                                                 : LCBG_METHOD_INVERSE_RAY_CARTESIAN;
    return true;
}

// Stage 2b, one thread per queued epoch: the caustic-distance test
// (:4699-4718) and the point-source derivative test (:4719-4763) on the
// epoch's own quintic roots; epochs that pass neither are queued for the
// hexadecapole, epochs close to a caustic for inverse-ray shooting.
constexpr int kNearTestBlock = 256;

__global__ void __launch_bounds__(kNearTestBlock) k_fs_near_test(const CurveRec* __restrict__ curves,
                                                                 FsOptions opt, FsEpochs ep)
{
    const int count = static_cast<int>(ep.counters[kQNearCount]);
    const int stride = static_cast<int>(gridDim.x) * kNearTestBlock;
    for (int base = static_cast<int>(blockIdx.x) * kNearTestBlock; base < count; base += stride) {
        const int item = base + static_cast<int>(threadIdx.x);
        int64_t e = -1;
        bool hex = false, ir = false;
        if (item < count) {
            e = ep.near_list[item];
            const CurveRec& cv = curves[ep.curve_of[e]];
            const double rho = cv.rho;
            const double a_ps = ep.a_ps[e];
            const double refined = ep.dist[e];
            if (refined < opt.hex_threshold * rho) {
                ir = mark_inverse_ray(opt, ep, e, a_ps);
            } else if (opt.hex_threshold > 0.0) {
                double safety = 1.0;
                if (rho >= 1.0e-3 && isfinite(refined) && rho > 0.0) {
                    const double dist_ratio = refined / rho;
                    const double t = opt.hex_threshold / dmax(dist_ratio, opt.hex_threshold);
                    safety = dmax(1.0, 30.0 * t * t * t);
                }
                // the adaptive mode asks the tolerances of the point-source
                // magnification (:4731-4738)
                const double requested = opt.adaptive ? opt.reltol + opt.tol / dmax(fabs(a_ps), 1.0)
                                         : opt.adaptive_hex > 0.0 ? opt.adaptive_hex : 1.0e-3;
                const Pt src = ep.src[e];
                const Geom g = make_geom(cv.s, cv.q_eff, src.x, src.y);
                cx roots[5];
                const double* rr = ep.roots + 10 * e;
#pragma unroll
                for (int k = 0; k < 5; ++k) roots[k] = cx{rr[2 * k], rr[2 * k + 1]};
                // binary_mag0_with_derivatives on the same quintic as the
                // point-source solve: identical roots
                const unsigned mask = select_physical(g, roots);
                const double a_d = ps_from_roots(g, roots, mask).mag;
                const double ind = derivative_indicator(g, roots, mask);
                const double dre = ind * rho * rho / dmax(fabs(a_d), 1.0e-10);
                const double thr = requested / safety;
                if (isfinite(dre) && dre <= thr) {
                    ep.mag[e] = a_d;
                    ep.method[e] = LCBG_METHOD_POINT_SOURCE;
                    ep.err[e] = dre * dmax(fabs(a_d), 1.0);  // :4742-4751
                } else {
                    hex = true;
                }
            } else {
                hex = true;
            }
        }
        const int hs = wave_append(&ep.counters[kQHexCount], hex);
        if (hex) ep.hex_list[hs] = e;
        const int is = wave_append(&ep.counters[kQIrCount], ir);
        if (ir) ep.ir_list[is] = e;
    }
}

// Stage 2c, hexadecapole_binary (:792-847): the twelve point-source solves of
// an epoch on twelve consecutive lanes, five epochs per wave (lanes 60-63
// idle); the first lane of each group sums the samples in the reference's
// order and applies the hexadecapole test (:4765-4803).
constexpr int kHexPerWave = 5;

__global__ void __launch_bounds__(kWave) k_fs_hex(const CurveRec* __restrict__ curves, FsOptions opt,
                                                  FsEpochs ep)
{
    const int lane = lane_id();
    const int group = lane < kHexPerWave * 12 ? lane / 12 : kHexPerWave - 1;
    const int k = lane - 12 * group;  // sample of the group (>= 12 on lanes 60-63)
    const int count = static_cast<int>(ep.counters[kQHexCount]);
    for (int base = static_cast<int>(blockIdx.x) * kHexPerWave; base < count;
         base += static_cast<int>(gridDim.x) * kHexPerWave) {
        const int item = base + group;
        const bool valid = item < count && k < 12;
        int64_t e = 0;
        double my = 0.0;
        if (valid) {
            e = ep.hex_list[item];
            const CurveRec& cv = curves[ep.curve_of[e]];
            const Pt src = ep.src[e];
            const double rho = cv.rho;
            const double h = 0.70710678118654752440;
            const double sineta[8] = {0.0, 1.0, 0.0, -1.0, h, h, -h, -h};
            const double coseta[8] = {1.0, 0.0, -1.0, 0.0, h, -h, -h, h};
            const int i = k / 3;
            const int kind = k - 3 * i;
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
        const int g0 = 12 * group;
        double a1p = 0.0, a2p = 0.0, a1x = 0.0;
        for (int i = 0; i < 4; ++i) {
            a1p += shfl(my, g0 + 3 * i);
            a2p += shfl(my, g0 + 3 * i + 1);
            a1x += shfl(my, g0 + 3 * i + 2);
        }
        bool ir = false;
        if (valid && k == 0) {
            const CurveRec& cv = curves[ep.curve_of[e]];
            const double a_ps = ep.a_ps[e];
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
            double thr = opt.adaptive_hex;
            if (opt.adaptive) {
                // the tolerances of the hexadecapole magnification over the
                // caustic-distance safety factor (:4771-4790)
                const double requested = opt.reltol + opt.tol / dmax(fabs(hmag), 1.0);
                const double refined = ep.dist[e];
                double safety = 1.0;
                if (cv.rho >= 1.0e-3 && isfinite(refined) && cv.rho > 0.0) {
                    const double t = opt.hex_threshold / dmax(refined / cv.rho, opt.hex_threshold);
                    safety = dmax(1.0, 30.0 * t * t * t);
                }
                const double limit = requested / safety;
                thr = thr < limit ? thr : limit;
            }
            if (rel <= thr) {
                ep.mag[e] = hmag;
                ep.method[e] = LCBG_METHOD_HEXADECAPOLE;
                ep.err[e] = rel * dmax(fabs(hmag), 1.0);  // :4792-4802
            } else {
                ir = mark_inverse_ray(opt, ep, e, a_ps);
            }
        }
        const int is = wave_append(&ep.counters[kQIrCount], ir);
        if (ir) ep.ir_list[is] = e;
    }
}

// The first epoch whose magnification is not finite: counters[kFirstFailWord]
// = max over such epochs e of n - e (0: none).  The reference's array call stops
// at that epoch (src/lcbinint/lcbinint.cpp:110-119).
constexpr int kScanBlock = 256;

__global__ void __launch_bounds__(kScanBlock) k_first_nonfinite(const double* __restrict__ mag, int64_t n,
                                                                 unsigned* counters)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kScanBlock;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kScanBlock; base < n; base += stride) {
        const int64_t e = base + static_cast<int64_t>(threadIdx.x);
        const bool bad = e < n && !isfinite(mag[e]);
        const uint64_t m = ballot(bad);
        if (m) {
            const int64_t first = base + (static_cast<int64_t>(threadIdx.x) & ~int64_t{kWave - 1}) + first_set(m);
            if (lane_id() == 0) atomicMax(&counters[kFirstFailWord], static_cast<unsigned>(n - first));
            return;  // later epochs of this thread's range come after it
        }
    }
}

lcbg_status launch_first_nonfinite(const double* mag, int64_t n, unsigned* counters, hipStream_t st)
{
    if (n <= 0) return LCBG_OK;
    const int64_t blocks = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_first_nonfinite, dim3(static_cast<unsigned>(blocks < 2048 ? blocks : 2048)),
                       dim3(kScanBlock), 0, st, mag, n, counters);
    return check_launch();
}

// Per-epoch records of a finished run, the fields of lcbi_result
// (include/lcbinint/lcbinint.h:106-116) as evaluate_binary_light_curve_info
// (python/lcbinint_pybind.cpp:1211-1253) and LensModel::magnification
// (src/lcbinint/model/lens_model.cpp:186-234) fill them.
__global__ void __launch_bounds__(kScanBlock) k_epoch_results(const CurveRec* __restrict__ curves,
                                                               FsOptions opt, FsEpochs ep, double tol,
                                                               double reltol, lcbg_result* out)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kScanBlock;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kScanBlock + threadIdx.x; e < ep.n; e += stride) {
        const CurveRec cv = curves[ep.curve_of[e]];
        const Pt src = ep.src[e];
        const double mag = ep.mag[e];
        const int method = ep.method[e];
        int images = 0;
        if (!(cv.s == 0.0 || cv.q_eff <= 0.0)) {
            const Geom g = make_geom(cv.s, cv.q_eff, src.x, src.y);
            cx roots[5];
            const double* rr = ep.roots + 10 * e;
#pragma unroll
            for (int k = 0; k < 5; ++k) roots[k] = cx{rr[2 * k], rr[2 * k + 1]};
            // the LensModel path counts binary_images; the static path binary_mag0's
            images = (opt.ps_images && cv.rho != 0.0) ? binary_images_count(g, roots)
                                                       : popcount64(select_physical(g, roots));
        }
        const double err = ep.err[e];
        bool converged = isfinite(mag);
        // adaptive Cartesian epochs: the level k_ir_cart_adapt recorded
        const int lv =
            (ep.level && opt.adaptive_source_bins != 0 && method == LCBG_METHOD_INVERSE_RAY_CARTESIAN) ? ep.level[e] : 0;
        // fixed_inverse_ray_binary (finite_source_magnifier.cpp:3290-3294)
        if (method == LCBG_METHOD_INVERSE_RAY_CARTESIAN)
            converged = converged && err <= 0.999 * (tol + reltol * dmax(fabs(mag), 1.0)) &&
                        !(lv & kLevelUnconverged);
        lcbg_result r;
        r.magnification = mag;
        r.point_source_magnification = ep.a_ps[e];
        r.finite_source_magnification = mag;
        r.source_x = src.x;
        r.source_y = src.y;
        r.image_count = images;
        r.method = method;
        r.error_estimate = err;
        r.refinement_level = lv & (kLevelUnconverged - 1);
        r.converged = converged ? 1 : 0;
        out[e] = r;
    }
}

lcbg_status launch_epoch_results(const CurveRec* d_curves, const FsOptions& opt, const FsEpochs& ep,
                                 double tol, double reltol, lcbg_result* out, hipStream_t st)
{
    if (ep.n <= 0) return LCBG_OK;
    const int64_t blocks = (ep.n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_epoch_results, dim3(static_cast<unsigned>(blocks < 4096 ? blocks : 4096)),
                       dim3(kScanBlock), 0, st, d_curves, opt, ep, tol, reltol, out);
    return check_launch();
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
                             const FsOptions& opt, const FsEpochs& ep, hipStream_t st, bool one_epoch)
{
    if (n_curves <= 0) return LCBG_OK;
    if (one_epoch) {
        const unsigned blocks = static_cast<unsigned>((n_curves + kStage1Block - 1) / kStage1Block);
        hipLaunchKernelGGL(k_fs_stage1_flat, dim3(blocks), dim3(kStage1Block), 0, st, d_curves, n_curves, cache,
                           opt, ep);
    } else {
        hipLaunchKernelGGL(k_fs_stage1, dim3(static_cast<unsigned>(n_curves)), dim3(kStage1Block), 0,
                           st, d_curves, cache, opt, ep);
    }
    return check_launch();
}

lcbg_status launch_fs_stage2(const CurveRec* d_curves, int n_cus, const FsCache& cache,
                             const FsOptions& opt, const FsEpochs& ep, int n_waves, hipStream_t st)
{
    if (cache.bins <= kLdsBins) {
        hipLaunchKernelGGL(k_fs_near_dist_curve, dim3(static_cast<unsigned>(n_cus > 0 ? n_cus : 1)),
                           dim3(kNearCurveBlock), 0, st, d_curves, cache, opt, ep);
    } else {
        hipLaunchKernelGGL(k_fs_near_dist, dim3(static_cast<unsigned>(n_waves)), dim3(kWave), 0, st,
                           d_curves, cache, opt, ep);
    }
    lcbg_status s = check_launch();
    if (s != LCBG_OK) return s;
    hipLaunchKernelGGL(k_fs_near_test, dim3(static_cast<unsigned>(n_waves / 4)), dim3(kNearTestBlock), 0,
                       st, d_curves, opt, ep);
    s = check_launch();
    if (s != LCBG_OK) return s;
    hipLaunchKernelGGL(k_fs_hex, dim3(static_cast<unsigned>(n_waves)), dim3(kWave), 0, st, d_curves, opt, ep);
    return check_launch();
}

}  // namespace lcbg
