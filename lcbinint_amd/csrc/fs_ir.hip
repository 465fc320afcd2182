// Inverse-ray shooting for the epochs the dispatcher could not settle:
// fixed_inverse_ray_binary (finite_source_magnifier.cpp:3042-3172) without the
// adaptive refinement, i.e. image seeds, then the Cartesian row scan or the
// polar flood fill on the reference's own lattice.
//
// One 64-lane wave per epoch, resident waves pulling epochs from a queue.  The
// epoch's control program (seed bookkeeping, row records, gap repair, overlap
// rules, the flood-fill queue) runs uniformly on all lanes; the lanes share the
// data-parallel work:
//   * seed search: critical sources come precomputed per curve; probe batches
//     (8 caustic probes, 80 angular, 400 boundary, 192 interior) are solved one
//     probe per lane, then merged in the reference's order;
//   * Cartesian rows: 64 consecutive lattice cells per step, the cell
//     coordinates reproducing the reference's repeated `x += incr` rounding
//     exactly (see seq_coord), the row end found by ballot;
//   * polar: radial runs and neighbour rows evaluated 64 cells at a time.
#include "fs_launch.hpp"
#include "fs_math.hpp"
#include "fs_wave.hpp"
#include "pq_heap.hpp"

namespace lcbg {

// ------------------------------------------------------------------ context

struct Ctx {
    Lens L;
    double s, q;   // separation and effective mass ratio as binary_mag receives them
    Pt src;
    double rho, r2, inv_r2;
    double ld_c, ld_d;
    bool uniform;
    int lane;
    Pt* seeds;
    int n_seeds;
    bool overflow;
    long long steps;  // wave steps spent on this epoch
    long long budget; // wave steps the epoch may take (FsOptions::step_budget)
    int last_phase;   // seed phases run: 0 images, 1 first crossing, 2 branches,
                      // 3 boundary, 4 interior, 5 caustic segments (diagnostics)
    // algorithmic work of the epoch (uniform): quintic and quartic solves, the
    // lattice cells the reference's scans evaluate (plain / Jacobian-guarded
    // Cartesian, polar)
    long long n_quintic, n_quartic, n_cart, n_cart_guarded, n_polar;
    long long n_rows;  // Cartesian rows closed plus polar runs recorded
    int row_cap;       // row records of this slot (Cartesian)
    int run_cap;       // visited runs of this slot (polar)
    bool cap_overflow; // a slot capacity (not the seed list or work budget) ran out
    // first-crossing scan hand-off to the probe farm (seed search only)
    const uint8_t* farm;  // probe flags of this search (resumed scan), or null
    int p1_start;         // first scan position whose probes are applied
    int defer_budget;     // inside positions to probe before deferring (<0: never)
    int defer_at;         // scan position the search was deferred at, or -1
    int defer_slot;       // its heavy-list slot
    unsigned* heavy_count;
    bool seeds_changed;   // a probe added or replaced a seed
    bool hints;           // center images are seed hints (LensModel path)
    // LegacyAreaDiagnostics of the Cartesian scan (finite_source_magnifier.cpp:1006-1015):
    // processed images, fold-guarded images, counted rows, gap repairs,
    // overlap hits, largest row-edge jump in cells
    int d_processed, d_fold, d_boundary_rows, d_gaps, d_overlaps;
    double d_max_jump;
};

// Counts one wave step; past the budget the epoch is abandoned with an error.
__device__ __forceinline__ bool step_ok(Ctx& C)
{
    C.steps = uni(C.steps + 1);
    if (C.steps > C.budget) C.overflow = true;
    return !C.overflow;
}

// ------------------------------------------------------- probe image search

// binary_image_candidates (point_source_magnifier.cpp:739-778) restricted to
// selected_point_images (finite_source_magnifier.cpp:1873-1889): roots sorted
// by residual, kept while residual^2 < 1e-12.
__device__ __forceinline__ int probe_r3_images(const Ctx& C, Pt probe, Pt* out)
{
    if (C.s == 0.0 || C.q <= 0.0) return 0;
    const Geom g = make_geom(C.s, C.q, probe.x, probe.y);
    cx roots[5];
    solve_quintic(g, roots);
    double res[5];
    int ord[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        res[i] = __builtin_sqrt(residual_squared(g, roots[i]));
        ord[i] = i;
    }
    // stable insertion sort by residual (std::sort of 5 elements)
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
    int n = 0;
    for (int i = 0; i < 5; ++i) {
        const int k = ord[i];
        if (res[k] * res[k] < 1.0e-12) out[n++] = Pt{roots[k].re, roots[k].im};
    }
    return n;
}

__device__ __forceinline__ void push_seed(Ctx& C, Pt p)
{
    if (C.n_seeds >= kSeedCap) {
        C.overflow = true;
        return;
    }
    wave_sync();
    wave_store(&C.seeds[C.n_seeds], p);
    C.n_seeds = uni(C.n_seeds + 1);
    wave_sync();
}

// Index of the first seed in [0, n) within 0.25 rho of `img`, or -1.
__device__ __forceinline__ int first_near_seed(const Ctx& C, Pt img, int n)
{
    const double lim = 0.0625 * C.r2;
    for (int base = 0; base < n; base += kWave) {
        const int k = base + C.lane;
        bool hit = false;
        if (k < n) hit = dist2(img, C.seeds[k]) < lim;
        const uint64_t m = ballot(hit);
        if (m) return base + first_set(m);
    }
    return -1;
}

// One probe's contribution, append_valid_probe_image_seeds
// (finite_source_magnifier.cpp:1449-1498), replayed uniformly from the images a
// lane found.  valid = probe inside the disk and more than three images.
__device__ void merge_probe(Ctx& C, int n_img, const Pt* img, const double* md)
{
    if (!step_ok(C)) return;
    const int n_before = C.n_seeds;
    const double lim = C.r2 * (1.0 + 1.0e-8);
    for (int i = 0; i < n_img; ++i) {
        if (md[i] > lim) continue;
        const int si = first_near_seed(C, img[i], n_before);
        if (si >= 0) {
            const double em = mapped_dist2(C.L, C.seeds[si].x, C.seeds[si].y, C.src);
            if (uni(md[i] + 1.0e-16 < em)) {
                wave_sync();  // every lane has read seeds[si] before it changes
                wave_store(&C.seeds[si], img[i]);
                wave_sync();
                C.seeds_changed = true;
            }
        } else {
            push_seed(C, img[i]);
            C.seeds_changed = true;
        }
    }
}

// Per-lane probe evaluation: inside-disk test, R3 images, mapped distances.
struct ProbeResult {
    int n;
    bool solved;  // the probe was inside and its quintic was solved
    Pt img[5];
    double md[5];
};

__device__ __forceinline__ ProbeResult eval_probe(const Ctx& C, Pt probe)
{
    ProbeResult r;
    r.n = 0;
    r.solved = false;
    if (!(C.rho > 0.0) || dist2(probe, C.src) >= C.r2 * (1.0 + 1.0e-10)) return r;
    Pt im[5];
    r.solved = true;
    const int n = probe_r3_images(C, probe, im);
    if (n <= 3) return r;
    r.n = n;
    for (int i = 0; i < n; ++i) {
        r.img[i] = im[i];
        r.md[i] = mapped_dist2(C.L, im[i].x, im[i].y, C.src);
    }
    return r;
}

// Replays one lane's probe result uniformly.
__device__ __forceinline__ void merge_lane(Ctx& C, const ProbeResult& mine, int src_lane)
{
    const int n = uni(shfl(mine.n, src_lane));
    if (n <= 3) return;
    Pt img[5];
    double md[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        img[i].x = uni(shfl(mine.img[i].x, src_lane));
        img[i].y = uni(shfl(mine.img[i].y, src_lane));
        md[i] = uni(shfl(mine.md[i], src_lane));
    }
    merge_probe(C, n, img, md);
}

// Probe k of a caustic point cs (append_caustic_probe_image_seeds,
// finite_source_magnifier.cpp:1500-1563): k < 8 along the unit vector (ux, uy)
// from the source centre, steps 0.02, 0.05, 0.15, 0.35 rho, outward then
// inward; k >= 8 the angular probes at radii 0.02, 0.05, 0.10, 0.20, 0.35 rho.
__device__ __forceinline__ Pt caustic_probe_point(const Ctx& C, Pt cs, double ux, double uy, int k,
                                                  const FsTables& tab)
{
    if (k < 8) {
        const int si = k >> 1;
        const double st = si == 0 ? 0.02 * C.rho : si == 1 ? 0.05 * C.rho : si == 2 ? 0.15 * C.rho
                                                                                   : 0.35 * C.rho;
        return (k & 1) ? Pt{cs.x - ux * st, cs.y - uy * st} : Pt{cs.x + ux * st, cs.y + uy * st};
    }
    const int a = k - 8;
    const int ri = a / kAngularProbes;
    const double st = ri == 0 ? 0.02 * C.rho : ri == 1 ? 0.05 * C.rho : ri == 2 ? 0.10 * C.rho
                    : ri == 3 ? 0.20 * C.rho : 0.35 * C.rho;
    const Pt t = tab.angular_trig[a % kAngularProbes];
    return Pt{cs.x + st * t.x, cs.y + st * t.y};
}

// Direction of the caustic probes of cs; false when cs is at the centre or
// not inside the disk (no probes).
__device__ __forceinline__ bool caustic_probe_axis(const Ctx& C, Pt cs, double* ux, double* uy)
{
    const double dx = cs.x - C.src.x;
    const double dy = cs.y - C.src.y;
    const double distance = hypot_glibc(dx, dy);
    if (distance <= 0.0 || distance >= C.rho) return false;
    *ux = dx / distance;
    *uy = dy / distance;
    return true;
}

// append_caustic_probe_image_seeds (finite_source_magnifier.cpp:1500-1563).
// With probe flags `fl` (a resumed scan) only the probes flagged as finding
// more than three images are solved again; the others cannot change the seeds.
__device__ void caustic_probes(Ctx& C, Pt cs, const FsTables& tab, const uint8_t* fl)
{
    double ux = 0.0, uy = 0.0;
    if (!uni(caustic_probe_axis(C, cs, &ux, &uy) ? 1 : 0)) return;
    {
        // 4 steps x 2 directions in the reference's order, one probe per lane
        ProbeResult mine;
        mine.n = 0;
        mine.solved = false;
        if (C.lane < 8 && (!fl || fl[C.lane])) mine = eval_probe(C, caustic_probe_point(C, cs, ux, uy, C.lane, tab));
        C.n_quintic = uni(C.n_quintic + popcount64(ballot(mine.solved)));
        for (int k = 0; k < 8; ++k) merge_lane(C, mine, k);
    }
    if (C.n_seeds <= 3) {
        for (int base = 0; base < 5 * kAngularProbes; base += kWave) {
            const int k = base + C.lane;
            ProbeResult mine;
            mine.n = 0;
            mine.solved = false;
            if (k < 5 * kAngularProbes && (!fl || fl[8 + k]))
                mine = eval_probe(C, caustic_probe_point(C, cs, ux, uy, 8 + k, tab));
            C.n_quintic = uni(C.n_quintic + popcount64(ballot(mine.solved)));
            const int cnt = (5 * kAngularProbes - base) < kWave ? (5 * kAngularProbes - base) : kWave;
            for (int j = 0; j < cnt; ++j) {
                merge_lane(C, mine, j);
                if (C.n_seeds > 3) return;
            }
        }
    }
}

// nearest_critical_source_distance2 (finite_source_magnifier.cpp:1643-1664) at
// an arbitrary phase; the phase's cos/sin come from the device libm.
__device__ double nearest_critical(const Ctx& C, double phase, Pt* nearest)
{
    double sn, cs;
    sincos(phase, &sn, &cs);
    cx roots[4];
    critical_roots(C.L, cx{cs, sn}, roots);
    double best = __builtin_inf();
    for (int r = 0; r < 4; ++r) {
        const Pt p = map_real(C.L, roots[r].re, roots[r].im);
        const double d2 = dist2(p, C.src);
        if (d2 < best) {
            best = d2;
            *nearest = p;
        }
    }
    return best;
}

// refine_nearest_critical_source (finite_source_magnifier.cpp:1666-1707)
__device__ Pt refine_critical(Ctx& C, double center, double step)
{
    double left = center - step;
    double right = center + step;
    const double golden = 0.61803398874989484820;
    double x1 = right - golden * (right - left);
    double x2 = left + golden * (right - left);
    Pt n1{0.0, 0.0}, n2{0.0, 0.0};
    C.n_quartic = uni(C.n_quartic + 34);
    double f1 = uni(nearest_critical(C, x1, &n1));
    double f2 = uni(nearest_critical(C, x2, &n2));
    for (int iter = 0; iter < 32; ++iter) {
        if (f1 > f2) {
            left = x1;
            x1 = x2;
            f1 = f2;
            n1 = n2;
            x2 = left + golden * (right - left);
            f2 = uni(nearest_critical(C, x2, &n2));
        } else {
            right = x2;
            x2 = x1;
            f2 = f1;
            n2 = n1;
            x1 = right - golden * (right - left);
            f1 = uni(nearest_critical(C, x1, &n1));
        }
    }
    return f1 < f2 ? n1 : n2;
}

// augmented_image_seeds (finite_source_magnifier.cpp:1709-1871) without seed
// hints, then augment_seeds_from_caustic_branches (:4845-4926).
__device__ void image_seeds(Ctx& C, const Pt* crit, const uint8_t* assign, const Pt* br,
                            const int* br_cell, const int* br_order, const CausticScalars& sc, int bins,
                            const FsTables& tab, const double* roots_ps)
{
    // Phase 0: binary_images (point_source_magnifier.cpp:723-778) of the source
    {
        const Geom g = make_geom(C.s, C.q, C.src.x, C.src.y);
        cx roots[5];
        double res[5];
        int ord[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            roots[i] = cx{roots_ps[2 * i], roots_ps[2 * i + 1]};
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
        if (C.hints) {
            // the LensModel path passes these images as seed hints, kept if
            // they map inside the disk (append_valid_seed, :1430-1447)
            for (int i = 0; i < physical; ++i) {
                const Pt z{roots[ord[i]].re, roots[ord[i]].im};
                if (!(mapped_dist2(C.L, z.x, z.y, C.src) > C.r2 * (1.0 + 1.0e-8))) push_seed(C, z);
            }
        }
        if (C.n_seeds == 0)
            for (int i = 0; i < physical; ++i) push_seed(C, Pt{roots[ord[i]].re, roots[ord[i]].im});
    }
    if (!(C.rho > 0.0) || C.last_phase < 1) return;
    const double phase_step = 2.0 * kPi / static_cast<double>(kSeedPhases);

    // Phase 1: strided scan (sample = j*40 + k; k outer, j inner), stopping at
    // the first phase whose caustic probes add a seed.
    bool found = false;
    double best_d2 = __builtin_inf();
    int best_ord = 0x7fffffff;
    constexpr int kOrd = kScanPositions;  // scan positions: (k*35 + j)*4 + root
    int probed = 0;                        // inside positions probed so far
    for (int base = 0; base < kOrd && !found; base += kWave) {
        const int o = base + C.lane;
        double d2 = __builtin_inf();
        bool inside = false;
        if (o < kOrd) {
            const int ph = o >> 2;
            const int r = o & 3;
            const int kphi = ph / 35;
            const int jphi = ph - kphi * 35;
            const int sample = jphi * 40 + kphi;
            d2 = dist2(crit[4 * sample + r], C.src);
            // positions before a resumed scan's start were probed without effect
            inside = d2 < C.r2 && d2 > 0.0 && o >= C.p1_start;
        }
        // running best in scan order (ties keep the earliest)
        {
            double v = d2;
            int idx = o < kOrd ? o : 0x7fffffff;
            wave_argmin(v, idx);
            v = uni(v);
            idx = uni(idx);
            if (v < best_d2) {
                best_d2 = v;
                best_ord = idx;
            }
        }
        uint64_t m = ballot(inside);
        // positions whose probes can change the seeds (all of them, unless the
        // probe farm flagged this search's probes)
        uint64_t live = m;
        if (C.farm) {
            bool any = false;
            if (inside) {
                const uint64_t* f8 = reinterpret_cast<const uint64_t*>(
                    C.farm + static_cast<size_t>(o) * kProbesPerPosition);
                uint64_t acc = 0;
#pragma unroll
                for (int w = 0; w < kProbesPerPosition / 8; ++w) acc |= f8[w];
                any = acc != 0;
            }
            live = ballot(any);
        }
        int cur_phase = -1;
        bool phase_found = false;
        while (m) {
            const int b = first_set(m);
            m &= m - 1;
            const int o2 = base + b;
            const int ph = o2 >> 2;
            if (ph != cur_phase) {
                if (cur_phase >= 0 && phase_found) break;
                cur_phase = ph;
                phase_found = false;
            }
            if (C.defer_budget >= 0 && probed >= C.defer_budget && !C.seeds_changed) {
                // a long scan with no effect so far: hand the positions from
                // o2 on to the probe farm and resume later
                const int h = wave_fetch_inc(C.heavy_count);
                if (h < kHeavyCap) {
                    C.defer_at = o2;
                    C.defer_slot = h;
                    return;
                }
                C.defer_budget = -1;
            }
            ++probed;
            const int before = C.n_seeds;
            if ((live >> b) & 1ull) {
                const int kphi = ph / 35;
                const int jphi = ph - kphi * 35;
                const int sample = jphi * 40 + kphi;
                caustic_probes(C, crit[4 * sample + (o2 & 3)], tab,
                               C.farm ? C.farm + static_cast<size_t>(o2) * kProbesPerPosition : nullptr);
            }
            phase_found = uni(C.n_seeds > before);
        }
        if (cur_phase >= 0 && phase_found) {
            found = true;
            // the scan stops here; best_d2 is not used once a crossing is found
        }
    }
    if (!found && best_d2 < C.r2 && best_d2 > 0.0) {
        const int ph = best_ord >> 2;
        const int kphi = ph / 35;
        const int jphi = ph - kphi * 35;
        const int sample = jphi * 40 + kphi;
        const double best_phase = phase_step * static_cast<double>(sample);
        const Pt cs = refine_critical(C, best_phase, phase_step);
        caustic_probes(C, cs, tab, nullptr);
    }

    if (C.last_phase < 2) return;
    // Phase 2: branch-tracked scan, probes on outside->inside transitions and
    // every 20th inside sample for large sources.  A branch the greedy matching
    // left unassigned (255) is outside (:1843-1846).  The reference starts the
    // matching from positions at (1e30, 1e30) with a best distance of 1e60, and
    // every squared distance to them is 2e60, so no branch is ever assigned and
    // the phase adds nothing; the whole scan is skipped when phase 0 has no
    // assignment, since the unmatched positions then never change.
    const bool any_assigned = uni(static_cast<int>(assign[0] & assign[1] & assign[2] & assign[3]) != 255);
    if (C.n_seeds >= 5 && any_assigned) {
        const bool arc = C.rho >= 2.0e-2;
        for (int base = 0; base < kOrd; base += kWave) {
            const int o = base + C.lane;
            bool fire = false;
            if (o < kOrd) {
                const int sample = o >> 2;
                const int bi = o & 3;
                const int j = assign[4 * sample + bi];
                bool now = false;
                if (j != 255) {
                    const double d2 = dist2(crit[4 * sample + j], C.src);
                    now = d2 < C.r2 && d2 > 0.0;
                }
                bool was = false;
                if (sample > 0) {
                    const int jp = assign[4 * (sample - 1) + bi];
                    if (jp != 255) {
                        const double dp = dist2(crit[4 * (sample - 1) + jp], C.src);
                        was = dp < C.r2 && dp > 0.0;
                    }
                }
                fire = (now && !was) || (arc && now && (sample % 20 == 0));
            }
            uint64_t m = ballot(fire);
            while (m) {
                const int b = first_set(m);
                m &= m - 1;
                const int o2 = base + b;
                const int sample = o2 >> 2;
                const int j = assign[4 * sample + (o2 & 3)];
                caustic_probes(C, crit[4 * sample + j], tab, nullptr);
            }
        }
    }

    if (C.last_phase < 3) return;
    // boundary probes at 0.98 rho (:1565-1591), stopping at 128 seeds
    if (C.n_seeds < 128) {
        const double pr = C.rho * (1.0 - 0.02);
        for (int base = 0; base < kBoundaryProbes && C.n_seeds < 128; base += kWave) {
            const int k = base + C.lane;
            ProbeResult mine;
            mine.n = 0;
            mine.solved = false;
            if (k < kBoundaryProbes) {
                const Pt t = tab.boundary_trig[k];
                mine = eval_probe(C, Pt{C.src.x + pr * t.x, C.src.y + pr * t.y});
            }
            C.n_quintic = uni(C.n_quintic + popcount64(ballot(mine.solved)));
            const int cnt = (kBoundaryProbes - base) < kWave ? (kBoundaryProbes - base) : kWave;
            for (int j = 0; j < cnt && C.n_seeds < 128; ++j) merge_lane(C, mine, j);
        }
    }
    if (C.last_phase < 4) return;
    // interior probes (:1593-1620) for rho >= 0.02
    if (C.rho >= 2.0e-2) {
        const double fr[3] = {0.25, 0.5, 0.75};
        for (int ri = 0; ri < 3; ++ri) {
            const double pr = C.rho * fr[ri];
            const Pt t = tab.interior_trig[C.lane];
            const ProbeResult mine = eval_probe(C, Pt{C.src.x + pr * t.x, C.src.y + pr * t.y});
            C.n_quintic = uni(C.n_quintic + popcount64(ballot(mine.solved)));
            for (int j = 0; j < kInteriorProbes; ++j) merge_lane(C, mine, j);
        }
    }

    if (C.last_phase < 5) return;
    // Phase 3 (:4845-4926), only with fewer than five seeds: probe past the
    // nearest caustic segment; a longer image list replaces the seeds.
    if (C.n_seeds < 5 && sc.n >= 2) {
        const double seg_search = C.rho + sc.max_seg_len;
        const int ix0 = iclamp(static_cast<int>((C.src.x - seg_search - sc.min_x) / sc.step_x), 0, kGrid - 1);
        const int ix1 = iclamp(static_cast<int>((C.src.x + seg_search - sc.min_x) / sc.step_x), 0, kGrid - 1);
        const int iy0 = iclamp(static_cast<int>((C.src.y - seg_search - sc.min_y) / sc.step_y), 0, kGrid - 1);
        const int iy1 = iclamp(static_cast<int>((C.src.y + seg_search - sc.min_y) / sc.step_y), 0, kGrid - 1);
        const int n = sc.n;
        const int nk = 4 * bins;
        // Candidates are visited in (cell, branch, position) order, the order
        // of br_order; the seed count is tested only when the cell changes, as
        // in the reference.  Each lane takes one candidate of a 64-candidate
        // chunk and solves its probe; the results are then applied in order,
        // and the solves past a stop are discarded.
        int last_key = -1;
        bool stop = false;
        for (int base = 0; base < nk && !stop; base += kWave) {
            const int i = base + C.lane;
            bool cand = false;
            bool probe_ok = false;
            int cell = -1;
            Pt probe{0.0, 0.0};
            if (i < nk) {
                const int k = br_order[i];
                cell = br_cell[k];
                const int iy = cell / kGrid;
                const int ix = cell - iy * kGrid;
                const int b = k / bins;
                const int pos = k - b * bins;
                const int next = (pos + 1) % n;
                if (ix >= ix0 && ix <= ix1 && iy >= iy0 && iy <= iy1 && next != 0) {
                    const Pt* bb = br + static_cast<size_t>(b) * bins;
                    const Pt p0 = bb[pos], p1 = bb[next];
                    if (point_segment_distance(C.src, p0, p1) < C.rho) {
                        cand = true;
                        const double seg_dx = p1.x - p0.x;
                        const double seg_dy = p1.y - p0.y;
                        const double seg_len2 = seg_dx * seg_dx + seg_dy * seg_dy;
                        const double t = seg_len2 > 0.0
                                             ? dclamp(((C.src.x - p0.x) * seg_dx + (C.src.y - p0.y) * seg_dy) / seg_len2,
                                                      0.0, 1.0)
                                             : 0.0;
                        const Pt nearest{p0.x + t * seg_dx, p0.y + t * seg_dy};
                        const double distance = __builtin_sqrt(dist2(nearest, C.src));
                        if (distance > 0.0 && distance < C.rho) {
                            const double step = C.rho * 0.05 / distance;
                            probe = Pt{nearest.x + (nearest.x - C.src.x) * step,
                                       nearest.y + (nearest.y - C.src.y) * step};
                            probe_ok = dist2(probe, C.src) < C.rho * C.rho;
                        }
                    }
                }
            }
            uint64_t cm = ballot(cand);
            if (!cm) continue;
            Pt im[5] = {};
            int ni = 0;
            if (probe_ok) ni = probe_r3_images(C, probe, im);
            while (cm) {
                const int j = first_set(cm);
                cm &= cm - 1;
                const int cj = uni(shfl(cell, j));
                if (cj != last_key) {
                    if (C.n_seeds >= 5) {
                        stop = true;
                        break;
                    }
                    last_key = cj;
                }
                if (uni(shfl(probe_ok ? 1 : 0, j)) == 0) continue;
                C.n_quintic = uni(C.n_quintic + 1);
                const int nij = uni(shfl(ni, j));
                if (nij > C.n_seeds) {
                    Pt imj[5];
#pragma unroll
                    for (int q = 0; q < 5; ++q) {
                        imj[q].x = uni(shfl(im[q].x, j));
                        imj[q].y = uni(shfl(im[q].y, j));
                    }
                    wave_sync();
                    for (int q = 0; q < nij; ++q) wave_store(&C.seeds[q], imj[q]);
                    C.n_seeds = nij;
                    wave_sync();
                }
            }
        }
    }
}

// ------------------------------------------------------------ Cartesian scan

// Lane k receives x0 after k sequential additions of `step` (the reference
// advances lattice cells with `x += incr`, finite_source_magnifier.cpp:2004-2010).
// Every lane from a verified base lane b takes base + (k - b) d with
// d = (base + step) - base, the increment that repeats while the additions
// stay in one binade; each lane then checks its value against the previous
// lane's value plus step.  Values up to the first failing lane are exact by
// induction from the base, so that lane becomes the next base, holding the
// previous lane's value plus step.  The loop runs once per binade crossing
// (at most 64 times) and yields exactly the sequential sums.
__device__ __forceinline__ double seq_lanes(double x0, double step)
{
    const int lane = lane_id();
    double base = x0;
    int b = 0;
    double cand = x0;
    for (int it = 0; it < kWave; ++it) {
        const double d = (base + step) - base;
        if (lane == b)
            cand = base;
        else if (lane > b)
            cand = base + static_cast<double>(lane - b) * d;
        const double prev = lane_prev(cand);
        const uint64_t bad = ballot(lane > b && !(cand == prev + step));
        if (!bad) break;
        const int f = first_set(bad);
        base = read_lane(cand, f - 1) + step;
        b = f;
    }
    return cand;
}

// Two chains in one wave: lane k of [0, a) receives xa after k sequential
// additions of sa, lane k of [a, 64) receives xb after k - a sequential
// additions of sb; each chain is verified as in seq_lanes.  0 < a < 64.
__device__ __forceinline__ double seq_lanes2(double xa, double sa, double xb, double sb, int a)
{
    const int lane = lane_id();
    const bool hi = lane >= a;
    const double step = hi ? sb : sa;
    const uint64_t lo_mask = (1ull << a) - 1ull;
    int b0 = 0, b1 = a;
    double v0 = xa, v1 = xb;
    double cand = hi ? xb : xa;
    for (int it = 0; it < kWave; ++it) {
        const int b = hi ? b1 : b0;
        const double base = hi ? v1 : v0;
        const double d = (base + step) - base;
        if (lane == b)
            cand = base;
        else if (lane > b)
            cand = base + static_cast<double>(lane - b) * d;
        const double prev = lane_prev(cand);
        const uint64_t bad = ballot(lane > b && !(cand == prev + step));
        if (!bad) break;
        const uint64_t bad0 = bad & lo_mask;
        const uint64_t bad1 = bad & ~lo_mask;
        if (bad0) {
            const int f = first_set(bad0);
            v0 = read_lane(cand, f - 1) + sa;
            b0 = f;
        }
        if (bad1) {
            const int f = first_set(bad1);
            v1 = read_lane(cand, f - 1) + sb;
            b1 = f;
        }
    }
    return cand;
}

// The reference keeps, per seed, an unordered_map from row key to the rows
// pushed with that key (finite_source_magnifier.cpp:917, 1989-1999).  Here the
// map is a chained hash: head[key & (kRowHash-1)] -> newest node, node.next ->
// older.  A node also holds the row's final xmin/xmax (a pushed row's record
// never changes afterwards), so a duplicate check reads one node per step.
// `dirty` (in memory, so it survives an abandoned epoch) counts the keys whose
// heads are set; they are cleared before the next seed's scan.
struct RowNode {
    int key, row, next, pad;
    double xmin, xmax;
};

struct Rows {
    double* r;     // [row*5 + field]
    RowNode* node; // [row_cap] pushed rows in push order
    int* head;     // [kRowHash]
    int* dirty;    // keys pushed since the heads were last cleared (<0: none)
    int size;      // rows allocated (ensure)
    int pushed;
};

__device__ __forceinline__ void rows_reset(Ctx& C, Rows& R)
{
    wave_sync();
    const int nd = uni(*R.dirty);
    for (int k = C.lane; k < nd; k += kWave) R.head[R.node[k].key & (kRowHash - 1)] = -1;
    wave_sync();
    wave_store(R.dirty, 0);
    wave_sync();
    R.size = 0;
    R.pushed = 0;
}

enum { XMIN = 0, XMAX = 1, AX = 2, YY = 3, DYS = 4 };

struct Cart {
    double incr, inv_incr;
    int jac_sign;
    long long max_steps;
};

__device__ __forceinline__ bool rows_ensure(Ctx& C, Rows& R, int index)
{
    if (index < R.size) return true;
    if (index >= C.row_cap) {
        C.overflow = true;
        C.cap_overflow = true;
        return false;
    }
    const int from = R.size * 5;
    const int to = (index + 1) * 5;
    wave_sync();
    for (int k = from + C.lane; k < to; k += kWave) R.r[k] = 0.0;
    R.size = index + 1;
    wave_sync();
    return true;
}

// Stores a row field; the barriers order it after every lane's earlier reads of
// the row records and before any later read.
__device__ __forceinline__ void rset(Ctx& C, Rows& R, int row, int f, double v)
{
    wave_sync();
    wave_store(&R.r[row * 5 + f], v);
    wave_sync();
}

__device__ __forceinline__ double rget(const Rows& R, int row, int f) { return R.r[row * 5 + f]; }

// One lattice cell: inside flag and brightness weight
// (finite_source_magnifier.cpp:1922-1958).
__device__ __forceinline__ bool cart_cell(const Ctx& C, const Cart& K, double x, double y, double* w)
{
    double md, dz2;
    if (K.jac_sign == 0) {
        md = mapped_dist2(C.L, x, y, C.src);
        dz2 = md;
    } else {
        double jac;
        md = mapped_dist2_jac(C.L, x, y, C.src, &jac);
        const int es = sign_of(jac);
        dz2 = (es != -K.jac_sign) ? md : C.r2 + 1.0;
    }
    const bool in = dz2 <= C.r2;
    *w = in ? (C.uniform ? 1.0 : ld_brightness(md * C.inv_r2, C.ld_c, C.ld_d)) : 0.0;
    return in;
}

// ------------------------------------------------------ local refinement tape
//
// LocalRefinementTape (finite_source_magnifier.cpp:977-1004): while the
// adaptive mode scans an epoch, every evaluated cell whose mapped distance lies
// within the boundary band |d^2 - rho^2| <= max(4 rho incr, 4 incr^2) of the
// source edge (and, for guarded cells, on the seed's side of the critical
// curve) is recorded with its weight (:1952-1961), in scan order, once per
// lattice key (llround(x / incr), llround(y / incr)), up to `cap` cells.  Once
// full the tape takes nothing more, so the scans record only until then.
struct Tape {
    double* x;
    double* y;
    double* w;
    long long* kx;
    long long* ky;
    int n, cap;
    double band2, step;
};

__device__ __forceinline__ bool tape_open(const Tape* T) { return T && T->n < T->cap; }

// Appends the cells of the lanes in `rec` in lane order (the scan's visiting
// order within one step), each unless its key is already on the tape.
__device__ void tape_add(Ctx& C, Tape& T, bool rec, double x, double y, double w)
{
    uint64_t m = ballot(rec);
    while (m && T.n < T.cap) {
        const int j = first_set(m);
        m &= m - 1ull;
        const double cx = read_lane(x, j);
        const double cw = read_lane(w, j);
        const long long kx = llround(cx / T.step);
        const long long ky = llround(y / T.step);
        bool dup = false;
        for (int base = 0; base < T.n && !dup; base += kWave) {
            const int i = base + C.lane;
            dup = ballot(i < T.n && T.kx[i] == kx && T.ky[i] == ky) != 0;
        }
        if (dup) continue;
        wave_sync();
        wave_store(&T.x[T.n], cx);
        wave_store(&T.y[T.n], y);
        wave_store(&T.w[T.n], cw);
        wave_store(&T.kx[T.n], kx);
        wave_store(&T.ky[T.n], ky);
        wave_sync();
        T.n = uni(T.n + 1);
    }
}

// cart_cell, also returning the mapped distance and whether the cell may enter
// the tape as an outside cell (jac_ok of :1927-1937).
__device__ __forceinline__ bool cart_cell_md(const Ctx& C, const Cart& K, double x, double y, double* w,
                                             double* md_out, bool* jok_out)
{
    double md, dz2;
    bool jok = true;
    if (K.jac_sign == 0) {
        md = mapped_dist2(C.L, x, y, C.src);
        dz2 = md;
    } else {
        double jac;
        md = mapped_dist2_jac(C.L, x, y, C.src, &jac);
        jok = sign_of(jac) != -K.jac_sign;
        dz2 = jok ? md : C.r2 + 1.0;
    }
    const bool in = dz2 <= C.r2;
    *w = in ? (C.uniform ? 1.0 : ld_brightness(md * C.inv_r2, C.ld_c, C.ld_d)) : 0.0;
    *md_out = md;
    *jok_out = jok;
    return in;
}

// cartesian_image_area_impl (finite_source_magnifier.cpp:1892-2014).  The
// scan keeps the current row's xmin/xmax and the previous row's xmin in
// registers (written through to the row records), and each lane keeps its
// share of the row's brightness sum, reduced once when the row closes; a
// row's sum is nonzero exactly when one of its counted weights is, which is
// what the scan's decisions test.
//
// A row is a right scan from x0 followed by a left scan from x0 - incr.  The
// first step of a row evaluates both at once: `split` lanes continue the right
// scan and the other lanes hold the first cells of the left scan, which are
// used once the right scan has ended (a cell's value does not depend on the
// scan state; only the decisions taken on it do, and they are taken in the
// reference's order).  `split` follows the right-scan length of the previous
// row.
//
// `watch` (LDS, or null): the scan stops with C.overflow set at the first row
// close that finds it no longer equal to `expect`.
//
// The scan counts its cells (guard), wave steps and rows in locals; cart_scan
// adds them to the epoch's counts when it returns.  `room` is the epoch's
// remaining step budget.
//
// With kTape, the cells each step visits in the reference's order (up to and
// including the cell that ends the run) go through the tape while it is open.
template <bool kTape>
__device__ __forceinline__ double cart_scan_rows(Ctx& C, const Cart& K, Rows& R, Pt seed, double dy, int& yi,
                                                 double prev_xmin_in, const int* watch, int expect,
                                                 long long& guard, int& steps, int& rows, int room, Tape* T)
{
    const double incr = K.incr;
    double acc = 0.0;     // this lane's share of the row's brightness sum (countx)
    bool any_w = false;   // countx != 0
    double countall = 0.0;
    bool last_in = 99999999.9 <= C.r2;
    double x = seed.x;    // first cell of the right scan
    double y = seed.y;
    double x0 = seed.x;
    const long long budget = K.max_steps - 1;
    int split = kWave / 2;
    if (!rows_ensure(C, R, yi)) return __builtin_nan("");
    // a row's xmin is set once its right scan ends, and its xmax before any
    // read whenever the row counts (the right scan sets it at its end unless
    // its first cell is out, and the left scan then sets it at its first
    // inside cell); a row that counts nothing returns before reading either.
    // So the record's old values are never needed: the scan starts them at 0.
    double cur_xmin = 0.0;
    double cur_xmax = 0.0;
    double prev_xmin = prev_xmin_in;  // row yi - 1's xmin (0 for yi = 0)
    for (;;) {
        // the row's key and its chain head, loaded now for the duplicate check at
        // the row's close (only this row's own push changes the heads)
        const int row_key = static_cast<int>(y * K.inv_incr + 2000000);
        const int head_pf = R.head[row_key & (kRowHash - 1)];
        // ---------------------------------------------------------- right scan
        int stash_n = 0;      // left-scan cells held in lanes [64 - stash_n, 64)
        double sx = 0.0, sw = 0.0;
        bool s_in = false;
        double s_md = 0.0;    // their mapped distances and guard results (tape)
        bool s_jok = false;
        int right_cells = 0;
        for (bool first = true;; first = false) {
            if (++steps > room) {
                C.overflow = true;
                return __builtin_nan("");
            }
            const long long remaining = budget - guard;
            if (remaining <= 0) return countall;
            const bool combined = first && remaining >= kWave;
            const int n_eval = combined ? split : (remaining < kWave ? static_cast<int>(remaining) : kWave);
            const double xk = combined ? seq_lanes2(x, incr, x0 + (-incr), -incr, split) : seq_lanes(x, incr);
            double w = 0.0;
            bool in = false;
            double md = 0.0;
            bool jok = false;
            if constexpr (kTape) {
                if (combined || C.lane < n_eval) in = cart_cell_md(C, K, xk, y, &w, &md, &jok);
            } else {
                if (combined || C.lane < n_eval) in = cart_cell(C, K, xk, y, &w);
            }
            if (combined) {
                stash_n = kWave - split;
                sx = xk;
                sw = w;
                s_in = in;
                if constexpr (kTape) {
                    s_md = md;
                    s_jok = jok;
                }
            }
            const bool valid = C.lane < n_eval;
            const int s = first_set(ballot(!in && valid));
            if constexpr (kTape) {
                if (tape_open(T)) {
                    const bool vis = valid && (s < 0 || C.lane <= s);
                    tape_add(C, *T, vis && jok && fabs(md - C.r2) <= T->band2, xk, y, w);
                }
            }
            if (s < 0) {
                const double wv = valid ? w : 0.0;
                acc += wv;
                any_w = any_w || ballot(wv != 0.0) != 0;
                guard += n_eval;
                right_cells += n_eval;
                last_in = true;
                x = read_lane(xk, n_eval - 1) + incr;
                continue;
            }
            const double wl = C.lane < s ? w : 0.0;
            acc += wl;
            any_w = any_w || ballot(wl != 0.0) != 0;
            guard += s + 1;
            right_cells += s + 1;
            const double xs = read_lane(xk, s);
            const bool prev_in = s > 0 ? true : last_in;
            if (prev_in) {
                cur_xmax = xs;
                rset(C, R, yi, XMAX, xs);
            }
            break;
        }
        x = x0;
        cur_xmin = x + (-incr);
        rset(C, R, yi, XMIN, cur_xmin);
        last_in = false;
        x = x + (-incr);
        // ----------------------------------------------------------- left scan
        const double thr = yi != 0 ? prev_xmin - (-incr) : 0.0;
        bool use_stash = stash_n > 0 && budget - guard >= stash_n;
        for (;;) {
            if (++steps > room) {
                C.overflow = true;
                return __builtin_nan("");
            }
            const long long remaining = budget - guard;
            if (remaining <= 0) return countall;
            int lo, n_eval;
            double xk, w;
            bool in;
            double md = 0.0;
            bool jok = false;
            if (use_stash) {
                lo = kWave - stash_n;
                n_eval = stash_n;
                xk = sx;
                w = sw;
                in = s_in;
                if constexpr (kTape) {
                    md = s_md;
                    jok = s_jok;
                }
                use_stash = false;
            } else {
                lo = 0;
                n_eval = remaining < kWave ? static_cast<int>(remaining) : kWave;
                xk = seq_lanes(x, -incr);
                w = 0.0;
                in = false;
                if constexpr (kTape) {
                    if (C.lane < n_eval) in = cart_cell_md(C, K, xk, y, &w, &md, &jok);
                } else {
                    if (C.lane < n_eval) in = cart_cell(C, K, xk, y, &w);
                }
            }
            const int end = lo + n_eval;
            const bool valid = C.lane >= lo && C.lane < end;
            const bool inv = in && valid;
            const bool cb = any_w;
            const uint64_t in_m = ballot(inv);
            // the row's sum becomes nonzero after lane g (-1: not in this step)
            const int g = first_set(ballot(inv && w != 0.0));
            const bool cond = (yi != 0) && (xk >= thr);
            const bool stop = valid && !in && (cb || (g >= 0 && C.lane > g) || yi == 0 || !cond);
            const int s = first_set(ballot(stop));
            if constexpr (kTape) {
                if (tape_open(T)) {
                    const bool vis = valid && (s < 0 || C.lane <= s);
                    tape_add(C, *T, vis && jok && fabs(md - C.r2) <= T->band2, xk, y, w);
                }
            }
            const int limit = s >= 0 ? s : end;
            const double wl = (C.lane < limit && inv) ? w : 0.0;
            acc += wl;
            if (!cb) {
                // every in-cell met while the sum is zero sets xmax = x + incr:
                // the last is the first nonzero-weight cell, else the last in-cell
                int j = -1;
                if (g >= 0 && g < limit) {
                    j = g;
                } else {
                    const uint64_t mm = in_m & (limit >= kWave ? ~0ull : ((1ull << limit) - 1ull));
                    if (mm) j = 63 - __builtin_clzll(mm);
                }
                if (j >= 0) {
                    cur_xmax = read_lane(xk, j) - (-incr);
                    rset(C, R, yi, XMAX, cur_xmax);
                }
            }
            any_w = any_w || ballot(wl != 0.0) != 0;
            if (s < 0) {
                guard += n_eval;
                last_in = ((in_m >> (end - 1)) & 1ull) != 0;
                x = read_lane(xk, end - 1) + (-incr);
                continue;
            }
            guard += s - lo + 1;
            const double xs = read_lane(xk, s);
            const bool prev_in = s > lo ? (((in_m >> (s - 1)) & 1ull) != 0) : last_in;
            if (prev_in) {
                cur_xmin = xs;
                rset(C, R, yi, XMIN, xs);
            }
            break;
        }
        // ------------------------------------------------------- close the row
        if (watch && uni(__hip_atomic_load(watch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) != expect) {
            C.overflow = true;
            return __builtin_nan("");
        }
        ++rows;
        const double countx = uni(wave_sum(acc));
        countall += countx;
        rset(C, R, yi, AX, countx);
        rset(C, R, yi, YY, y);
        rset(C, R, yi, DYS, countx == 0.0 ? -dy : dy);
        if (countx == 0.0) return countall;
        const int key = row_key;
        const int newest = uni(head_pf);
        {
            bool dup = false;
            for (int p = newest; p >= 0 && !dup;) {
                const RowNode nd = R.node[p];
                if (uni(nd.key) == key)
                    dup = cur_xmin + incr < uni(nd.xmax) && cur_xmax - incr > uni(nd.xmin);
                p = uni(nd.next);
            }
            if (dup) return countall - countx;
        }
        if (R.pushed >= C.row_cap) {
            C.overflow = true;
            C.cap_overflow = true;
            return __builtin_nan("");
        }
        wave_sync();
        {
            RowNode nd;
            nd.key = key;
            nd.row = yi;
            nd.next = newest;
            nd.pad = 0;
            nd.xmin = cur_xmin;
            nd.xmax = cur_xmax;
            if (C.lane == 0) R.node[R.pushed] = nd;
            wave_store(&R.head[key & (kRowHash - 1)], R.pushed);
            wave_store(R.dirty, R.pushed + 1);
        }
        R.pushed = uni(R.pushed + 1);
        wave_sync();
        ++yi;
        const bool fresh = yi >= R.size;  // a new record is all zero
        if (!rows_ensure(C, R, yi)) return __builtin_nan("");
        prev_xmin = cur_xmin;
        x0 = cur_xmax;
        (void)fresh;
        cur_xmin = 0.0;
        cur_xmax = 0.0;
        x = x0 - incr;
        y = y + dy;
        acc = 0.0;
        any_w = false;
        last_in = false;
        x = x + incr;
        split = iclamp(right_cells + 4, 8, 32);
    }
}

template <bool kTape>
__device__ __forceinline__ double cart_scan(Ctx& C, const Cart& K, Rows& R, Pt seed, double dy, int& yi,
                                            double prev_xmin, const int* watch, int expect, Tape* T)
{
    long long guard = 0;
    int steps = 0, rows = 0;
    const long long left = C.budget - C.steps;
    const int room = left < 0 ? 0 : static_cast<int>(left);
    const double r = cart_scan_rows<kTape>(C, K, R, seed, dy, yi, prev_xmin, watch, expect, guard, steps, rows,
                                           room, T);
    C.steps = uni(C.steps + steps);
    if (K.jac_sign == 0)
        C.n_cart = uni(C.n_cart + guard);
    else
        C.n_cart_guarded = uni(C.n_cart_guarded + guard);
    C.n_rows = uni(C.n_rows + rows);
    return r;
}

// ----------------------------------------------------------- Cartesian images
//
// inverse_ray_cartesian_binary_mag (finite_source_magnifier.cpp:2470-2768)
// processes the seeds in order: an unflagged seed gets its image scanned, then
// the overlap rule flags the later seeds that image covers and subtracts the
// earlier images it covers.  The scan of an image depends only on its seed
// (its rows start empty), so images can be computed anywhere, in any order,
// and accounted in seed order afterwards; the accounting reads the image's
// rows, its area and its work counts.  image_area() is the scan of one image,
// account_image() its accounting.

struct ImageRun {
    double area;
    int nyi;  // rows [0, nyi) take part in the overlap test
    int gaps;         // gap-repair scans started (diagnostics->gap_repairs, :2578-2644)
    double max_jump;  // largest |row edge step| / incr over counted rows (:2571-2573)
};

__device__ __forceinline__ void cart_params(Cart& K, const Ctx& C, int source_bins)
{
    const double nbin = static_cast<double>(source_bins > 1 ? source_bins : 1);
    K.incr = C.rho / nbin;
    K.inv_incr = 1.0 / K.incr;
    const long long sb = source_bins;
    K.max_steps = (100000LL > sb * sb * 2000LL) ? 100000LL : sb * sb * 2000LL;
    K.jac_sign = 0;
}

__device__ __forceinline__ Rows slot_rows(const CartScratch& scr, int slot)
{
    Rows R;
    R.r = scr.rows + static_cast<size_t>(slot) * scr.row_cap * 5;
    int* rk = scr.row_key + static_cast<size_t>(slot) * row_ints_per_slot(scr.row_cap);
    R.node = reinterpret_cast<RowNode*>(rk);
    R.head = rk + 8 * scr.row_cap;
    R.dirty = rk + 8 * scr.row_cap + kRowHash;
    R.size = 0;
    R.pushed = 0;
    return R;
}

// Team of kTeamWaves waves sharing one workgroup (k_ir_cart).  Each wave runs
// its own epochs; a wave whose epoch has run kOfferSteps wave steps offers the
// next unflagged seeds of that epoch as image jobs, and a teammate between
// epochs computes such an image in its own slot.  Job status words:
//   kJobFree          unused
//   kJobOpen          offered, not taken
//   kJobRun + h       wave h computes it
//   kJobDone + h      wave h has finished; its rows stay untouched until the
//                     owner sets the word back to kJobFree
//   kJobCancel + h    the owner no longer needs it (wave h stops at its next
//                     poll and does not wait)
// Only the owner takes a job back (Open -> Free), cancels or releases it; only
// a helper claims it (Open -> Run) and finishes it (Run -> Done).  A wave only
// ever waits for a job that another wave is computing, and a computation waits
// for nothing, so every wait ends; each wait also has a time limit.
enum : int { kJobFree = 0, kJobOpen = 1, kJobRun = 16, kJobDone = 32, kJobCancel = 48 };
constexpr long long kOfferSteps = 2048;
constexpr long long kTeamWaitTicks = 6000000000ll;  // 60 s of the 100 MHz clock

struct TeamOut {
    double area, max_jump;
    int nyi, bad;  // bad: 1 work budget or slot exceeded, 2 slot capacity
    int gaps;
    long long steps, n_cart, n_cart_guarded, n_rows;
};

struct Team {
    int active;                          // waves still taking queue items
    int ready;                           // wave 0 has started its first epoch
    int item[kTeamWaves][kTeamJobs];     // queue item (epoch) of each job
    int seed[kTeamWaves][kTeamJobs];     // its seed
    int status[kTeamWaves][kTeamJobs];
    TeamOut out[kTeamWaves];             // result of the job each helper finished
    // each wave's epoch-loop state while it scans an image (the scan does not
    // use it, so it is kept here rather than in registers)
    struct {
        double area;
        long long e, t_begin, idle_from;
        int item, nbin, first_item, count, reserved, bits;
    } park[kTeamWaves];
};

__device__ __forceinline__ int lds_get(const int* p)
{
    return uni(__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
}

__device__ __forceinline__ void lds_put(int* p, int v)
{
    __hip_atomic_store(p, uni(v), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Compare-and-swap on behalf of the wave: lane 0 swaps, the other lanes compare
// against a value no status word holds.
__device__ __forceinline__ bool lds_cas(int* p, int expected, int desired)
{
    int exp = lane_id() == 0 ? expected : -1;
    __hip_atomic_compare_exchange_strong(p, &exp, desired, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return __builtin_amdgcn_readfirstlane(exp == expected ? 1 : 0) != 0;
}

__device__ __forceinline__ void team_sleep() { __builtin_amdgcn_s_sleep(8); }

// The owner drops job k: an offered job is withdrawn, a running one cancelled,
// a finished one released.
__device__ __forceinline__ void team_drop(Team& T, int w, int k)
{
    int* st = &T.status[w][k];
    for (int tries = 0; tries < 8; ++tries) {
        const int s = lds_get(st);
        if (s == kJobFree || s >= kJobCancel) return;
        if (s == kJobOpen) {
            if (lds_cas(st, s, kJobFree)) return;
        } else if (s < kJobDone) {
            if (lds_cas(st, s, kJobCancel + (s - kJobRun))) return;
        } else {
            lds_put(st, kJobFree);
            return;
        }
    }
}

// Job of the owner's seed j that is still live (offered, running or done), or -1.
__device__ __forceinline__ int team_find(Team& T, int w, int j)
{
    for (int k = 0; k < kTeamJobs; ++k) {
        const int s = lds_get(&T.status[w][k]);
        if (s != kJobFree && s < kJobCancel && lds_get(&T.seed[w][k]) == j) return k;
    }
    return -1;
}

// Offers the owner's next unflagged seeds after `cur` in the free job entries.
__device__ __forceinline__ void team_offer(Team& T, int w, int item, const int* flags, int cur, int n_img,
                                           int& next_offer)
{
    if (next_offer <= cur) next_offer = cur + 1;
    for (int k = 0; k < kTeamJobs && next_offer < n_img; ++k) {
        const int s = lds_get(&T.status[w][k]);
        if (s != kJobFree && s < kJobCancel) continue;
        while (next_offer < n_img && (uni(flags[next_offer]) & 1)) ++next_offer;
        if (next_offer >= n_img) break;
        lds_put(&T.item[w][k], item);
        lds_put(&T.seed[w][k], next_offer);
        lds_put(&T.status[w][k], kJobOpen);
        ++next_offer;
    }
}

// Claims an offered job of a teammate: its owner p and job k.
__device__ __forceinline__ bool team_claim(Team& T, int w, int& p, int& k)
{
    for (int d = 1; d < kTeamWaves; ++d) {
        const int q = (w + d) % kTeamWaves;
        for (int j = 0; j < kTeamJobs; ++j) {
            if (lds_get(&T.status[q][j]) != kJobOpen) continue;
            if (lds_cas(&T.status[q][j], kJobOpen, kJobRun + w)) {
                p = q;
                k = j;
                return true;
            }
        }
    }
    return false;
}

// Checks between scans: a helper stops once its job is no longer its own; the
// owner of a long epoch offers its next images.
struct ImageHook {
    Team* team;
    int w;
    const int* watch;  // helper: its job's status word, which must stay `expect`
    int expect;
    const int* flags;  // owner: offering state
    int item, cur, n_img;
    int* next_offer;
    bool armed;
};

__device__ __forceinline__ bool hook_poll(Ctx& C, ImageHook& hk)
{
    if (hk.watch && lds_get(hk.watch) != hk.expect) {
        C.overflow = true;
        return false;
    }
    if (hk.armed && C.steps >= kOfferSteps) {
        team_offer(*hk.team, hk.w, hk.item, hk.flags, hk.cur, hk.n_img, *hk.next_offer);
        hk.armed = false;
    }
    return true;
}

// State of the gap-repair examination (image_area stage 2) that gap_batch
// advances.
struct GapState {
    int row, cs;
    double dxmax, dxmin;
    int yi, nyi, gaps;
    double max_jump;
};

// Several gap-repair scans per wave step.  The examination of rows
// row .. nyi - 2 (image_area stage 2, :2578-2644) lists the next scans: start
// cell, row and threshold come from the row records, which stay as they are
// while the scans find nothing.  The first row of each listed scan is then
// evaluated in one wave step, kc lanes per scan: its start cell and the one
// after it (the right scan) and the kc - 2 cells before it (the left scan),
// each lane forming its coordinate with the additions the one-scan path makes
// (start + incr, or start + (-incr) repeated).  The scans are resolved in order
// with the row rules of cart_scan_rows.  A scan whose first row holds no weight
// ends there, with no area, no row pushed and yi back where it was: it is
// committed here, with the records, counts and examination state that the
// one-scan path leaves behind.  The first scan that finds weight, runs past its
// lanes or meets a limit ends the batch and is left to the one-scan path.
// Returns the scans committed, or -1 with C.overflow set.
__device__ __forceinline__ int gap_batch(Ctx& C, const Cart& K, Rows& R, GapState& G, const int* watch, int expect)
{
    const double incr = K.incr;
    const double g = 1.1 * incr;
    const int last = G.nyi - 2;  // the last row is left to the one-scan path
    if (G.row < 0 || G.row > last) return 0;
    const int base_row = G.row;
    const int span_rows = 11;  // records of rows base_row .. base_row + 11 in lanes 0..59
    const int nrec = (base_row + span_rows <= G.nyi - 1 ? span_rows : G.nyi - 1 - base_row) + 1;
    double rec = 0.0;
    if (C.lane < 5 * nrec) rec = R.r[base_row * 5 + C.lane];
    auto fld = [&](int row, int f) { return uni(read_lane(rec, (row - base_row) * 5 + f)); };
    // the scans, lane j holding scan j's start, row y, threshold, dy, the
    // start record's xmin / xmax, and its (row, case)
    constexpr int kMaxScans = 8;
    double v_sx = 0.0, v_sy = 0.0, v_thr = 0.0, v_dy = 0.0, v_mn = 0.0, v_mx = 0.0;
    int v_rc = 0;
    int nd = 0;
    int r = base_row, k = G.cs;
    double dxmax = G.dxmax, dxmin = G.dxmin;
    double mj = G.max_jump;
    for (int it = 0; it < 64 && nd < kMaxScans; ++it) {
        if (k >= 4) {
            if (r + 1 > last || r + 2 - base_row >= nrec) break;
            ++r;
            dxmax = fld(r + 1, XMAX) - fld(r, XMAX);
            dxmin = fld(r + 1, XMIN) - fld(r, XMIN);
            k = fld(r + 1, AX) > 0.0 ? 0 : 4;
            if (k == 0) mj = dmax(mj, dmax(fabs(dxmax), fabs(dxmin)) / incr);
            continue;
        }
        const int kk = k++;
        bool have = false;
        double x = 0.0, y = 0.0, d = 0.0, a = 0.0, b = 0.0;
        if (kk == 0 && dxmax > g) {
            x = fld(r + 1, XMAX);
            y = fld(r, YY);
            a = fld(r, XMAX);
            b = fld(r + 1, XMAX);
            d = -fld(r, DYS);
            have = true;
        } else if (kk == 1 && dxmin > g) {
            x = fld(r + 1, XMIN) - incr;
            y = fld(r + 1, YY);
            a = fld(r, XMIN);
            b = fld(r + 1, XMIN);
            d = fld(r, DYS);
            have = true;
        } else if (kk == 2 && dxmin < -g) {
            x = fld(r, XMIN) - incr;
            y = fld(r, YY);
            a = fld(r + 1, XMIN);
            b = fld(r, XMIN);
            d = -fld(r, DYS);
            have = true;
        } else if (kk == 3 && dxmax < -g) {
            x = fld(r + 1, XMAX) + incr;
            y = fld(r + 1, YY);
            a = fld(r + 1, XMAX);
            b = fld(r, XMAX);
            d = fld(r, DYS);
            have = true;
        }
        if (!have) continue;
        const bool mine = C.lane == nd;
        v_sx = mine ? x : v_sx;
        v_sy = mine ? y : v_sy;
        v_thr = mine ? a - (-incr) : v_thr;  // thr = prev_xmin - (-incr), prev_xmin = the start record's xmin
        v_dy = mine ? d : v_dy;
        v_mn = mine ? a : v_mn;
        v_mx = mine ? b : v_mx;
        v_rc = mine ? (r * 4 + kk) : v_rc;
        ++nd;
    }
    if (nd == 0) return 0;
    // lanes per scan: the widest left scan (start down to the threshold) of
    // the listed scans, plus the right scan's two cells
    double span_d = 0.0;
    if (C.lane < nd) span_d = (v_sx - v_thr) * K.inv_incr;
    const double span_max = uni(wave_max(C.lane < nd ? span_d : 0.0));
    const int need = static_cast<int>(span_max) + 4;
    const int kc = need <= 8 ? 8 : (need <= 16 ? 16 : 32);
    const int ns = nd < kWave / kc ? nd : kWave / kc;
    // each lane's cell: c >= 0 the right scan's start + c incr, c < 0 the left
    // scan's start + (-incr) |c| times
    const int my_j = C.lane / kc;
    const int my_t = C.lane - my_j * kc;
    const int c = my_t - (kc - 2);
    const double sx = shfl(v_sx, my_j);
    const double sy = shfl(v_sy, my_j);
    double xk = sx;
    if (c > 0) xk = xk + incr;
    for (int i = 0; i < -c; ++i) xk = xk + (-incr);
    const double thr = shfl(v_thr, my_j);
    double w = 0.0;
    bool in = false;
    const bool live = my_j < ns;
    if (live) in = cart_cell(C, K, xk, sy, &w);
    const uint64_t inm = ballot(in);
    const uint64_t nzm = ballot(in && w != 0.0);
    const uint64_t belowm = ballot(live && !(xk >= thr));  // left of the threshold (cond false)
    const uint64_t kmask = kc >= 64 ? ~0ull : ((1ull << kc) - 1ull);
    // resolve in order
    int n = 0;
    long long visited = 0;
    int xmn_lane = -1;               // the last scan's xmin cell
    int xmx_lane = -1, xmx_plus = 0;  // the last xmax a scan set: its cell, and whether it is cell + incr
    const int p = kc - 2;  // the start's lane within a scan
    for (int j = 0; j < ns; ++j) {
        const uint64_t bits = (inm >> (j * kc)) & kmask;
        const uint64_t nzb = (nzm >> (j * kc)) & kmask;
        const uint64_t blw = (belowm >> (j * kc)) & kmask;
        // right scan: p, p + 1
        const uint64_t outs = ~bits & kmask & ~((1ull << p) - 1ull);
        if (!outs) break;
        const int s_r = first_set(outs);
        bool nzs = (nzb & ((1ull << s_r) - 1ull) & ~((1ull << p) - 1ull)) != 0;
        int xmx_t = -1, xmx_inc = 0;
        if (s_r > p) {
            xmx_t = s_r;  // xmax = the first out cell
            xmx_inc = 0;
        }
        int stop_t = -1;
        for (int t = p - 1; t >= 0; --t) {
            if ((bits >> t) & 1ull) {
                if (!nzs) {
                    xmx_t = t;  // xmax = cell + incr
                    xmx_inc = 1;
                }
                if ((nzb >> t) & 1ull) nzs = true;
            } else if (nzs || ((blw >> t) & 1ull)) {
                stop_t = t;
                break;
            }
        }
        if (stop_t < 0 || nzs) break;
        if (G.yi + 1 >= C.row_cap) break;
        const bool prev_in = stop_t + 1 <= p - 1 && ((bits >> (stop_t + 1)) & 1ull);
        xmn_lane = j * kc + (prev_in ? stop_t : p - 1);
        if (xmx_t >= 0) {
            xmx_lane = j * kc + xmx_t;
            xmx_plus = xmx_inc;
        }
        visited += (s_r - p + 1) + (p - stop_t);
        n = j + 1;
    }
    if (n == 0) return 0;
    // commit: each scan wrote its start record at yi and its one row at
    // yi + 1, then yi came back; the records hold the last scan's writes
    if (!rows_ensure(C, R, G.yi + 1)) return -1;
    const double l_mn = uni(read_lane(v_mn, n - 1));
    const double l_mx = uni(read_lane(v_mx, n - 1));
    const double l_dy = uni(read_lane(v_dy, n - 1));
    const double l_y = uni(read_lane(v_sy, n - 1));
    const int l_rc = __builtin_amdgcn_readlane(v_rc, n - 1);
    const double x_mn = uni(read_lane(xk, xmn_lane));
    double x_mx = 0.0;
    if (xmx_lane >= 0) {
        const double cell = uni(read_lane(xk, xmx_lane));
        x_mx = xmx_plus ? cell - (-incr) : cell;
    }
    wave_sync();
    {
        int idx = -1;
        double v = 0.0;
        const int f = C.lane;
        if (f == 0) { idx = G.yi * 5 + XMIN; v = l_mn; }
        else if (f == 1) { idx = G.yi * 5 + XMAX; v = l_mx; }
        else if (f == 2) { idx = G.yi * 5 + DYS; v = l_dy; }
        else if (f == 3) { idx = (G.yi + 1) * 5 + XMIN; v = x_mn; }
        else if (f == 4 && xmx_lane >= 0) { idx = (G.yi + 1) * 5 + XMAX; v = x_mx; }
        else if (f == 5) { idx = (G.yi + 1) * 5 + AX; v = 0.0; }
        else if (f == 6) { idx = (G.yi + 1) * 5 + YY; v = l_y; }
        else if (f == 7) { idx = (G.yi + 1) * 5 + DYS; v = -l_dy; }
        if (idx >= 0) R.r[idx] = v;
    }
    wave_sync();
    G.gaps += n;
    G.max_jump = mj;
    C.n_rows = uni(C.n_rows + n);
    if (K.jac_sign == 0)
        C.n_cart = uni(C.n_cart + visited);
    else
        C.n_cart_guarded = uni(C.n_cart_guarded + visited);
    C.steps = uni(C.steps + 2 * n);
    // the examination continues after the last committed scan
    const int lrow = l_rc >> 2;
    G.dxmax = fld(lrow + 1, XMAX) - fld(lrow, XMAX);
    G.dxmin = fld(lrow + 1, XMIN) - fld(lrow, XMIN);
    G.row = lrow;
    G.cs = (l_rc & 3) + 1;
    if (watch && uni(__hip_atomic_load(watch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) != expect) {
        C.overflow = true;
        return -1;
    }
    return n;
}

// The scans of one image (:2520-2680): up from the seed, down from the row
// below it, then the gap-repair scans of rows 0..nyi-1, four cases per row,
// all through one call site of cart_scan.  A failure leaves C.overflow set.
// With kTape the scans record into T; gap-repair scans run one at a time
// while it is open, so that they record in the reference's order.
template <bool kTape = false>
__device__ __forceinline__ ImageRun image_area(Ctx& C, Cart& K, Rows& R, Pt seed, ImageHook& hk,
                                               Tape* T = nullptr)
{
    ImageRun out{__builtin_nan(""), 0, 0, 0.0};
    const double incr = K.incr;
    int gaps = 0;
    double max_jump = 0.0;
    rows_reset(C, R);
    if (!rows_ensure(C, R, 1)) return out;
    int yi = 0;
    const double J_seed = jacobian_real(C.L, seed.x, seed.y);
    K.jac_sign = fabs(J_seed) < 0.02 ? sign_of(J_seed) : 0;
    int stage = 0;
    int row = -1;  // gap-repair row being examined
    int cs = 4;    // next case of `row` (4: the row is done)
    double dxmax = 0.0, dxmin = 0.0;
    int nyi = 0;
    double areai = 0.0, areabound = 0.0;
    const double g = 1.1 * incr;
    int gap_cool = 0, gap_fails = 0;  // backs off gap_batch where it commits nothing
    for (;;) {
        if (!hook_poll(C, hk)) return out;
        Pt start{0.0, 0.0};
        double sdy = 0.0;
        double prev_xmin = 0.0;  // the xmin just written to row yi - 1 (stages 1, 2)
        if (stage == 0) {
            rset(C, R, 0, XMIN, seed.x);
            rset(C, R, 0, XMAX, seed.x);
            start = seed;
            sdy = incr;
        } else if (stage == 1) {
            sdy = -incr;
            if (!rows_ensure(C, R, yi)) return out;
            start = Pt{rget(R, 0, XMAX), seed.y + sdy};
            prev_xmin = uni(rget(R, 0, XMIN));
            rset(C, R, yi, XMIN, prev_xmin);
            rset(C, R, yi, XMAX, rget(R, 0, XMAX));
            rset(C, R, yi, YY, rget(R, 0, YY));
            rset(C, R, yi, DYS, sdy);
            ++yi;
        } else {
            bool have = false;
            if (gap_cool <= 0 && row >= 0 && row <= nyi - 2 && !(kTape && tape_open(T))) {
                GapState G{row, cs, dxmax, dxmin, yi, nyi, gaps, max_jump};
                const int nb = gap_batch(C, K, R, G, hk.watch, hk.expect);
                if (nb < 0) return out;
                if (nb > 0) {
                    row = G.row;
                    cs = G.cs;
                    dxmax = G.dxmax;
                    dxmin = G.dxmin;
                    gaps = G.gaps;
                    max_jump = G.max_jump;
                    gap_fails = 0;
                    continue;
                }
                ++gap_fails;
                gap_cool = gap_fails < 4 ? (1 << gap_fails) : 16;
            } else if (gap_cool > 0) {
                --gap_cool;
            }
            while (!have) {
                if (cs >= 4) {
                    if (row >= 0 && row == nyi - 1 && areabound > 0.0 && yi > nyi) nyi = yi;
                    ++row;
                    // rows before the last whose four cases all fail change
                    // nothing: skip them 64 at a time (their records and the
                    // next row's are final while the loop runs)
                    while (row < nyi - 1) {
                        const int rr = row + C.lane;
                        bool act = rr == nyi - 1;
                        double jump = 0.0;
                        if (rr < nyi - 1 && rget(R, rr + 1, AX) > 0.0) {
                            const double ex = rget(R, rr + 1, XMAX) - rget(R, rr, XMAX);
                            const double en = rget(R, rr + 1, XMIN) - rget(R, rr, XMIN);
                            act = ex > g || en > g || en < -g || ex < -g;
                            jump = dmax(fabs(ex), fabs(en)) / incr;
                        }
                        // every row below nyi - 1 is examined by the reference's
                        // loop and its records are final here: the maximum may
                        // take in rows past the first active one
                        max_jump = dmax(max_jump, uni(wave_max(jump)));
                        const uint64_t mm = ballot(act);
                        if (mm) {
                            row += first_set(mm);
                            break;
                        }
                        row += kWave;
                    }
                    if (row >= nyi) break;
                    if (!rows_ensure(C, R, row + 1)) return out;
                    dxmax = rget(R, row + 1, XMAX) - rget(R, row, XMAX);
                    dxmin = rget(R, row + 1, XMIN) - rget(R, row, XMIN);
                    cs = rget(R, row + 1, AX) > 0.0 ? 0 : 4;
                    if (cs == 0) max_jump = dmax(max_jump, uni(dmax(fabs(dxmax), fabs(dxmin)) / incr));
                    continue;
                }
                const int k = cs++;
                double nxmin = 0.0, nxmax = 0.0;
                if (k == 0 && dxmax > g) {
                    start = Pt{rget(R, row + 1, XMAX), rget(R, row, YY)};
                    nxmin = rget(R, row, XMAX);
                    nxmax = rget(R, row + 1, XMAX);
                    sdy = -rget(R, row, DYS);
                    have = true;
                } else if (k == 1 && dxmin > g) {
                    start = Pt{rget(R, row + 1, XMIN) - incr, rget(R, row + 1, YY)};
                    nxmin = rget(R, row, XMIN);
                    nxmax = rget(R, row + 1, XMIN);
                    sdy = rget(R, row, DYS);
                    have = true;
                } else if (k == 2 && dxmin < -g) {
                    start = Pt{rget(R, row, XMIN) - incr, rget(R, row, YY)};
                    nxmin = rget(R, row + 1, XMIN);
                    nxmax = rget(R, row, XMIN);
                    sdy = -rget(R, row, DYS);
                    have = true;
                } else if (k == 3 && dxmax < -g) {
                    start = Pt{rget(R, row + 1, XMAX) + incr, rget(R, row + 1, YY)};
                    nxmin = rget(R, row + 1, XMAX);
                    nxmax = rget(R, row, XMAX);
                    sdy = rget(R, row, DYS);
                    have = true;
                }
                if (have) {
                    ++gaps;
                    start.x = uni(start.x);
                    start.y = uni(start.y);
                    sdy = uni(sdy);
                    nxmin = uni(nxmin);
                    nxmax = uni(nxmax);
                    if (!rows_ensure(C, R, yi)) return out;
                    prev_xmin = nxmin;
                    rset(C, R, yi, XMIN, nxmin);
                    rset(C, R, yi, XMAX, nxmax);
                    rset(C, R, yi, DYS, sdy);
                    ++yi;
                }
            }
            if (!have) break;
        }
        const double r = cart_scan<kTape>(C, K, R, start, sdy, yi, prev_xmin, hk.watch, hk.expect, T);
        if (C.overflow) return out;
        if (stage == 0) {
            areai = r;
            stage = 1;
        } else if (stage == 1) {
            areai += r;
            nyi = yi;
            stage = 2;
        } else {
            areai += r;
            areabound += r;
            if (r <= 0.0) --yi;
        }
    }
    out.area = areai;
    out.nyi = nyi;
    out.gaps = gaps;
    out.max_jump = max_jump;
    return out;
}

// Accounting of image i (:2674-2739): its area is added, then each other seed
// the image's rows cover is flagged (a later one is skipped; an earlier one's
// area is subtracted once).  A seed flagged to be skipped has its job dropped.
__device__ __forceinline__ void account_image(Ctx& C, const Cart& K0, int i, const Rows& R, ImageRun img,
                                              int source_bins, double& area, double* areaimage, int* flags,
                                              Team* team, int w)
{
    const Pt seed = C.seeds[i];
    const double J_seed = jacobian_real(C.L, seed.x, seed.y);
    const int jac_sign = fabs(J_seed) < 0.02 ? sign_of(J_seed) : 0;
    const double m = 0.5 * K0.incr * 1.01;
    const int n_img = C.n_seeds;
    const int nyi = img.nyi;
    area += img.area;
    // diagnostics of this image (:2536-2540, 2673-2680)
    {
        int counted = 0;
        for (int base = 0; base < nyi; base += kWave) {
            const int row = base + C.lane;
            counted += popcount64(ballot(row < nyi && rget(R, row, AX) > 0.0));
        }
        C.d_processed = uni(C.d_processed + 1);
        C.d_fold = uni(C.d_fold + (jac_sign != 0 ? 1 : 0));
        C.d_boundary_rows = uni(C.d_boundary_rows + counted);
        C.d_gaps = uni(C.d_gaps + img.gaps);
        C.d_max_jump = uni(dmax(C.d_max_jump, img.max_jump));
    }
    wave_sync();
    wave_store(&areaimage[i], img.area);
    wave_sync();
    for (int other = 0; other < n_img; ++other) {
        if (other == i) continue;
        const Pt pos = C.seeds[other];
        if (jac_sign != 0) {
            const double Jo = jacobian_real(C.L, pos.x, pos.y);
            if (sign_of(Jo) == -jac_sign) continue;
            if (C.rho >= 4.0e-3 && source_bins >= 35 && other > i && fabs(J_seed) < 1.0e-3 && fabs(Jo) < 0.02)
                continue;
        }
        bool hit = false;
        for (int base = 0; base < nyi && !hit; base += kWave) {
            const int row = base + C.lane;
            bool h = false;
            if (row < nyi && rget(R, row, AX) > 0.0) {
                const double yy = rget(R, row, YY);
                h = pos.y >= yy - m && pos.y <= yy + m && pos.x >= rget(R, row, XMIN) - m &&
                    pos.x <= rget(R, row, XMAX) + m;
            }
            hit = ballot(h) != 0;
        }
        if (hit) {
            C.d_overlaps = uni(C.d_overlaps + 1);
            const int fl = flags[other];
            wave_sync();
            if (other < i) {
                if (!(fl & 2)) {
                    area -= areaimage[other];
                    wave_store(&flags[other], fl | 2);
                }
            } else {
                wave_store(&flags[other], fl | 1);
            }
            wave_sync();
            if (other > i && team) {
                const int k = team_find(*team, w, other);
                if (k >= 0) team_drop(*team, w, k);
            }
        }
    }
}

__device__ __forceinline__ double cart_scale(const Ctx& C, int source_bins)
{
    const double nbin = static_cast<double>(source_bins > 1 ? source_bins : 1);
    return source_flux(C.rho, C.ld_c, C.ld_d) / (C.rho * C.rho) * nbin * nbin;
}

// high_magnification_floor_coefficient (finite_source_magnifier.cpp:1017-1063).
__device__ double cart_floor_coefficient(const Ctx& C, double mag)
{
    const double r = C.rho;
    const double am = fabs(mag);
    const int seeds = C.n_seeds, gaps = C.d_gaps, ov = C.d_overlaps;
    const double jump = C.d_max_jump;
    if (r >= 0.1 && (gaps > 0 || jump > 50.0)) return 0.06;
    if (r < 1.0e-2 && seeds <= 5 && am <= 10.0 && gaps > 0 && jump > 20.0) return 0.035;
    if (r < 1.0e-2 && seeds <= 5 && am <= 10.0 && jump > 1000.0) return 0.04;
    if (r < 1.0e-2 && seeds >= 16 && ov >= 8 && gaps >= 100 && jump > 1000.0 && am <= 50.0) return 0.035;
    if (r < 1.0e-2 && am <= 300.0) return 0.02;
    if (r < 1.0e-2 && am <= 1000.0) return 0.04;
    if (r < 1.0e-2 && seeds <= 5 && gaps > 0 && jump > 20.0) return 0.04;
    if (am <= 80.0 || gaps <= 1000) return 0.0;
    return r >= 1.0e-2 ? 0.04 : 0.05;
}

// The error estimate of a Cartesian epoch: cartesian_area_error_indicator
// (:1065-1106) raised to the high-magnification floor
// (inverse_ray_cartesian_binary_mag, :2745-2756).
__device__ double cart_error_estimate(const Ctx& C, int source_bins, double mag)
{
    const int bins = source_bins > 1 ? source_bins : 1;
    const double r = C.rho;
    const double flux = source_flux(r, C.ld_c, C.ld_d);
    double est = 0.0;
    if (isfinite(flux) && flux > 0.0 && r > 0.0) {
        const int seeds = C.n_seeds;
        const double gap_w = r >= 2.0e-2 ? 0.005 : ((r < 1.0e-2 && seeds >= 16) ? 0.015 : 0.03);
        const double ov_w = r >= 2.0e-2 ? 0.005 : 0.02;
        const double bnd_w = r >= 2.0e-2 ? 0.012 : 0.03;
        double cells = bnd_w * static_cast<double>(C.d_boundary_rows) + gap_w * static_cast<double>(C.d_gaps) +
                       ov_w * static_cast<double>(C.d_overlaps) +
                       0.02 * static_cast<double>(C.d_fold > 0 ? C.d_fold : 0) +
                       0.002 * static_cast<double>(seeds - 5 > 0 ? seeds - 5 : 0);
        if (r >= 1.0e-3 && seeds >= 16 && C.d_overlaps >= 8) {
            const double island_w = r < 1.0e-2 ? 0.001 : 0.0005;
            cells += island_w * static_cast<double>(seeds) * static_cast<double>(C.d_overlaps);
        }
        if (r >= 1.0e-3 && C.d_max_jump > 10.0) {
            const double jump_w = r >= 2.0e-2 ? 0.2 : 4.0;
            cells += jump_w * log10(C.d_max_jump / 10.0);
        }
        const double nb = static_cast<double>(bins);
        const double cell_area = (r / nb) * (r / nb);
        const double e = 1.25 * cells * cell_area / flux;
        est = isfinite(e) ? e : 0.0;
    }
    const double fc = cart_floor_coefficient(C, mag);
    if (fc > 0.0) est = dmax(est, fabs(mag) * (fc / static_cast<double>(bins)));
    return est;
}

// The whole Cartesian area of one epoch by one wave, images in seed order
// (the epoch-at-a-time form of k_ir_cart's loop, for single-epoch kernels).
template <bool kTape = false>
__device__ double cartesian_mag(Ctx& C, int source_bins, const CartScratch& scr, int slot, Tape* T = nullptr)
{
    const int n_img = C.n_seeds;
    if (n_img == 0 || !(C.rho > 0.0)) return __builtin_nan("");
    Cart K;
    cart_params(K, C, source_bins);
    double* areaimage = scr.areaimage + static_cast<size_t>(slot) * kSeedCap;
    int* flags = scr.overlap + static_cast<size_t>(slot) * kSeedCap;
    wave_sync();
    for (int k = C.lane; k < n_img; k += kWave) {
        areaimage[k] = 0.0;
        flags[k] = 0;  // bit 0: overlap, bit 1: subtracted
    }
    wave_sync();
    Rows R = slot_rows(scr, slot);
    ImageHook hk{nullptr, 0, nullptr, 0, nullptr, 0, 0, 0, nullptr, false};
    double area = 0.0;
    for (int i = 0; i < n_img; ++i) {
        if (flags[i] & 1) continue;
        const ImageRun img = image_area<kTape>(C, K, R, C.seeds[i], hk, T);
        if (C.overflow) return __builtin_nan("");
        account_image(C, K, i, R, img, source_bins, area, areaimage, flags, nullptr, 0);
    }
    return area / cart_scale(C, source_bins);
}

// ---------------------------------------------------------------- polar fill

// Visited radial runs, the reference's per-row interval lists
// (PolarVisitedCellIntervals, finite_source_magnifier.cpp:909) as chains hashed
// on the phi row: head[iphi & (kPhiHash-1)] -> newest run, runs[4r+3] -> older.
struct Polar {
    double dr, dphi;
    int phi_bins;
    int* head;   // [kPhiHash] -> newest run with that hash, -1 if none
    int* runs;   // [run*4]: iphi, left, right, next
    int n_runs;
};

__device__ __forceinline__ bool polar_visited(const Polar& P, int ir, int iphi)
{
    if (ir < 0) return true;
    for (int r = P.head[iphi & (kPhiHash - 1)]; r >= 0; r = P.runs[4 * r + 3]) {
        if (P.runs[4 * r] == iphi && ir >= P.runs[4 * r + 1] && ir <= P.runs[4 * r + 2]) return true;
    }
    return false;
}

// cell_inside of inverse_ray_polar_boundary_binary (:1241-1254); the angle's
// cos/sin come from the device libm.
__device__ __forceinline__ bool polar_inside(const Ctx& C, const Polar& P, int ir, double cph,
                                             double sph, double* dz2)
{
    if (ir < 0) return false;
    const double radius = (static_cast<double>(ir) + 0.5) * P.dr;
    const Pt w = map_real(C.L, radius * cph, radius * sph);
    const double d = dist2(w, C.src);
    *dz2 = d;
    return d <= C.r2;
}

__device__ __forceinline__ void phi_trig(const Polar& P, int iphi, double* c, double* s)
{
    const double phi = (static_cast<double>(iphi) + 0.5) * P.dphi;
    sincos(phi, s, c);
}

// Extends an inside, unvisited cell to its radial run (:1290-1297), records
// it and returns its brightness-weighted sum of radii.
__device__ double polar_add_run(Ctx& C, Polar& P, int ir, int iphi, double cph, double sph)
{
    // left
    int left = ir;
    for (;;) {
        if (!step_ok(C)) return 0.0;
        const int c = left - 1 - C.lane;
        bool ok = false;
        double dz2;
        if (c >= 0) ok = !polar_visited(P, c, iphi) && polar_inside(C, P, c, cph, sph, &dz2);
        const uint64_t m = ballot(ok);
        const int stop = first_set(~m);
        if (stop < 0) {
            left -= kWave;
            continue;
        }
        left -= stop;
        break;
    }
    int right = ir;
    for (;;) {
        if (!step_ok(C)) return 0.0;
        const int c = right + 1 + C.lane;
        double dz2;
        const bool ok = !polar_visited(P, c, iphi) && polar_inside(C, P, c, cph, sph, &dz2);
        const uint64_t m = ballot(ok);
        const int stop = first_set(~m);
        if (stop < 0) {
            right += kWave;
            continue;
        }
        right += stop;
        break;
    }
    if (P.n_runs >= C.run_cap) {
        C.overflow = true;
        C.cap_overflow = true;
        return 0.0;
    }
    // the run's cells and the two outside cells that end it
    C.n_polar = uni(C.n_polar + (right - left + 3));
    C.n_rows = uni(C.n_rows + 1);
    const int old_head = uni(P.head[iphi & (kPhiHash - 1)]);
    wave_sync();
    {
        const int r = P.n_runs;
        wave_store(&P.runs[4 * r + 0], iphi);
        wave_store(&P.runs[4 * r + 1], left);
        wave_store(&P.runs[4 * r + 2], right);
        wave_store(&P.runs[4 * r + 3], old_head);
        wave_store(&P.head[iphi & (kPhiHash - 1)], r);
    }
    P.n_runs = uni(P.n_runs + 1);
    wave_sync();
    double sum = 0.0;
    for (int c = left + C.lane; c <= right; c += kWave) {
        double dz2 = 0.0;
        polar_inside(C, P, c, cph, sph, &dz2);
        const double radius = (static_cast<double>(c) + 0.5) * P.dr;
        const double b = C.uniform ? 1.0 : ld_brightness(dz2 * C.inv_r2, C.ld_c, C.ld_d);
        sum += b * radius;
    }
    return uni(wave_sum(sum));
}

__device__ __forceinline__ double wrap_angle(double angle)
{
    while (angle < 0.0) angle += 2.0 * kPi;
    while (angle >= 2.0 * kPi) angle -= 2.0 * kPi;
    return angle;
}

// inverse_ray_polar_boundary_binary (:1156-1319): the flood fill sums every
// cell of the 4-connected inside components that contain a start cell; here
// the components are grown run by run (scan-line fill over neighbour rows).
__device__ double polar_mag(Ctx& C, int source_bins, double grid_ratio, int* head, int* runs)
{
    const int n_img = C.n_seeds;
    if (n_img == 0) return __builtin_nan("");
    Polar P;
    P.dr = C.rho / static_cast<double>(source_bins);
    double max_r = 1.0;
    for (int i = 0; i < n_img; ++i) {
        const Pt s = C.seeds[i];
        max_r = dmax(max_r, hypot_glibc(s.x, s.y) + 4.0 * C.rho);
    }
    const double pbd = ceil(2.0 * kPi * max_r / (P.dr * grid_ratio));
    if (!(pbd <= static_cast<double>(kPhiCap))) {
        C.overflow = true;
        return __builtin_nan("");
    }
    const int pb = static_cast<int>(pbd);
    P.phi_bins = pb > 16 ? pb : 16;
    P.dphi = 2.0 * kPi / static_cast<double>(P.phi_bins);
    const double flux = source_flux(C.rho, C.ld_c, C.ld_d);
    if (!isfinite(flux)) return __builtin_nan("");
    P.head = head;
    P.runs = runs;
    P.n_runs = 0;
    double total = 0.0;

    // start cells (find_polar_inside_start, :1321-1364): 242 candidates in
    // the reference's shell / direction / radial order, evaluated in parallel
    for (int i = 0; i < n_img && !C.overflow; ++i) {
        const Pt seed = C.seeds[i];
        const double seed_radius = hypot_glibc(seed.x, seed.y);
        const double seed_phi = wrap_angle(atan2(seed.y, seed.x));
        const int seed_ir = static_cast<int>(floor(seed_radius / P.dr));
        const int seed_iphi = iclamp(static_cast<int>(seed_phi / P.dphi), 0, P.phi_bins - 1);
        int found_ord = -1;
        for (int base = 0; base < 242 && found_ord < 0; base += kWave) {
            const int o = base + C.lane;
            bool in = false;
            if (o < 242) {
                // decode o -> (shell, dir, d_ir); shell s has 2*(2s+1) entries
                int shell = 0, rem = o;
                while (rem >= 2 * (2 * shell + 1)) {
                    rem -= 2 * (2 * shell + 1);
                    ++shell;
                }
                const int dir = rem < (2 * shell + 1) ? -1 : 1;
                const int d_ir = (rem % (2 * shell + 1)) - shell;
                const int d_iphi = (shell - abs(d_ir)) * dir;
                const int ir = seed_ir + d_ir;
                if (ir >= 0) {
                    int iphi = (seed_iphi + d_iphi) % P.phi_bins;
                    if (iphi < 0) iphi += P.phi_bins;
                    double cph, sph, dz2;
                    phi_trig(P, iphi, &cph, &sph);
                    in = polar_inside(C, P, ir, cph, sph, &dz2);
                }
            }
            const uint64_t m = ballot(in);
            if (m) found_ord = base + first_set(m);
        }
        C.n_polar = uni(C.n_polar + (found_ord < 0 ? 242 : found_ord + 1));
        if (found_ord < 0) continue;
        int shell = 0, rem = found_ord;
        while (rem >= 2 * (2 * shell + 1)) {
            rem -= 2 * (2 * shell + 1);
            ++shell;
        }
        const int dir = rem < (2 * shell + 1) ? -1 : 1;
        const int d_ir = (rem % (2 * shell + 1)) - shell;
        const int d_iphi = (shell - abs(d_ir)) * dir;
        const int sir = seed_ir + d_ir;
        int siphi = (seed_iphi + d_iphi) % P.phi_bins;
        if (siphi < 0) siphi += P.phi_bins;
        const double grid_radius = (static_cast<double>(sir) + 0.5) * P.dr;
        const double image_phi = (static_cast<double>(siphi) + 0.5) * P.dphi;
        const int fl = static_cast<int>(floor(grid_radius / P.dr));
        const int ir = fl > 0 ? fl : 0;
        const int iphi = iclamp(static_cast<int>(image_phi / P.dphi), 0, P.phi_bins - 1);
        double cph, sph, dz2;
        phi_trig(P, iphi, &cph, &sph);
        if (polar_visited(P, ir, iphi) || !polar_inside(C, P, ir, cph, sph, &dz2)) continue;
        total += polar_add_run(C, P, ir, iphi, cph, sph);
    }
    // grow: each recorded run visits its two angular neighbour rows
    for (int q = 0; q < P.n_runs && !C.overflow; ++q) {
        const int p = runs[4 * q + 0];
        const int L = runs[4 * q + 1];
        const int Rr = runs[4 * q + 2];
        for (int side = 0; side < 2; ++side) {
            int nb = side == 0 ? p - 1 : p + 1;
            if (nb < 0) nb += P.phi_bins;
            if (nb >= P.phi_bins) nb -= P.phi_bins;
            double cph, sph;
            phi_trig(P, nb, &cph, &sph);
            int c = L;
            while (c <= Rr && step_ok(C)) {
                const int cc = c + C.lane;
                bool ok = false;
                if (cc <= Rr) {
                    double dz2;
                    ok = !polar_visited(P, cc, nb) && polar_inside(C, P, cc, cph, sph, &dz2);
                }
                const uint64_t m = ballot(ok);
                if (!m) {
                    const int span = Rr - c + 1;
                    C.n_polar = uni(C.n_polar + (span < kWave ? span : kWave));
                    c += kWave;
                    continue;
                }
                C.n_polar = uni(C.n_polar + first_set(m));
                const int start = c + first_set(m);
                total += polar_add_run(C, P, start, nb, cph, sph);
                c = runs[4 * (P.n_runs - 1) + 2] + 1;
            }
        }
    }
    // reset the rows this epoch touched
    wave_sync();
    for (int q = C.lane; q < P.n_runs; q += kWave) head[runs[4 * q] & (kPhiHash - 1)] = -1;
    wave_sync();
    if (C.overflow) return __builtin_nan("");
    const double image_flux = total * P.dr * P.dphi;
    return image_flux / flux;
}

// ------------------------------------------------------------------ kernels

__device__ __forceinline__ void init_ctx(Ctx& C, const CurveRec& cv, Pt src, Pt* seeds)
{
    C.s = cv.s;
    C.q = cv.q_eff;
    C.L = make_lens(cv.s, cv.q_eff);
    C.src = src;
    C.rho = cv.rho;
    C.r2 = cv.rho * cv.rho;
    C.inv_r2 = 1.0 / C.r2;
    C.ld_c = cv.ld_c;
    C.ld_d = cv.ld_d;
    C.uniform = cv.ld_c == 0.0 && cv.ld_d == 0.0;
    C.lane = lane_id();
    C.seeds = seeds;
    C.n_seeds = 0;
    C.overflow = false;
    C.steps = 0;
    C.budget = kStepBudget;
    C.last_phase = 5;
    C.n_quintic = 0;
    C.n_quartic = 0;
    C.n_cart = 0;
    C.n_cart_guarded = 0;
    C.n_polar = 0;
    C.n_rows = 0;
    C.row_cap = 0;
    C.run_cap = 0;
    C.cap_overflow = false;
    C.farm = nullptr;
    C.p1_start = 0;
    C.defer_budget = -1;
    C.defer_at = -1;
    C.defer_slot = -1;
    C.heavy_count = nullptr;
    C.seeds_changed = false;
    C.hints = false;
    C.d_processed = 0;
    C.d_fold = 0;
    C.d_boundary_rows = 0;
    C.d_gaps = 0;
    C.d_overlaps = 0;
    C.d_max_jump = 0.0;
}

__device__ __forceinline__ unsigned long long* work_totals(const FsEpochs& ep)
{
    return reinterpret_cast<unsigned long long*>(ep.counters + kWorkOffset);
}

// Adds a uniform amount to a run total (every lane adds, all but lane 0 zero).
__device__ __forceinline__ void add_total(unsigned long long* p, long long v)
{
    atomicAdd(p, lane_id() == 0 ? static_cast<unsigned long long>(v) : 0ull);
}

__device__ __forceinline__ void flush_work(const Ctx& C, const FsEpochs& ep)
{
    unsigned long long* w = work_totals(ep);
    if (C.n_quintic) add_total(&w[0], C.n_quintic);
    if (C.n_quartic) add_total(&w[1], C.n_quartic);
    if (C.n_cart) add_total(&w[2], C.n_cart);
    if (C.n_cart_guarded) add_total(&w[3], C.n_cart_guarded);
    if (C.n_polar) add_total(&w[4], C.n_polar);
    if (C.n_rows) add_total(&w[7], C.n_rows);
}

__device__ __forceinline__ int curve_nbin(const CurveRec& cv, const FsOptions& opt)
{
    return cv.source_bins > 0 ? cv.source_bins : opt.source_bins;
}

// Seeds of epoch e into C (fixed_inverse_ray_binary's seed step,
// finite_source_magnifier.cpp:3042-3100).
__device__ __forceinline__ void epoch_seeds(Ctx& C, const FsCache& cache, const FsTables& tab,
                                            const FsEpochs& ep, int c, int64_t e)
{
    const int bins = cache.bins;
    const CausticScalars sc = cache.sc[c];
    image_seeds(C, cache.crit + static_cast<size_t>(c) * kSeedPhases * 4,
                cache.assign + static_cast<size_t>(c) * kSeedPhases * 4,
                cache.br + static_cast<size_t>(c) * bins * 4,
                cache.br_cell + static_cast<size_t>(c) * bins * 4,
                cache.br_order + static_cast<size_t>(c) * bins * 4, sc, bins, tab,
                ep.roots + 10 * e);
}

__device__ __forceinline__ double polar_epoch(Ctx& C, const CurveRec& cv, const FsOptions& opt,
                                              int* head, int* runs)
{
    if (C.overflow) return __builtin_nan("");
    const int pbins = opt.polar_source_bins > 0 ? opt.polar_source_bins : curve_nbin(cv, opt);
    const int psb = pbins > 1 ? pbins : 1;
    double ratio = opt.polar_grid_ratio > 0.0 ? opt.polar_grid_ratio : opt.grid_ratio;
    ratio = dmax(ratio, 1.0e-12);
    if (opt.mode == 4) ratio = dmax(ratio, 12.0);
    return polar_mag(C, psb, ratio, head, runs);
}

__device__ __forceinline__ double cart_epoch(Ctx& C, const CurveRec& cv, const FsOptions& opt,
                                             const CartScratch& scr, int slot)
{
    if (C.overflow) return __builtin_nan("");
    C.row_cap = scr.row_cap;
    return cartesian_mag(C, curve_nbin(cv, opt), scr, slot);
}

// Cost bucket of an inverse-ray epoch, largest first.  The key is the epoch's
// measured ticks in the previous run when there is one (a fit evaluates the
// same epochs again and again at nearby parameters), else an estimate: the
// lattice cells of an epoch scale as |A_ps| nbin^2 (times about 32 ticks).
__device__ __forceinline__ int cost_bucket(const FsEpochs& ep, int64_t e, int nbin)
{
    const float learned = ep.sched ? ep.sched[e] : 0.0f;
    const double est = learned > 0.0f ? static_cast<double>(learned)
                                      : 32.0 * fabs(ep.a_ps[e]) * static_cast<double>(nbin) * static_cast<double>(nbin);
    int lg = est >= 1.0 ? ilogb(est) : 0;
    lg = lg < kCostBuckets - 1 ? lg : kCostBuckets - 1;
    return kCostBuckets - 1 - lg;
}

__global__ void k_ir_bucket_count(const CurveRec* __restrict__ curves, FsOptions opt, FsEpochs ep,
                                  IrLists L)
{
    const int n = static_cast<int>(ep.counters[kQIrCount]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int64_t e = ep.ir_list[i];
        const CurveRec cv = curves[ep.curve_of[e]];
        const int b = cost_bucket(ep, e, curve_nbin(cv, opt));
        L.bucket_of[i] = b;
        atomicAdd(&ep.counters[kBucketOffset + b], 1u);
    }
}

__global__ void k_ir_bucket_scan(FsEpochs ep)
{
    if (threadIdx.x != 0) return;
    unsigned run = 0;
    for (int b = 0; b < kCostBuckets; ++b) {
        const unsigned c = ep.counters[kBucketOffset + b];
        ep.counters[kBucketOffset + kCostBuckets + b] = run;
        run += c;
    }
}

__global__ void k_ir_bucket_scatter(FsEpochs ep, IrLists L)
{
    const int n = static_cast<int>(ep.counters[kQIrCount]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned pos = atomicAdd(&ep.counters[kBucketOffset + kCostBuckets + L.bucket_of[i]], 1u);
        L.ir_sorted[pos] = ep.ir_list[i];
    }
}

// Seed search of every inverse-ray epoch (largest first); the seeds go to the
// shared pool and the epoch to the Cartesian or polar queue.  Pass 0 runs
// every search and defers the long first-crossing scans to the probe farm;
// pass 1 resumes the deferred ones with the farm's probe flags.
__global__ void __launch_bounds__(kWave) k_ir_seeds(const CurveRec* __restrict__ curves,
                                                     FsCache cache, FsTables tab, FsOptions opt,
                                                     FsEpochs ep, IrLists L, SeedScratch scr, int pass)
{
    Pt* slot_seeds = scr.seeds + static_cast<size_t>(blockIdx.x) * kSeedCap;
    const HeavyScratch& hv = scr.heavy;
    const int n_heavy = uni(static_cast<int>(ep.counters[kQHeavyCount]));
    const int n_ir = uni(static_cast<int>(ep.counters[kQIrCount]));
    const int lead = opt.lead_items < n_ir ? opt.lead_items : n_ir;
    const int count = pass == kSeedLead ? lead
                    : pass == kSeedMain ? n_ir - lead
                                        : (n_heavy < kHeavyCap ? n_heavy : kHeavyCap);
    unsigned* cursor = &ep.counters[pass == kSeedLead ? kQLeadCursor : pass == kSeedMain ? kQIrCursor : kQHeavyCursor];
    for (;;) {
        const int task = next_item(cursor);
        if (task >= count) break;
        const int item = pass == kSeedLead ? task : pass == kSeedMain ? lead + task : uni(hv.item[task]);
        const int64_t e = L.ir_sorted[item];
        const long long t_begin = static_cast<long long>(wall_clock64());
        const int c = ep.curve_of[e];
        const CurveRec cv = curves[c];
        Ctx C;
        init_ctx(C, cv, ep.src[e], slot_seeds);
        C.budget = opt.step_budget;
        C.hints = opt.ps_images != 0;
        if (pass != kSeedResume) {
            C.defer_budget = opt.heavy_budget;
            C.heavy_count = &ep.counters[kQHeavyCount];
        } else {
            C.farm = hv.flags + static_cast<size_t>(task) * kScanPositions * kProbesPerPosition;
            C.p1_start = uni(hv.start[task]);
        }
        epoch_seeds(C, cache, tab, ep, c, e);
        if (C.defer_at >= 0) {
            // deferred: the farm and pass 1 finish this search
            wave_store(&hv.item[C.defer_slot], item);
            wave_store(&hv.start[C.defer_slot], C.defer_at);
            flush_work(C, ep);
            const long long ticks = static_cast<long long>(wall_clock64()) - t_begin;
            add_total(&work_totals(ep)[5], ticks);
            if (ep.cost) {
                wave_store(&ep.cost[e], static_cast<int64_t>(ticks));
                wave_store(&ep.cost[2 * ep.n + e], static_cast<int64_t>(C.n_quintic));
            }
            if (ep.sched) wave_store(&ep.sched[e], static_cast<float>(ticks));
            wave_sync();
            continue;
        }
        const int n = C.n_seeds;
        bool ok = !C.overflow;
        int off = 0;
        if (ok && n > 0) {
            off = wave_fetch_add(&ep.counters[kQPoolCursor], static_cast<unsigned>(n));
            if (static_cast<unsigned>(off) + static_cast<unsigned>(n) > L.pool_cap) {
                ok = false;
                atomicOr(&ep.counters[kQFlags], kErrSeedPool);
            }
        }
        if (ok) {
            wave_sync();
            for (int k = C.lane; k < n; k += kWave) L.seed_pool[off + k] = slot_seeds[k];
            wave_store(&L.seed_off[item], off);
            wave_store(&L.seed_n[item], n);
            const bool polar = uni(ep.method[e]) == LCBG_METHOD_INVERSE_RAY_POLAR;
            // the leading items start on large slots while the other searches
            // run; the rest wait for every search to end (an epoch that outgrows
            // a small slot is retried on a large one)
            const bool large = pass == kSeedLead;
            const int qword = polar ? (large ? kQPolarBigCount : kQPolarCount)
                                    : (large ? kQCartBigCount : kQCartCount);
            int* qarr = polar ? (large ? L.polar_big : L.polar_q) : (large ? L.cart_big : L.cart_q);
            const int pos = wave_fetch_inc(&ep.counters[qword]);
            wave_store(&qarr[pos], item);
        } else {
            wave_store(&L.seed_n[item], -1);
            wave_store(&ep.mag[e], __builtin_nan(""));
            if (C.overflow) atomicOr(&ep.counters[kQFlags], C.steps > C.budget ? kErrBudget : kErrSeedCap);
        }
        flush_work(C, ep);
        const long long ticks = static_cast<long long>(wall_clock64()) - t_begin;
        add_total(&work_totals(ep)[5], ticks);
        if (ep.cost) {
            const int64_t prior = pass != kSeedResume ? 0 : ep.cost[e];
            const int64_t prior_q = pass != kSeedResume ? 0 : ep.cost[2 * ep.n + e];
            wave_store(&ep.cost[e], prior + static_cast<int64_t>(ticks));
            wave_store(&ep.cost[2 * ep.n + e], prior_q + static_cast<int64_t>(C.n_quintic));
        }
        if (ep.sched)
            wave_store(&ep.sched[e], (pass != kSeedResume ? 0.0f : ep.sched[e]) + static_cast<float>(ticks));
        wave_sync();
    }
}

// Inside scan positions of each deferred search from its start on, in scan
// order (one wave per search).
__global__ void __launch_bounds__(kWave) k_ir_heavy_list(const CurveRec* __restrict__ curves,
                                                          FsCache cache, FsEpochs ep, IrLists L,
                                                          HeavyScratch hv)
{
    const int h = blockIdx.x;
    const int n_heavy = uni(static_cast<int>(ep.counters[kQHeavyCount]));
    if (h >= n_heavy || h >= kHeavyCap) return;
    const int lane = lane_id();
    const int item = uni(hv.item[h]);
    const int start = uni(hv.start[h]);
    const int64_t e = L.ir_sorted[item];
    const int c = ep.curve_of[e];
    const double rho = curves[c].rho;
    const double r2 = rho * rho;
    const Pt src = ep.src[e];
    const Pt* crit = cache.crit + static_cast<size_t>(c) * kSeedPhases * 4;
    uint16_t* list = hv.list + static_cast<size_t>(h) * kScanPositions;
    int n = 0;
    for (int base = start & ~(kWave - 1); base < kScanPositions; base += kWave) {
        const int o = base + lane;
        bool inside = false;
        if (o < kScanPositions && o >= start) {
            const int ph = o >> 2;
            const int kphi = ph / 35;
            const int jphi = ph - kphi * 35;
            const int sample = jphi * 40 + kphi;
            const double d2 = dist2(crit[4 * sample + (o & 3)], src);
            inside = d2 < r2 && d2 > 0.0;
        }
        const uint64_t m = ballot(inside);
        if (inside) list[n + popcount64(m & ((1ull << lane) - 1ull))] = static_cast<uint16_t>(o);
        n += popcount64(m);
    }
    wave_store(&hv.n_list[h], n);
}

// The probe farm: every caustic and angular probe of the listed positions of
// every deferred search, one probe per lane, kFarmChunk positions per task;
// flags[o * kProbesPerPosition + k] = probe k of position o is inside the disk
// and finds more than three images.
__global__ void __launch_bounds__(kWave) k_ir_farm(const CurveRec* __restrict__ curves,
                                                    FsCache cache, FsTables tab, FsEpochs ep,
                                                    IrLists L, HeavyScratch hv)
{
    const int n_heavy = uni(static_cast<int>(ep.counters[kQHeavyCount]));
    const int tasks = (n_heavy < kHeavyCap ? n_heavy : kHeavyCap) * kFarmChunks;
    long long solved = 0;
    for (;;) {
        const int t = next_item(&ep.counters[kQFarmCursor]);
        if (t >= tasks) break;
        const int h = t / kFarmChunks;
        const int i0 = (t - h * kFarmChunks) * kFarmChunk;
        const int nl = uni(hv.n_list[h]);
        if (i0 >= nl) continue;
        const int np = nl - i0 < kFarmChunk ? nl - i0 : kFarmChunk;
        const int item = uni(hv.item[h]);
        const int64_t e = L.ir_sorted[item];
        const int c = ep.curve_of[e];
        Ctx C;
        init_ctx(C, curves[c], ep.src[e], nullptr);
        const Pt* crit = cache.crit + static_cast<size_t>(c) * kSeedPhases * 4;
        const uint16_t* list = hv.list + static_cast<size_t>(h) * kScanPositions + i0;
        uint8_t* flags = hv.flags + static_cast<size_t>(h) * kScanPositions * kProbesPerPosition;
        for (int q0 = 0; q0 < np * kProbesPerPosition; q0 += kWave) {
            const int q = q0 + C.lane;
            if (q >= np * kProbesPerPosition) continue;
            const int pi = q / kProbesPerPosition;
            const int k = q - pi * kProbesPerPosition;
            const int o = list[pi];
            const int ph = o >> 2;
            const int kphi = ph / 35;
            const int jphi = ph - kphi * 35;
            const int sample = jphi * 40 + kphi;
            const Pt cs = crit[4 * sample + (o & 3)];
            bool valid = false;
            double ux, uy;
            if (caustic_probe_axis(C, cs, &ux, &uy)) {
                const ProbeResult r = eval_probe(C, caustic_probe_point(C, cs, ux, uy, k, tab));
                valid = r.n > 3;
                solved += r.solved ? 1 : 0;
            }
            flags[static_cast<size_t>(o) * kProbesPerPosition + k] = valid ? 1 : 0;
        }
    }
    if (solved) atomicAdd(&work_totals(ep)[0], static_cast<unsigned long long>(solved));
}

// The area queues in item order, i.e. largest cost key first: the seed passes
// append items in the order their searches end (deferred searches last).  One
// workgroup per queue (queues q0, q0 + 1, ...: small Cartesian, small polar,
// large Cartesian, large polar), a bitonic sort in LDS; a queue longer than
// kSortCap keeps its order.
constexpr int kSortCap = 16384;

__global__ void __launch_bounds__(1024) k_ir_sort_queues(FsEpochs ep, IrLists L, int q0)
{
    __shared__ int key[kSortCap];
    const int qn = q0 + static_cast<int>(blockIdx.x);
    int* q = qn == 0 ? L.cart_q : qn == 1 ? L.polar_q : qn == 2 ? L.cart_big : L.polar_big;
    const int word = qn == 0 ? kQCartCount : qn == 1 ? kQPolarCount : qn == 2 ? kQCartBigCount : kQPolarBigCount;
    const int n = static_cast<int>(ep.counters[word]);
    if (n < 2 || n > kSortCap) return;
    int m = 2;
    while (m < n) m <<= 1;
    for (int i = threadIdx.x; i < m; i += blockDim.x) key[i] = i < n ? q[i] : 0x7fffffff;
    __syncthreads();
    for (int size = 2; size <= m; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const int a = key[i], b = key[j];
                    if ((a > b) == up) {
                        key[i] = b;
                        key[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) q[i] = key[i];
}

// Cartesian areas of the queued epochs, in three passes (small slots; large
// slots for the largest epochs, beside the small pass; large slots for the
// epochs that outgrew a small slot).  A workgroup is a team of kTeamWaves
// waves (Team above), each with its own slot blockIdx.x * kTeamWaves + w.
// Every wave loops: help a teammate if a job is offered, else run its own
// epoch, whose images it computes itself or takes from the teammate that
// computed them, accounting them in seed order.
__global__ void __launch_bounds__(kTeamWaves * kWave, 3) k_ir_cart(const CurveRec* __restrict__ curves,
                                                                FsOptions opt, FsEpochs ep, IrLists L,
                                                                CartScratch scr, int pass)
{
    __shared__ Team team;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    const int slot = static_cast<int>(blockIdx.x) * kTeamWaves + w;
    for (int k = 0; k < kTeamJobs; ++k) {
        wave_store(&team.item[w][k], 0);
        wave_store(&team.seed[w][k], 0);
        wave_store(&team.status[w][k], static_cast<int>(kJobFree));
    }
    if (w == 0) {
        wave_store(&team.active, kTeamWaves);
        wave_store(&team.ready, 0);
    }
    __syncthreads();

    const bool large = pass != kPassSmall;
    const int* queue = pass == kPassSmall ? L.cart_q : pass == kPassBig ? L.cart_big : L.cart_retry;
    const int qword = pass == kPassSmall ? kQCartCount : pass == kPassBig ? kQCartBigCount : kQCartRetryCount;
    unsigned* cursor = &ep.counters[qword + 1];
    int count = uni(static_cast<int>(ep.counters[qword]));
    // the first gridDim.x items (the largest) go one per team, to its wave 0,
    // so that each has teammates running smaller epochs, free to help; all
    // other items come from the queue cursor
    int reserved = count < static_cast<int>(gridDim.x) ? count : static_cast<int>(gridDim.x);
    int first_item = (w == 0 && static_cast<int>(blockIdx.x) < reserved) ? static_cast<int>(blockIdx.x) : -1;
    // the other waves take their first item once wave 0 has offered the images
    // of its (reserved) first epoch
    bool waited = w == 0;
    double* areaimage = scr.areaimage + static_cast<size_t>(slot) * kSeedCap;
    int* flags = scr.overlap + static_cast<size_t>(slot) * kSeedCap;
    const Rows mine = slot_rows(scr, slot);
    Ctx C;
    Cart K;
    bool have = false;       // an own epoch is in progress
    bool exhausted = false;  // the queue is empty
    int item = -1, nbin = 1, cur = -1, next_offer = 0;
    int64_t e = 0;
    long long t_begin = 0, idle_from = 0;
    double area = 0.0;
    for (;;) {
        int mode = 1;  // 1: own image `cur`, 2: a teammate's job, 3: helper h's image `cur`
        int p = -1, jk = -1, h = -1;
        if (!waited) {
            const long long t0 = static_cast<long long>(wall_clock64());
            while (lds_get(&team.ready) == 0 && static_cast<long long>(wall_clock64()) - t0 < kTeamWaitTicks)
                team_sleep();
            waited = true;
        }
        if (!have) {
            if (team_claim(team, w, p, jk)) {
                mode = 2;
            } else if (!exhausted) {
                int qi = first_item;
                if (qi >= 0)
                    first_item = -1;
                else
                    qi = reserved + next_item(cursor);
                if (qi >= count) {
                    exhausted = true;
                    if (w == 0) lds_put(&team.ready, 1);
                    atomicAdd(&team.active, lane_id() == 0 ? -1 : 0);
                    idle_from = static_cast<long long>(wall_clock64());
                    continue;
                }
                item = uni(queue[qi]);
                e = L.ir_sorted[item];
                t_begin = static_cast<long long>(wall_clock64());
                const CurveRec cv = curves[ep.curve_of[e]];
                init_ctx(C, cv, ep.src[e], L.seed_pool + uni(L.seed_off[item]));
                C.budget = opt.step_budget;
                C.n_seeds = uni(L.seed_n[item]);
                C.row_cap = scr.row_cap;
                nbin = curve_nbin(cv, opt);
                cart_params(K, C, nbin);
                wave_sync();
                for (int k = C.lane; k < C.n_seeds; k += kWave) {
                    areaimage[k] = 0.0;
                    flags[k] = 0;  // bit 0: overlap, bit 1: subtracted
                }
                wave_sync();
                area = 0.0;
                cur = -1;
                next_offer = 0;
                have = true;
                // a reserved item is among the largest: its later images are
                // offered at once (the first is the owner's)
                if (qi < reserved) team_offer(team, w, item, flags, 0, C.n_seeds, next_offer);
                if (w == 0) lds_put(&team.ready, 1);
                continue;
            } else {
                if (lds_get(&team.active) == 0) break;
                if (static_cast<long long>(wall_clock64()) - idle_from > kTeamWaitTicks) break;
                team_sleep();
                continue;
            }
        } else {
            int i = cur + 1;
            while (i < C.n_seeds && (uni(flags[i]) & 1)) ++i;
            const bool degenerate = C.n_seeds == 0 || !(C.rho > 0.0);
            if (i >= C.n_seeds || C.overflow || degenerate) {
                // the epoch's result
                const double mag = (C.overflow || degenerate) ? __builtin_nan("") : area / cart_scale(C, nbin);
                for (int k = 0; k < kTeamJobs; ++k) team_drop(team, w, k);
                if (C.cap_overflow && !large) {
                    const int pos = wave_fetch_inc(&ep.counters[kQCartRetryCount]);
                    wave_store(&L.cart_retry[pos], item);
                } else {
                    wave_store(&ep.mag[e], mag);
                    wave_store(&ep.err[e], isfinite(mag) ? cart_error_estimate(C, nbin, mag) : __builtin_nan(""));
                    if (ep.cost) {
                        const int64_t n = ep.n;
                        wave_store(&ep.cost[5 * n + e], static_cast<int64_t>(C.d_processed));
                        wave_store(&ep.cost[6 * n + e], static_cast<int64_t>(C.d_fold));
                        wave_store(&ep.cost[7 * n + e], static_cast<int64_t>(C.d_boundary_rows));
                        wave_store(&ep.cost[8 * n + e], static_cast<int64_t>(C.d_gaps));
                        wave_store(&ep.cost[9 * n + e], static_cast<int64_t>(C.d_overlaps));
                        wave_store(&ep.cost[10 * n + e], static_cast<int64_t>(__double_as_longlong(C.d_max_jump)));
                    }
                    if (C.overflow) atomicOr(&ep.counters[kQFlags], C.cap_overflow ? kErrRowCap : kErrBudget);
                }
                flush_work(C, ep);
                const long long ticks = static_cast<long long>(wall_clock64()) - t_begin;
                add_total(&work_totals(ep)[6], ticks);
                if (ep.cost) {
                    wave_store(&ep.cost[ep.n + e], ep.cost[ep.n + e] + static_cast<int64_t>(ticks));
                    wave_store(&ep.cost[3 * ep.n + e], ep.cost[3 * ep.n + e] + static_cast<int64_t>(C.n_rows));
                    wave_store(&ep.cost[4 * ep.n + e], ep.cost[4 * ep.n + e] + static_cast<int64_t>(C.steps));
                }
                if (ep.sched) wave_store(&ep.sched[e], ep.sched[e] + static_cast<float>(ticks));
                wave_sync();
                have = false;
                continue;
            }
            cur = i;
            const int k = team_find(team, w, i);
            if (k >= 0 && !lds_cas(&team.status[w][k], kJobOpen, kJobFree)) {
                // a teammate has taken this image: wait for its result
                const long long t0 = static_cast<long long>(wall_clock64());
                for (;;) {
                    const int s = lds_get(&team.status[w][k]);
                    if (s >= kJobDone && s < kJobCancel) {
                        h = s - kJobDone;
                        break;
                    }
                    if (static_cast<long long>(wall_clock64()) - t0 > kTeamWaitTicks) break;
                    team_sleep();
                }
                if (h < 0) {
                    C.overflow = true;
                    atomicOr(&ep.counters[kQFlags], kErrTeam);
                    team_drop(team, w, k);
                    continue;
                }
                jk = k;
                mode = 3;
            }
            if (mode == 1 && C.steps >= kOfferSteps) team_offer(team, w, item, flags, cur, C.n_seeds, next_offer);
        }
        ImageRun img{__builtin_nan(""), 0, 0, 0.0};
        Rows R = mine;
        if (mode != 3) {
            int j = cur;
            ImageHook hk{&team, w, nullptr, 0, flags, item, cur, C.n_seeds, &next_offer, mode == 1};
            if (mode == 2) {
                // the teammate's epoch, in this wave's context and slot; the
                // job's words are read while it is still this wave's (an entry
                // is rewritten only after its status has changed)
                const int pitem = lds_get(&team.item[p][jk]);
                j = lds_get(&team.seed[p][jk]);
                const int64_t pe = L.ir_sorted[pitem];
                const CurveRec cv = curves[ep.curve_of[pe]];
                init_ctx(C, cv, ep.src[pe], L.seed_pool + uni(L.seed_off[pitem]));
                C.budget = opt.step_budget;
                C.n_seeds = uni(L.seed_n[pitem]);
                C.row_cap = scr.row_cap;
                cart_params(K, C, curve_nbin(cv, opt));
                hk.watch = &team.status[p][jk];
                hk.expect = kJobRun + w;
                hk.armed = false;
                if (j >= C.n_seeds) C.overflow = true;
            }
            {
                auto& pk = team.park[w];
                wave_store(&pk.area, area);
                wave_store(&pk.e, static_cast<long long>(e));
                wave_store(&pk.t_begin, t_begin);
                wave_store(&pk.idle_from, idle_from);
                wave_store(&pk.item, item);
                wave_store(&pk.nbin, nbin);
                wave_store(&pk.first_item, first_item);
                wave_store(&pk.count, count);
                wave_store(&pk.reserved, reserved);
                wave_store(&pk.bits, (have ? 1 : 0) | (exhausted ? 2 : 0) | (waited ? 4 : 0));
            }
            if (!C.overflow) img = image_area(C, K, R, C.seeds[j], hk);
            {
                const auto& pk = team.park[w];
                area = uni(pk.area);
                e = uni(pk.e);
                t_begin = uni(pk.t_begin);
                idle_from = uni(pk.idle_from);
                item = uni(pk.item);
                nbin = uni(pk.nbin);
                first_item = uni(pk.first_item);
                count = uni(pk.count);
                reserved = uni(pk.reserved);
                const int bits = uni(pk.bits);
                have = (bits & 1) != 0;
                exhausted = (bits & 2) != 0;
                waited = (bits & 4) != 0;
            }
            if (mode == 2) {
                TeamOut& o = team.out[w];
                wave_store(&o.area, img.area);
                wave_store(&o.nyi, img.nyi);
                wave_store(&o.gaps, img.gaps);
                wave_store(&o.max_jump, img.max_jump);
                wave_store(&o.bad, (C.overflow ? 1 : 0) | (C.cap_overflow ? 2 : 0));
                wave_store(&o.steps, C.steps);
                wave_store(&o.n_cart, C.n_cart);
                wave_store(&o.n_cart_guarded, C.n_cart_guarded);
                wave_store(&o.n_rows, C.n_rows);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (lds_cas(&team.status[p][jk], kJobRun + w, kJobDone + w)) {
                    // the owner reads this slot's rows until it sets the word back
                    const long long t0 = static_cast<long long>(wall_clock64());
                    while (lds_get(&team.status[p][jk]) == kJobDone + w) {
                        if (static_cast<long long>(wall_clock64()) - t0 > kTeamWaitTicks) {
                            atomicOr(&ep.counters[kQFlags], kErrTeam);
                            return;
                        }
                        team_sleep();
                    }
                }
                continue;
            }
            if (C.overflow) continue;
        } else {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const TeamOut& o = team.out[h];
            img.area = uni(o.area);
            img.nyi = uni(o.nyi);
            img.gaps = uni(o.gaps);
            img.max_jump = uni(o.max_jump);
            const int bad = uni(o.bad);
            C.steps = uni(C.steps + uni(o.steps));
            C.n_cart = uni(C.n_cart + uni(o.n_cart));
            C.n_cart_guarded = uni(C.n_cart_guarded + uni(o.n_cart_guarded));
            C.n_rows = uni(C.n_rows + uni(o.n_rows));
            if (C.steps > C.budget || (bad & 1)) C.overflow = true;
            if (bad & 2) C.cap_overflow = true;
            R = slot_rows(scr, static_cast<int>(blockIdx.x) * kTeamWaves + h);
            if (C.overflow) {
                team_drop(team, w, jk);
                continue;
            }
        }
        account_image(C, K, cur, R, img, nbin, area, areaimage, flags, &team, w);
        if (mode == 3) team_drop(team, w, jk);
    }
}

// ------------------------------------------------ adaptive Cartesian epochs
//
// fixed_inverse_ray_binary's Cartesian branch in adaptive mode
// (finite_source_magnifier.cpp:3071-3296): the area scan records the
// refinement tape; an epoch whose scan shows the global topology pattern is
// scanned again at a finer lattice (needs_global_topology_retry, :1108-1121;
// :3166-3196); then the tape's cells are refined, largest error first, until
// the summed error estimate meets the tolerance.  One wave per epoch: the
// queue, the sums and the decisions are the wave's uniform control program;
// the twenty weights of each split (four child centres, four corners each) are
// evaluated one per lane.

// Store of the refinement queue: every lane writes the uniform value.
struct WaveStore {
    template <class T>
    __device__ void operator()(T* p, T v) const { *p = uni(v); }
};

// cell_weight (:3090-3099): brightness of the source point the image point
// maps to, zero outside the disk.
__device__ __forceinline__ double refine_weight(const Ctx& C, double x, double y)
{
    const double md = mapped_dist2(C.L, x, y, C.src);
    if (md > C.r2) return 0.0;
    return C.uniform ? 1.0 : ld_brightness(md / C.r2, C.ld_c, C.ld_d);
}

// estimate_local_cell_error (:3100-3128) from the four corner weights
// (corner order (-,-), (+,-), (-,+), (+,+)).
__device__ __forceinline__ double corner_error(double w0, double w1, double w2, double w3, double coarse, double frac)
{
    double sum = 0.0;
    double mn = __builtin_inf();
    double mx = 0.0;
    int nz = 0;
    const double w[4] = {w0, w1, w2, w3};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        sum += w[k];
        mn = w[k] < mn ? w[k] : mn;
        mx = mx < w[k] ? w[k] : mx;
        nz += w[k] > 0.0 ? 1 : 0;
    }
    const double avg = 0.25 * sum;
    const double center = fabs(avg - coarse);
    const double variation = 0.5 * (mx - mn);
    const double boundary = (nz > 0 && nz < 4) ? 0.25 * mx : 0.0;
    double m = center;
    if (m < variation) m = variation;
    if (m < boundary) m = boundary;
    return frac * m;
}

struct Refined {
    double mag, err;
    int level;
    bool full;  // the cell pool ran out (the epoch fails)
};

// The local refinement loop (:3201-3289) over the tape's n cells of side
// `incr`, after an area scan that gave `mag` with error estimate `base_err`
// (diagnostics.estimated_error) at `active` bins per radius; `nbin` is the
// curve's own setting.  Cells are kept in the slot's pool, the queue in
// (hk, hv).
__device__ Refined refine_cells(const Ctx& C, const FsOptions& opt, const Tape& T, double incr, int nbin, int active,
                                double mag, double base_err, int level, double* cx, double* cy, double* ccoarse,
                                double* cerr, int* cdepth, int cell_cap, double* hk, int* hv)
{
    Refined out{mag, base_err, level, false};
    const int n = T.n;
    const double rho = C.rho;
    const double scale = source_flux(rho, C.ld_c, C.ld_d) / (rho * rho) *
                         static_cast<double>(active) * static_cast<double>(active);
    const double safety = nbin < 32 ? 8.0 : rho >= 8.0e-3 ? 2.5 : rho >= 3.0e-3 ? 1.8 : 1.25;
    auto guarded = [&](double m) { return (opt.tol + opt.reltol * dmax(fabs(m), 1.0)) / safety; };
    const int max_bins = active > opt.max_source_bins ? active : opt.max_source_bins;
    int max_depth = 0;
    for (int b = active > 1 ? active : 1; b < max_bins && max_depth < 8; b *= 2) ++max_depth;
    const int max_splits = 4 * max_bins > 64 ? 4 * max_bins : 64;
    if (n + 4 * max_splits > cell_cap) {
        out.full = true;
        return out;
    }
    // the tape's cells: corner weights four lanes per cell, sixteen cells a step
    {
        const double half = 0.5 * incr;
        for (int base = 0; base < n; base += kWave / 4) {
            const int i = base + C.lane / 4;
            const int k = C.lane & 3;
            double w = 0.0;
            if (i < n) {
                const double x = T.x[i], y = T.y[i];
                const double px = (k & 1) ? x + half : x - half;
                const double py = (k & 2) ? y + half : y - half;
                w = refine_weight(C, px, py);
            }
            const double w1 = shfl(w, (C.lane & ~3) + 1);
            const double w2 = shfl(w, (C.lane & ~3) + 2);
            const double w3 = shfl(w, (C.lane & ~3) + 3);
            if (i < n && k == 0) {
                cx[i] = T.x[i];
                cy[i] = T.y[i];
                ccoarse[i] = T.w[i];
                cerr[i] = corner_error(w, w1, w2, w3, T.w[i], 1.0);
                cdepth[i] = 0;
            }
        }
    }
    wave_sync();
    PqHeap<WaveStore> H{hk, hv, 0, WaveStore{}};
    double lec = 0.0;  // local_error_counts
    for (int i = 0; i < n; ++i) {
        const double e = uni(cerr[i]);
        lec += e;
        H.push(e, i);
        H.n = uni(H.n);
    }
    wave_sync();
    double corr = 0.0;  // local_correction_counts
    int splits = 0;
    int pool = n;
    while (H.n > 0 && lec > guarded(mag + corr / scale) * scale && splits < max_splits) {
        double pk;
        const int pi = uni(H.pop(&pk));
        H.n = uni(H.n);
        wave_sync();
        const double perr = uni(cerr[pi]);
        const int pdepth = uni(cdepth[pi]);
        if (perr <= 0.0 || pdepth >= max_depth) break;
        const double px = uni(cx[pi]), py = uni(cy[pi]), pcoarse = uni(ccoarse[pi]);
        const double size = ldexp(incr, -pdepth);        // 0.5 x size, depth times
        const double pfrac = ldexp(1.0, -2 * pdepth);    // 0.25 x area_fraction, depth times
        const double offset = 0.25 * size;
        const double csize = 0.5 * size;
        const double cfrac = 0.25 * pfrac;
        const double half = 0.5 * csize;
        // lanes 0-3: child centres; lane 4 + 4 j + k: corner k of child j
        double w = 0.0;
        if (C.lane < 20) {
            const int j = C.lane < 4 ? C.lane : (C.lane - 4) >> 2;
            const double ccx = (j & 1) ? px + offset : px - offset;
            const double ccy = (j & 2) ? py + offset : py - offset;
            double x = ccx, y = ccy;
            if (C.lane >= 4) {
                const int k = (C.lane - 4) & 3;
                x = (k & 1) ? ccx + half : ccx - half;
                y = (k & 2) ? ccy + half : ccy - half;
            }
            w = refine_weight(C, x, y);
        }
        double coarse[4], err[4], ccxs[4], ccys[4];
        double contribution = 0.0, child_error = 0.0;
        int worst = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            coarse[j] = read_lane(w, j);
            err[j] = corner_error(read_lane(w, 4 + 4 * j), read_lane(w, 5 + 4 * j), read_lane(w, 6 + 4 * j),
                                  read_lane(w, 7 + 4 * j), coarse[j], cfrac);
            ccxs[j] = (j & 1) ? px + offset : px - offset;
            ccys[j] = (j & 2) ? py + offset : py - offset;
            contribution += cfrac * coarse[j];
            child_error += err[j];
            if (err[j] > err[worst]) worst = j;
        }
        corr += contribution - pfrac * pcoarse;
        lec += child_error - perr;
        out.level = out.level > pdepth + 1 ? out.level : pdepth + 1;
        ++splits;
        wave_sync();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            wave_store(&cx[pool + j], ccxs[j]);
            wave_store(&cy[pool + j], ccys[j]);
            wave_store(&ccoarse[pool + j], coarse[j]);
            wave_store(&cerr[pool + j], err[j]);
            wave_store(&cdepth[pool + j], pdepth + 1);
        }
        wave_sync();
        const bool all = lec <= guarded(mag + corr / scale) * scale || err[worst] <= 0.0 || pdepth + 1 >= max_depth;
        for (int j = 0; j < 4; ++j)
            if (all || j != worst) H.push(err[j], pool + j);
        if (!all) H.push(err[worst], pool + worst);
        H.n = uni(H.n);
        wave_sync();
        pool += 4;
    }
    if (out.level > 0) out.mag = mag + corr / scale;
    double e = safety * lec / scale;
    if (out.level == 0 && C.ld_c == 0.0 && C.ld_d == 0.0) e = e < base_err ? base_err : e;
    out.err = e;
    return out;
}

// Adaptive Cartesian epochs of one queue (the pass's queue of k_ir_cart), one
// wave per workgroup and slot; slot b of the kernel uses the Cartesian slot b
// of `scr` and the adaptive slot slot0 + b of `ad`.
__global__ void __launch_bounds__(kWave) k_ir_cart_adapt(const CurveRec* __restrict__ curves, FsOptions opt,
                                                          FsEpochs ep, IrLists L, CartScratch scr,
                                                          AdaptScratch ad, int slot0, int pass)
{
    extern __shared__ double adapt_lds[];
    const int slot = static_cast<int>(blockIdx.x);
    const int aslot = slot0 + slot;
    const bool large = pass != kPassSmall;
    const int* queue = pass == kPassSmall ? L.cart_q : pass == kPassBig ? L.cart_big : L.cart_retry;
    const int qword = pass == kPassSmall ? kQCartCount : pass == kPassBig ? kQCartBigCount : kQCartRetryCount;
    unsigned* cursor = &ep.counters[qword + 1];
    const int count = uni(static_cast<int>(ep.counters[qword]));
    const size_t tc = static_cast<size_t>(ad.tape_cap), cc = static_cast<size_t>(ad.cell_cap);
    double* tape = ad.tape + static_cast<size_t>(aslot) * tc * 3;
    long long* key = ad.key + static_cast<size_t>(aslot) * tc * 2;
    double* cell = ad.cell + static_cast<size_t>(aslot) * cc * 4;
    int* depth = ad.depth + static_cast<size_t>(aslot) * cc;
    double* hk = ad.heap_lds ? adapt_lds : ad.heap_key + static_cast<size_t>(aslot) * ad.heap_cap;
    int* hv = ad.heap_lds ? reinterpret_cast<int*>(adapt_lds + ad.heap_cap)
                          : ad.heap_val + static_cast<size_t>(aslot) * ad.heap_cap;
    Ctx C;
    for (;;) {
        const int qi = next_item(cursor);
        if (qi >= count) break;
        const int item = uni(queue[qi]);
        const int64_t e = L.ir_sorted[item];
        const long long t_begin = static_cast<long long>(wall_clock64());
        const CurveRec cv = curves[ep.curve_of[e]];
        init_ctx(C, cv, ep.src[e], L.seed_pool + uni(L.seed_off[item]));
        C.budget = opt.step_budget;
        C.n_seeds = uni(L.seed_n[item]);
        C.row_cap = scr.row_cap;
        const int nbin = curve_nbin(cv, opt);
        const bool adaptive = opt.adaptive != 0;
        const double rho = C.rho;
        Tape T;
        T.x = tape;
        T.y = tape + tc;
        T.w = tape + 2 * tc;
        T.kx = key;
        T.ky = key + tc;
        T.cap = adaptive ? (opt.max_source_bins > 64 ? opt.max_source_bins : 64) : 0;
        auto start_tape = [&](int bins) {
            const double incr = rho / static_cast<double>(bins > 1 ? bins : 1);
            T.n = 0;
            T.step = incr;
            const double a = 4.0 * rho * incr, b = 4.0 * incr * incr;
            T.band2 = a < b ? b : a;
        };
        start_tape(nbin);
        double mag = C.n_seeds > 0 ? cartesian_mag<true>(C, nbin, scr, slot, &T) : __builtin_nan("");
        double err = isfinite(mag) ? cart_error_estimate(C, nbin, mag) : __builtin_nan("");
        int active = nbin > 1 ? nbin : 1;
        int level = 0;
        // needs_global_topology_retry (:1108-1121) on the first scan's diagnostics
        if (!C.overflow && isfinite(mag) && opt.adaptive_source_bins != 0 && nbin < 50 &&
            opt.max_source_bins > nbin && rho < 1.0e-2 && C.n_seeds >= 16 && C.d_overlaps >= 100 &&
            C.d_gaps >= 1000 && C.d_max_jump >= 1000.0) {
            const int twice = 2 * (nbin > 1 ? nbin : 1);
            const int want = twice > 50 ? twice : 50;
            const int rbins = want < opt.max_source_bins ? want : opt.max_source_bins;
            if (rbins > nbin) {
                C.d_processed = 0;
                C.d_fold = 0;
                C.d_boundary_rows = 0;
                C.d_gaps = 0;
                C.d_overlaps = 0;
                C.d_max_jump = 0.0;
                start_tape(rbins);
                const double m2 = cartesian_mag<true>(C, rbins, scr, slot, &T);
                // the reference's retry is finite whenever the first scan is;
                // a retry that fails here (capacity, budget) fails the epoch
                mag = m2;
                err = isfinite(m2) ? cart_error_estimate(C, rbins, m2) : __builtin_nan("");
                active = rbins;
                level = 1;
            }
        }
        bool full = false;
        if (!C.overflow && isfinite(mag) && adaptive && T.n > 0) {
            const Refined r = refine_cells(C, opt, T, T.step, nbin, active, mag, err, level, cell, cell + cc,
                                           cell + 2 * cc, cell + 3 * cc, depth, ad.cell_cap, hk, hv);
            full = r.full;
            mag = r.mag;
            err = r.err;
            level = r.level;
        }
        if (C.cap_overflow && !large) {
            const int pos = wave_fetch_inc(&ep.counters[kQCartRetryCount]);
            wave_store(&L.cart_retry[pos], item);
        } else {
            const bool bad = C.overflow || full;
            wave_store(&ep.mag[e], bad ? __builtin_nan("") : mag);
            wave_store(&ep.err[e], bad ? __builtin_nan("") : err);
            wave_store(&ep.level[e], level | ((adaptive && active < 32) ? kLevelUnconverged : 0));
            if (ep.cost) {
                const int64_t n = ep.n;
                wave_store(&ep.cost[5 * n + e], static_cast<int64_t>(C.d_processed));
                wave_store(&ep.cost[6 * n + e], static_cast<int64_t>(C.d_fold));
                wave_store(&ep.cost[7 * n + e], static_cast<int64_t>(C.d_boundary_rows));
                wave_store(&ep.cost[8 * n + e], static_cast<int64_t>(C.d_gaps));
                wave_store(&ep.cost[9 * n + e], static_cast<int64_t>(C.d_overlaps));
                wave_store(&ep.cost[10 * n + e], static_cast<int64_t>(__double_as_longlong(C.d_max_jump)));
            }
            if (C.overflow) atomicOr(&ep.counters[kQFlags], C.cap_overflow ? kErrRowCap : kErrBudget);
            if (full) atomicOr(&ep.counters[kQFlags], kErrAdaptCap);
        }
        flush_work(C, ep);
        const long long ticks = static_cast<long long>(wall_clock64()) - t_begin;
        add_total(&work_totals(ep)[6], ticks);
        if (ep.cost) {
            wave_store(&ep.cost[ep.n + e], ep.cost[ep.n + e] + static_cast<int64_t>(ticks));
            wave_store(&ep.cost[3 * ep.n + e], ep.cost[3 * ep.n + e] + static_cast<int64_t>(C.n_rows));
            wave_store(&ep.cost[4 * ep.n + e], ep.cost[4 * ep.n + e] + static_cast<int64_t>(C.steps));
        }
        if (ep.sched) wave_store(&ep.sched[e], ep.sched[e] + static_cast<float>(ticks));
        wave_sync();
    }
}

// Polar flood fills of the queued epochs, in the same three passes.
__global__ void __launch_bounds__(kWave) k_ir_polar(const CurveRec* __restrict__ curves,
                                                     FsOptions opt, FsEpochs ep, IrLists L,
                                                     PolarScratch scr, int pass)
{
    int* head = scr.phi_head + static_cast<size_t>(blockIdx.x) * kPhiHash;
    int* runs = scr.runs + static_cast<size_t>(blockIdx.x) * scr.run_cap * 4;
    const bool large = pass != kPassSmall;
    const int* queue = pass == kPassSmall ? L.polar_q : pass == kPassBig ? L.polar_big : L.polar_retry;
    const int qword = pass == kPassSmall ? kQPolarCount : pass == kPassBig ? kQPolarBigCount : kQPolarRetryCount;
    unsigned* cursor = &ep.counters[qword + 1];
    const int count = uni(static_cast<int>(ep.counters[qword]));
    for (;;) {
        const int qi = next_item(cursor);
        if (qi >= count) break;
        const int item = uni(queue[qi]);
        const int64_t e = L.ir_sorted[item];
        const long long t_begin = static_cast<long long>(wall_clock64());
        const CurveRec cv = curves[ep.curve_of[e]];
        Ctx C;
        init_ctx(C, cv, ep.src[e], L.seed_pool + uni(L.seed_off[item]));
        C.budget = opt.step_budget;
        C.n_seeds = uni(L.seed_n[item]);
        C.run_cap = scr.run_cap;
        const double mag = polar_epoch(C, cv, opt, head, runs);
        if (C.cap_overflow && !large) {
            const int pos = wave_fetch_inc(&ep.counters[kQPolarRetryCount]);
            wave_store(&L.polar_retry[pos], item);
        } else {
            wave_store(&ep.mag[e], mag);
            // the polar method reports no error estimate (:3169-3175)
            wave_store(&ep.err[e], isfinite(mag) ? 0.0 : __builtin_nan(""));
            if (C.overflow) atomicOr(&ep.counters[kQFlags], C.cap_overflow ? kErrPolarCap : kErrBudget);
        }
        flush_work(C, ep);
        const long long ticks = static_cast<long long>(wall_clock64()) - t_begin;
        add_total(&work_totals(ep)[6], ticks);
        if (ep.cost) {
            wave_store(&ep.cost[ep.n + e], ep.cost[ep.n + e] + static_cast<int64_t>(ticks));
            wave_store(&ep.cost[3 * ep.n + e], ep.cost[3 * ep.n + e] + static_cast<int64_t>(C.n_rows));
            wave_store(&ep.cost[4 * ep.n + e], ep.cost[4 * ep.n + e] + static_cast<int64_t>(C.steps));
        }
        if (ep.sched) wave_store(&ep.sched[e], ep.sched[e] + static_cast<float>(ticks));
        wave_sync();
    }
}

// One epoch of the last run, on the first seed slot and the first large
// Cartesian / polar slot: with n_in < 0 the seeds are searched (phases up to
// last_phase) and written to seeds_io (*n_out = their count, up to kSeedCap);
// otherwise the n_in seeds of seeds_io are used.  *mag_out receives the
// inverse-ray magnification with method `method`.
__global__ void __launch_bounds__(kWave) k_ir_epoch_debug(const CurveRec* __restrict__ curves,
                                                          FsCache cache, FsTables tab,
                                                          FsOptions opt, FsEpochs ep,
                                                          IrScratch scr, int64_t e, int method,
                                                          int last_phase, Pt* seeds_io, int n_in,
                                                          int* n_out, double* mag_out)
{
    Pt* slot_seeds = scr.seed.seeds;
    const int c = ep.curve_of[e];
    const CurveRec cv = curves[c];
    Ctx C;
    init_ctx(C, cv, ep.src[e], slot_seeds);
    C.budget = opt.step_budget;
    C.hints = opt.ps_images != 0;
    if (last_phase >= 0) C.last_phase = last_phase;
    if (n_in < 0) {
        epoch_seeds(C, cache, tab, ep, c, e);
        wave_sync();
        for (int k = C.lane; k < C.n_seeds; k += kWave) seeds_io[k] = slot_seeds[k];
    } else {
        for (int k = C.lane; k < n_in; k += kWave) slot_seeds[k] = seeds_io[k];
        C.n_seeds = n_in;
    }
    wave_sync();
    wave_store(n_out, C.overflow ? -1 : C.n_seeds);
    double mag;
    if (method == LCBG_METHOD_INVERSE_RAY_POLAR) {
        C.run_cap = scr.polar_large.run_cap;
        mag = polar_epoch(C, cv, opt, scr.polar_large.phi_head, scr.polar_large.runs);
    } else {
        mag = cart_epoch(C, cv, opt, scr.cart_large, 0);
    }
    wave_store(mag_out, mag);
}

lcbg_status launch_fs_ir(const CurveRec* d_curves, const FsCache& cache, const FsTables& tab,
                         const FsOptions& opt, const FsEpochs& ep, const IrLists& lists,
                         const IrScratch& scr, hipStream_t st, const IrStreams& ss, bool farm)
{
    auto mark = [&](int k, hipStream_t s) {
        if (ss.kev) (void)hipEventRecord(ss.kev[k], s);
    };
    // `to` continues after the work enqueued so far on `from`
    auto fork = [&](hipEvent_t ev, hipStream_t from, hipStream_t to) -> lcbg_status {
        if (from == to) return LCBG_OK;
        lcbg_status r = map_hip(hipEventRecord(ev, from));
        if (r == LCBG_OK) r = map_hip(hipStreamWaitEvent(to, ev, 0));
        return r;
    };
    hipLaunchKernelGGL(k_ir_bucket_count, dim3(256), dim3(256), 0, st, d_curves, opt, ep, lists);
    hipLaunchKernelGGL(k_ir_bucket_scan, dim3(1), dim3(64), 0, st, ep);
    hipLaunchKernelGGL(k_ir_bucket_scatter, dim3(256), dim3(256), 0, st, ep, lists);
    mark(2 * kIrSeeds, st);
    // the leading items' seeds, then their areas on large slots (big stream)
    // while the other searches run
    hipLaunchKernelGGL(k_ir_seeds, dim3(static_cast<unsigned>(scr.seed.n_slots)), dim3(kWave), 0, st,
                       d_curves, cache, tab, opt, ep, lists, scr.seed, int{kSeedLead});
    hipLaunchKernelGGL(k_ir_sort_queues, dim3(2), dim3(1024), 0, st, ep, lists, 2);
    lcbg_status status = check_launch();
    if (status == LCBG_OK) status = fork(ss.fork_lead, st, ss.big);
    if (status != LCBG_OK) return status;
    // adaptive runs scan their Cartesian epochs with k_ir_cart_adapt, one wave
    // per slot, in the same passes
    const bool adapt = opt.adaptive_source_bins != 0;
    const unsigned adapt_lds = scr.adapt.heap_lds ? static_cast<unsigned>(scr.adapt.heap_cap) * 12u : 0u;
    auto cart_pass = [&](int pass, hipStream_t s) {
        const bool small = pass == kPassSmall;
        const CartScratch& cs = small ? scr.cart : scr.cart_large;
        if (adapt) {
            int n = small ? scr.adapt_small : scr.adapt_large;
            n = n < cs.n_slots ? n : cs.n_slots;  // the schedule may leave fewer Cartesian slots
            hipLaunchKernelGGL(k_ir_cart_adapt, dim3(static_cast<unsigned>(n)), dim3(kWave), adapt_lds, s,
                               d_curves, opt, ep, lists, cs, scr.adapt, small ? scr.adapt_large : 0, pass);
        } else {
            hipLaunchKernelGGL(k_ir_cart, dim3(static_cast<unsigned>(cs.n_slots / kTeamWaves)),
                               dim3(kTeamWaves * kWave), 0, s, d_curves, opt, ep, lists, cs, pass);
        }
    };
    mark(2 * kIrCartBig, ss.big);
    cart_pass(kPassBig, ss.big);
    mark(2 * kIrCartBig + 1, ss.big);
    status = check_launch();
    // the leading polar epochs on large slots (polar stream)
    if (status == LCBG_OK && ss.polar != st) status = map_hip(hipStreamWaitEvent(ss.polar, ss.fork_lead, 0));
    if (status != LCBG_OK) return status;
    mark(2 * kIrPolarBig, ss.polar);
    hipLaunchKernelGGL(k_ir_polar, dim3(static_cast<unsigned>(scr.polar_large.n_slots)), dim3(kWave), 0,
                       ss.polar, d_curves, opt, ep, lists, scr.polar_large, int{kPassBig});
    mark(2 * kIrPolarBig + 1, ss.polar);
    status = check_launch();
    if (status != LCBG_OK) return status;
    // all other searches; the deferred first-crossing scans: list, farm, resume
    hipLaunchKernelGGL(k_ir_seeds, dim3(static_cast<unsigned>(scr.seed.n_slots)), dim3(kWave), 0, st,
                       d_curves, cache, tab, opt, ep, lists, scr.seed, int{kSeedMain});
    if (farm) {
        hipLaunchKernelGGL(k_ir_heavy_list, dim3(kHeavyCap), dim3(kWave), 0, st, d_curves, cache, ep, lists,
                           scr.seed.heavy);
        hipLaunchKernelGGL(k_ir_farm, dim3(static_cast<unsigned>(scr.farm_slots)), dim3(kWave), 0, st,
                           d_curves, cache, tab, ep, lists, scr.seed.heavy);
        hipLaunchKernelGGL(k_ir_seeds, dim3(static_cast<unsigned>(scr.seed.n_slots)), dim3(kWave), 0, st,
                           d_curves, cache, tab, opt, ep, lists, scr.seed, int{kSeedResume});
    }
    hipLaunchKernelGGL(k_ir_sort_queues, dim3(2), dim3(1024), 0, st, ep, lists, 0);
    mark(2 * kIrSeeds + 1, st);
    status = check_launch();
    if (status == LCBG_OK) status = fork(ss.fork, st, ss.polar_small);
    if (status != LCBG_OK) return status;
    // the polar small-slot pass beside the Cartesian one and the polar
    // large-slot pass (its own stream: it need not wait for the largest epoch)
    mark(2 * kIrPolarSmall, ss.polar_small);
    hipLaunchKernelGGL(k_ir_polar, dim3(static_cast<unsigned>(scr.polar.n_slots)), dim3(kWave), 0,
                       ss.polar_small, d_curves, opt, ep, lists, scr.polar, int{kPassSmall});
    mark(2 * kIrPolarSmall + 1, ss.polar_small);
    status = check_launch();
    if (status != LCBG_OK) return status;
    mark(2 * kIrCartSmall, st);
    cart_pass(kPassSmall, st);
    mark(2 * kIrCartSmall + 1, st);
    status = check_launch();
    if (status == LCBG_OK) status = fork(ss.join_big, ss.big, st);
    if (status == LCBG_OK) status = fork(ss.join_polar, ss.polar, st);
    if (status == LCBG_OK) status = fork(ss.join_polar_small, ss.polar_small, st);
    if (status != LCBG_OK) return status;
    // the epochs that outgrew a small slot, on the (now free) large slots
    mark(2 * kIrCartRetry, st);
    cart_pass(kPassRetry, st);
    mark(2 * kIrCartRetry + 1, st);
    mark(2 * kIrPolarRetry, st);
    hipLaunchKernelGGL(k_ir_polar, dim3(static_cast<unsigned>(scr.polar_large.n_slots)), dim3(kWave), 0,
                       st, d_curves, opt, ep, lists, scr.polar_large, int{kPassRetry});
    mark(2 * kIrPolarRetry + 1, st);
    return check_launch();
}

// Resident workgroups per CU of the adaptive Cartesian kernel with `lds`
// bytes of dynamic LDS.
int adapt_resident_per_cu(unsigned lds)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(k_ir_cart_adapt), kWave,
                                                     lds) != hipSuccess || n < 1)
        n = 1;
    return n;
}

int ir_resident_per_cu(int kind)
{
    const void* fn = kind == 0 ? reinterpret_cast<const void*>(k_ir_seeds)
                   : kind == 1 ? reinterpret_cast<const void*>(k_ir_cart)
                   : kind == 2 ? reinterpret_cast<const void*>(k_ir_polar)
                               : reinterpret_cast<const void*>(k_ir_farm);
    const int threads = kind == 1 ? kTeamWaves * kWave : kWave;  // waves per workgroup
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, 0) != hipSuccess || n < 1) n = 1;
    return n * (threads / kWave);
}

lcbg_status launch_ir_epoch_debug(const CurveRec* d_curves, const FsCache& cache, const FsTables& tab,
                                  const FsOptions& opt, const FsEpochs& ep, const IrScratch& scr,
                                  int64_t e, int method, int last_phase, Pt2* seeds_io, int n_in,
                                  int* n_out, double* mag_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_ir_epoch_debug, dim3(1), dim3(kWave), 0, st, d_curves, cache, tab, opt,
                       ep, scr, e, method, last_phase, seeds_io, n_in, n_out, mag_out);
    return check_launch();
}

}  // namespace lcbg
