// Device-resident data of the finite-source pipeline, shared by the host
// runtime (which allocates it) and the kernels.
//
// Layout in HBM, per batch:
//   per curve   - caustic cache (caustic_points_at_phase points, tracked
//                 branches, grid cells, bounding box), the critical sources of
//                 the 1400-phase seed scan and their branch assignment;
//   per epoch   - source position, point-source magnification and roots,
//                 output magnification and method;
//   per slot    - scratch of one resident inverse-ray wave: seeds, Cartesian
//                 row records, polar visited runs.
#pragma once

#include <cstdint>

namespace lcbg {

struct Pt2 {
    double x, y;
};

constexpr int kSeedPhases = 1400;       // augmented_image_seeds samples (finite_source_magnifier.cpp:1742)
constexpr int kGrid = 128;              // caustic_cache_grid_size_ (finite_source_magnifier.hpp:129)
constexpr int kBoundaryProbes = 400;    // finite_source_magnifier.cpp:1577
constexpr int kAngularProbes = 16;      // finite_source_magnifier.cpp:1538
constexpr int kInteriorProbes = 64;     // finite_source_magnifier.cpp:1604

// Capacities of the per-slot scratch; exceeding one marks the epoch unsupported.
// Rows: a Cartesian epoch pushes one record per scanned row of every seed; a
// long arc near the Einstein ring at nbin 100 and rho 0.03 scans ~1e4 rows per
// seed.  Phi rows: the polar lattice has ceil(2 pi R / (dr g)) rows, about
// 2.1e6 at rho 1e-4, nbin 100, g 12 and R 4.
// Row keys and phi rows are found through hashed chains (kRowHash, kPhiHash
// heads), so no dense per-row array is needed.  Slots come in two tiers: many
// small ones for the first pass and a few large ones that redo the epochs
// which ran out of a small slot's rows or runs.
constexpr int kSeedCap = 1024;
constexpr int kRowCapSmall = 1 << 17;
constexpr int kRowCapLarge = 1 << 18;
constexpr int kRunCapSmall = 1 << 16;
constexpr int kRunCapLarge = 1 << 20;
constexpr int kRowHash = 1 << 16;
// Cartesian workgroups: waves per team, image jobs each wave can offer
constexpr int kTeamWaves = 4;
constexpr int kTeamJobs = 8;
constexpr int kPhiCap = 1 << 30;
constexpr int kPhiHash = 1 << 16;
// ints per slot of CartScratch::row_key: key, idx, next per row, the hash
// heads, and the count of keys whose heads are still set (padded to 64)
// Row-key nodes (8 ints each: key, row, next, pad, xmin, xmax as doubles),
// the chain heads, the dirty count.
constexpr long long row_ints_per_slot(int row_cap) { return 8ll * row_cap + kRowHash + 64; }
// cost buckets of the inverse-ray work order (largest a_ps * nbin^2 first)
constexpr int kCostBuckets = 32;

struct CausticScalars {
    double min_x, max_x, min_y, max_y, step_x, step_y, max_seg_len;
    int n;  // points per branch (0: cache empty)
    int pad;
};

struct FsTables {
    const Pt2* caus_trig;     // bins: cos, sin of 2 pi i / bins
    const Pt2* seed_trig;     // 1400: cos, sin of (2 pi / 1400) i
    const Pt2* boundary_trig; // 400: cos, sin of 2 pi i / 400
    const Pt2* angular_trig;  // 16
    const Pt2* interior_trig; // 64
};

struct FsCache {
    int bins;
    Pt2* pts;          // [curve][bins*4]   caustic points in root order
    Pt2* br;           // [curve][4*bins]   tracked branches
    int* pt_cell;      // [curve][bins*4]   grid cell (iy*128+ix) of each point
    int* br_cell;      // [curve][4*bins]
    int* br_order;     // [curve][4*bins]   branch positions sorted by (br_cell, index)
    int* br_row;       // [curve][129]      first br_order entry of each grid row; 4*bins
    Pt2* crit;         // [curve][1400*4]   critical sources of the seed scan
    uint8_t* assign;   // [curve][1400*4]   Phase-2 branch -> root index
    CausticScalars* sc;// [curve]
    int* near_n;       // [curve]            epochs stage 1 queued for stage 2 in the current run
};

struct FsOptions {
    int source_bins;
    int caustic_bins;
    int mode;
    int polar_source_bins;
    double grid_ratio;
    double polar_grid_ratio;
    double kinji;
    double hex_threshold;
    double adaptive_hex;
    // 1: the point-source magnification and center seeds of the LensModel path
    // (sum over binary_images, seeds filtered by append_valid_seed), used for
    // curves with parallax or orbital motion; 0: binary_mag0 of the static path
    int ps_images;
    // schedule (lcbg_schedule): inside positions a seed search probes before the
    // probe farm takes it over, and the leading items (largest cost key) whose
    // seeds are searched first and whose areas start on large slots at once
    int heavy_budget;
    int lead_items;
    // wave steps one inverse-ray epoch may take before it is abandoned
    // (lcbg_batch_set_work_budget; default kStepBudget)
    long long step_budget;
    // adaptive inverse ray (FiniteSourceSettings::adaptive_source_bins,
    // max_source_bins, finite_source_tol / reltol): adaptive_source_bins != 0
    // allows the topology retry at a finer lattice; `adaptive` (it and a
    // positive tolerance) also widens the bounding-box margin, derives the
    // point-source and hexadecapole thresholds from the tolerances, and refines
    // the Cartesian areas locally (finite_source_magnifier.cpp:3071-3296, 4679-4790)
    int adaptive_source_bins;
    int max_source_bins;
    int adaptive;
    double tol, reltol;
};

// Bounding-box margin of the dispatcher (finite_source_magnifier.cpp:4678-4685).
#if defined(__HIPCC__)
__host__ __device__
#endif
inline double bbox_margin(const FsOptions& o, double rho)
{
    const double m = o.kinji * rho;
    if (!o.adaptive) return m;
    const double wide = 60.0 * rho;
    return m < wide ? wide : m;
}

struct FsEpochs {
    int64_t n;
    const int64_t* offsets;
    const double* times;
    double* mag;
    int32_t* method;
    int32_t* curve_of;  // [epoch]
    Pt2* src;           // [epoch]
    double* a_ps;       // [epoch]
    double* roots;      // [epoch][10]
    double* dist;       // [epoch] refined caustic distance (IR epochs)
    int64_t* near_list;
    int64_t* ir_list;
    int64_t* hex_list;  // epochs queued for the hexadecapole test
    int64_t* near_curve;// [epoch] each curve's queued epochs in order, from the curve's
                        // first epoch on (FsCache::near_n of them)
    int64_t* near_chunks;// [chunk] (curve << 32) | first index of up to kNearChunk queued
                         // epochs in near_curve, one entry per chunk
    unsigned* counters; // [0] near count, [1] near cursor, [2] ir count, [3] ir cursor, [4] error flags
    int64_t* cost;      // [epoch] wall-clock ticks of each inverse-ray epoch, or null
    float* sched;       // [epoch] ticks of the epoch's seed search and area in the last
                        // run (0: none), the cost key of the next run's queues; or null
    double* dint;       // [epoch][6] tangents of the magnification (gradient runs)
    double* err;        // [epoch] error estimate of the method that settled the epoch
                        // (FiniteSourceResult::error_estimate, finite_source_magnifier.hpp:47-54)
    int32_t* level;     // [epoch] adaptive Cartesian epochs: refinement level, plus
                        // kLevelUnconverged when fewer than 32 bins were active
    const double* src_y;// [epoch] the impact parameter in place of the curve's u0, or null
                        // (curves with t0 = 0, tE = 1, alpha = 0 whose source positions
                        // come from a trajectory: times are x, src_y is y)
};

// Scratch of the resident waves of each inverse-ray kernel.
// Seed searches whose first-crossing scan probes more than kHeavyBudget
// caustic positions without a change hand the rest of the scan to the probe
// farm: every probe of the remaining inside positions is solved across the
// whole device, then the scan resumes and solves again only the probes that
// found more than three images (the only ones that can change the seeds).
constexpr int kHeavyCap = 256;
constexpr int kHeavyBudget = 16;
constexpr int kScanPositions = kSeedPhases * 4;             // (phase, root) positions
constexpr int kProbesPerPosition = 8 + 5 * kAngularProbes;  // caustic + angular probes
constexpr int kFarmChunk = 8;                               // positions per farm task
constexpr int kFarmChunks = (kScanPositions + kFarmChunk - 1) / kFarmChunk;

struct HeavyScratch {
    int* item;        // [kHeavyCap] inverse-ray item of the deferred search
    int* start;       // [kHeavyCap] first scan position left to probe
    int* n_list;      // [kHeavyCap] inside positions from `start` on
    uint16_t* list;   // [kHeavyCap][kScanPositions] those positions in scan order
    uint8_t* flags;   // [kHeavyCap][kScanPositions * kProbesPerPosition] probe found > 3 images
};

struct SeedScratch {
    int n_slots;
    Pt2* seeds;        // [slot][kSeedCap]
    HeavyScratch heavy;
};

struct CartScratch {
    int n_slots;
    int row_cap;
    double* areaimage; // [slot][kSeedCap]
    int* overlap;      // [slot][kSeedCap]
    double* rows;      // [slot][row_cap*5]: xmin, xmax, ax, y, dys
    int* row_key;      // [slot][row_ints_per_slot(row_cap)]: nodes, heads, dirty
};

// Workspace of the adaptive Cartesian kernel per slot: the refinement tape
// (cells recorded by the scans), the pool of refinement cells (the tape's, then
// four per split) and, when it does not fit in LDS, the refinement queue.
struct AdaptScratch {
    int tape_cap;      // max(64, max_source_bins)
    int cell_cap;      // tape_cap + 4 splits of the largest split budget
    int heap_cap;      // tape_cap + 3 splits + 1
    int heap_lds;      // 1: the queue lives in the kernel's dynamic LDS
    double* tape;      // [slot][tape_cap*3]: x, y, weight
    long long* key;    // [slot][tape_cap*2]: lattice keys
    double* cell;      // [slot][cell_cap*4]: centre x, y, coarse weight, error weight
    int* depth;        // [slot][cell_cap]
    double* heap_key;  // [slot][heap_cap] (heap_lds == 0)
    int* heap_val;     // [slot][heap_cap]
};

struct PolarScratch {
    int n_slots;
    int run_cap;
    int* phi_head;     // [slot][kPhiHash]: hashed run-chain heads
    int* runs;         // [slot][run_cap*4]: iphi, left, right, next
};

struct IrScratch {
    SeedScratch seed;
    CartScratch cart, cart_large;
    PolarScratch polar, polar_large;
    int farm_slots;  // resident waves of the probe farm
    // adaptive Cartesian kernel: its slots are the first adapt_large slots of
    // cart_large (adaptive slots [0, adapt_large)) and the first adapt_small
    // slots of cart (adaptive slots from adapt_large on); tape == null until
    // an adaptive run allocates them
    AdaptScratch adapt;
    int adapt_large, adapt_small;
};

// Work lists of the inverse-ray stage, in the epoch workspace.  Items index
// ir_sorted; each holds its seeds in the shared seed pool.
struct IrLists {
    int64_t* ir_sorted;  // [cap] inverse-ray epochs, largest estimated cost first
    int* bucket_of;      // [cap] cost bucket of ir_list[i]
    int* seed_off;       // [cap] first seed of item i in seed_pool
    int* seed_n;         // [cap] seed count of item i (-1: none, the epoch failed)
    int* cart_q;         // [cap] items for the Cartesian kernel
    int* polar_q;        // [cap] items for the polar kernel
    int* cart_retry;     // [cap] items that ran out of a small Cartesian slot
    int* polar_retry;    // [cap] items that ran out of a small polar slot
    int* cart_big;       // [cap] largest Cartesian items, large slots from the start
    int* polar_big;      // [cap] largest polar items
    int* grad_retry;     // [cap] items the gradient quadrature retries on large slots
    Pt2* seed_pool;      // [pool_cap]
    unsigned pool_cap;
};

// counters[kWorkOffset..] hold 8 uint64 work totals of the inverse-ray stage:
// quintic solves, quartic solves, Cartesian cells, guarded Cartesian cells,
// polar cells, seed-search and area ticks, rows; counters[kBucketOffset..]
// the cost-bucket counts and then their cursors.  All are reset each run.
// Queue words: near and inverse-ray epoch lists, error flags, the Cartesian
// and polar queues of the small-slot pass, of the large-slot pass for the
// largest epochs, and of the large-slot retry pass, the seed-pool cursor.
constexpr int kWorkOffset = 28;
constexpr int kBucketOffset = 44;
// counters[kFirstFailWord]: n - e of the first epoch e whose magnification is
// not finite (0: none), written by the run's last kernel.  Then the cursors of
// the gradient's run and chunk pools (GradWork).
constexpr int kFirstFailWord = kBucketOffset + 2 * kCostBuckets;
constexpr int kQGradRuns = kFirstFailWord + 1;
constexpr int kQGradChunks = kFirstFailWord + 2;
// the chunk tasks the enumeration publishes (reserved, taken) and the items it
// has finished (GradWork::tasks)
constexpr int kQGradTasks = kFirstFailWord + 3;
constexpr int kQGradTaskCursor = kFirstFailWord + 4;
constexpr int kQGradItemsDone = kFirstFailWord + 5;
constexpr int kCounterWords = kFirstFailWord + 6;
enum : int {
    kQNearCount = 0, kQNearCursor = 1, kQIrCount = 2, kQIrCursor = 3, kQFlags = 4,
    kQCartCount = 5, kQCartCursor = 6, kQPolarCount = 7, kQPolarCursor = 8,
    kQCartRetryCount = 9, kQCartRetryCursor = 10, kQPolarRetryCount = 11,
    kQPolarRetryCursor = 12, kQPoolCursor = 13, kQCartBigCount = 14, kQCartBigCursor = 15,
    kQPolarBigCount = 16, kQPolarBigCursor = 17, kQGradCursor = 18, kQGradRetryCount = 19,
    kQGradRetryCursor = 20, kQHeavyCount = 21, kQHeavyCursor = 22, kQFarmCursor = 23,
    kQLeadCursor = 24, kQHexCount = 25, kQNearChunkCount = 26, kQNearChunkCursor = 27
};
// derivatives per epoch: t0, tE, u0, alpha, s, q, rho, u_LD
constexpr int kGradParams = 8;
// value and six tangents of a magnification in the gradient kernels
constexpr int kGradSlots = 7;

// Work of the gradient's antialiased inverse-ray quadrature.  k_grad_ir
// enumerates each inverse-ray item's cells as runs along lattice rows and
// appends them to the run pool, with the 64-cell chunks that cover them to the
// chunk pool; k_grad_cells evaluates the weights of every chunk across the
// device; k_grad_finish sums each item's chunks in order.
struct GradWork {
    int* runs;         // [run_cap][4]: row, first cell, last cell, item
    int* run_cells;    // [run_cap] cells of the item's earlier runs
    int* chunk_item;   // [chunk_cap] item of each chunk
    double* partial;   // [chunk_cap][kGradSlots] weight sums of each chunk
    unsigned run_cap, chunk_cap;
    // per inverse-ray item, in the epoch workspace
    int* run_off;      // first run in the pool (-1: the item failed)
    int* run_n;
    int* chunk_off;    // first chunk in the pool
    int* cells;
    double* h;         // lattice cell side
    double* points;    // [item][kGradSlots] unresolved point images
    // chunk tasks: an enumerated item's chunks in groups of kGradTaskChunks,
    // published for the enumeration waves that have run out of items
    int* tasks;        // [task_cap][2]: first chunk, chunk count
    int* task_ready;   // [task_cap] 1 once published (cleared at the start of a run)
    uint8_t* chunk_done; // [chunk_cap] evaluated by a task, k_grad_cells skips it
                         // (cleared at the start of a run)
    unsigned task_cap;
};
constexpr int kGradTaskChunks = 16;
// queued epochs of one curve per stage-2 distance work item
constexpr int kNearChunk = 128;
// leading inverse-ray items searched first and scanned on large slots at once
constexpr int kLeadItems = 32;
// inverse-ray area passes: small slots, large slots for the largest epochs
// (beside the small pass), large slots for the epochs a small slot could not hold
enum : int { kPassSmall = 0, kPassBig = 1, kPassRetry = 2 };
// seed-search passes: all remaining searches, the deferred ones resumed, the leading items
enum : int { kSeedMain = 0, kSeedResume = 1, kSeedLead = 2 };

// error flags in counters[4]
constexpr unsigned kErrSeedCap = 1u;
constexpr unsigned kErrRowCap = 2u;
constexpr unsigned kErrPolarCap = 4u;
constexpr unsigned kErrUnsupported = 8u;
constexpr unsigned kErrBudget = 16u;
constexpr unsigned kErrSeedPool = 32u;
constexpr unsigned kErrTeam = 64u;  // a Cartesian team wait ran out of time
constexpr unsigned kErrGradPool = 128u;  // the gradient's run or chunk pool was full
constexpr unsigned kErrAdaptCap = 256u;  // the adaptive refinement's cell pool was full
// FsEpochs::level flag: an adaptive epoch scanned with fewer than 32 bins,
// never converged (finite_source_magnifier.cpp:3292-3294)
constexpr int kLevelUnconverged = 256;

// Work budget of one inverse-ray epoch in wave steps (64-cell chunks, probe
// merges, polar runs); an epoch that exceeds it ends with an error flag, so
// every resident wave finishes in bounded time.
constexpr long long kStepBudget = 1ll << 22;

}  // namespace lcbg
