// Host orchestration of the finite-source pipeline: per-batch device state
// (caustic caches, trig tables, epoch workspace, inverse-ray scratch) and the
// finite-source entry points of the C ABI.
//
// Compiled by g++: the phase and probe angle tables are rounded by glibc's
// cos/sin exactly as the reference's std::polar / std::cos / std::sin calls are
// (finite_source_magnifier.cpp:553, 1551-1553, 1579-1583, 1606-1610, 1748, 4025).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "fs_launch.hpp"
#include "traj.hpp"

namespace lcbg {

constexpr double kPiHost = 3.14159265358979323846;

struct FsState {
    FsCache cache{};
    FsTables tab{};
    FsOptions opt{};
    void* table_mem = nullptr;
    void* cache_mem = nullptr;
    // epoch workspace
    int64_t epoch_cap = 0;
    void* epoch_mem = nullptr;
    unsigned* counters = nullptr;
    int stage2_waves = 8192;  // waves of the per-epoch stage-2 launches (hexadecapole, derivative test)
    int64_t last_n = 0;  // epochs of the last run (their workspace is kept)
    double* last_mag = nullptr;     // the last run's output buffers (caller-owned)
    int32_t* last_method = nullptr;
    // per-epoch ticks of the last run, the next run's queue order (FsEpochs::sched)
    float* sched = nullptr;
    int64_t sched_cap = 0;
    int64_t sched_n = -1;  // epoch count they were measured on
    IrStreams ir{};  // side streams and events of the inverse-ray stage
    float ir_kernel_ms[kIrLaunches] = {};  // per-launch times of the last profiled run
};

namespace {

template <class T>
T* carve(char*& p, std::size_t count)
{
    const std::size_t align = 256;
    std::size_t off = reinterpret_cast<std::uintptr_t>(p) % align;
    if (off) p += align - off;
    T* out = reinterpret_cast<T*>(p);
    p += count * sizeof(T);
    return out;
}

std::size_t carve_size(std::size_t bytes) { return bytes + 256; }

constexpr int kCounterAlloc = 128;
constexpr int kTangentsHost = 6;  // tangent slots of the gradient kernels (dual.hpp)
static_assert(kCounterWords <= kCounterAlloc, "counter words");

// Seeds held at once by the inverse-ray stage: 8 per epoch of capacity, at
// least 2^20 (a typical inverse-ray epoch has 3-16 seeds and most epochs do
// not reach the inverse-ray stage at all).
int64_t pool_cap_for(int64_t cap)
{
    const int64_t v = 8 * cap;
    const int64_t lo = int64_t{1} << 20;
    const int64_t hi = (int64_t{1} << 31) - 1;
    return v < lo ? lo : (v > hi ? hi : v);
}

lcbg_status ensure_epochs(FsState* fs, int64_t n)
{
    if (n <= fs->epoch_cap && fs->epoch_mem) return LCBG_OK;
    if (fs->epoch_mem) {
        (void)hipDeviceSynchronize();
        (void)hipFree(fs->epoch_mem);
        fs->epoch_mem = nullptr;
    }
    int64_t cap = 1024;
    while (cap < n) cap *= 2;
    const std::size_t bytes = carve_size(cap * sizeof(int32_t)) + carve_size(cap * sizeof(Pt2)) +
                              carve_size(cap * sizeof(double)) * 2 +
                              carve_size(cap * 10 * sizeof(double)) +
                              carve_size(cap * sizeof(int64_t)) * 2 +
                              carve_size(kCounterAlloc * sizeof(unsigned)) +
                              carve_size(cap * sizeof(int64_t)) + carve_size(cap * sizeof(int)) * 10 +
                              carve_size(pool_cap_for(cap) * sizeof(Pt2)) +
                              carve_size(cap * kTangentsHost * sizeof(double)) +
                              carve_size(cap * sizeof(int64_t)) * 3 + carve_size(cap * sizeof(double)) +
                              carve_size(cap * sizeof(int)) * 4 + carve_size(cap * sizeof(double)) +
                              carve_size(cap * kGradSlots * sizeof(double)) + carve_size(cap * sizeof(int32_t));
    lcbg_status st = map_hip(hipMalloc(&fs->epoch_mem, bytes));
    if (st != LCBG_OK) {
        fs->epoch_mem = nullptr;
        return st;
    }
    fs->epoch_cap = cap;
    return LCBG_OK;
}

// The scheduling keys for an n-epoch run: kept from the last run over the same
// epoch count, zeroed (no key: the static estimate) otherwise.  A run being
// captured into a graph finds them prepared (sync = true before the capture).
lcbg_status ensure_sched(FsState* fs, int64_t n, hipStream_t stream, bool sync)
{
    if (n > fs->sched_cap || !fs->sched) {
        if (fs->sched) {
            (void)hipDeviceSynchronize();
            (void)hipFree(fs->sched);
            fs->sched = nullptr;
        }
        int64_t cap = 1024;
        while (cap < n) cap *= 2;
        lcbg_status st = map_hip(hipMalloc(&fs->sched, cap * sizeof(float)));
        if (st != LCBG_OK) {
            fs->sched = nullptr;
            fs->sched_cap = 0;
            return st;
        }
        fs->sched_cap = cap;
        fs->sched_n = -1;
    }
    if (fs->sched_n != n) {
        lcbg_status st = sync ? map_hip(hipMemset(fs->sched, 0, n * sizeof(float)))
                              : map_hip(hipMemsetAsync(fs->sched, 0, n * sizeof(float), stream));
        if (st != LCBG_OK) return st;
        fs->sched_n = n;
    }
    return LCBG_OK;
}

FsEpochs epoch_view(FsState* fs, int64_t n, const int64_t* offsets, const double* times, double* mag,
                    int32_t* method)
{
    char* p = static_cast<char*>(fs->epoch_mem);
    const int64_t cap = fs->epoch_cap;
    FsEpochs ep{};
    ep.n = n;
    ep.offsets = offsets;
    ep.times = times;
    ep.mag = mag;
    ep.method = method;
    ep.curve_of = carve<int32_t>(p, cap);
    ep.src = carve<Pt2>(p, cap);
    ep.a_ps = carve<double>(p, cap);
    ep.dist = carve<double>(p, cap);
    ep.roots = carve<double>(p, cap * 10);
    ep.near_list = carve<int64_t>(p, cap);
    ep.ir_list = carve<int64_t>(p, cap);
    ep.counters = carve<unsigned>(p, kCounterAlloc);
    return ep;
}

// The inverse-ray work lists, carved after the epoch arrays.
IrLists lists_view(FsState* fs)
{
    char* p = static_cast<char*>(fs->epoch_mem);
    const int64_t cap = fs->epoch_cap;
    (void)carve<int32_t>(p, cap);
    (void)carve<Pt2>(p, cap);
    (void)carve<double>(p, cap);
    (void)carve<double>(p, cap);
    (void)carve<double>(p, cap * 10);
    (void)carve<int64_t>(p, cap);
    (void)carve<int64_t>(p, cap);
    (void)carve<unsigned>(p, kCounterAlloc);
    IrLists L{};
    L.ir_sorted = carve<int64_t>(p, cap);
    L.bucket_of = carve<int>(p, cap);
    L.seed_off = carve<int>(p, cap);
    L.seed_n = carve<int>(p, cap);
    L.cart_q = carve<int>(p, cap);
    L.polar_q = carve<int>(p, cap);
    L.cart_retry = carve<int>(p, cap);
    L.polar_retry = carve<int>(p, cap);
    L.cart_big = carve<int>(p, cap);
    L.polar_big = carve<int>(p, cap);
    L.grad_retry = carve<int>(p, cap);
    L.pool_cap = static_cast<unsigned>(pool_cap_for(cap));
    L.seed_pool = carve<Pt2>(p, L.pool_cap);
    return L;
}

double* dint_view(FsState* fs)
{
    char* p = static_cast<char*>(fs->epoch_mem);
    const int64_t cap = fs->epoch_cap;
    (void)carve<int32_t>(p, cap);
    (void)carve<Pt2>(p, cap);
    (void)carve<double>(p, cap);
    (void)carve<double>(p, cap);
    (void)carve<double>(p, cap * 10);
    (void)carve<int64_t>(p, cap);
    (void)carve<int64_t>(p, cap);
    (void)carve<unsigned>(p, kCounterAlloc);
    (void)carve<int64_t>(p, cap);
    for (int k = 0; k < 10; ++k) (void)carve<int>(p, cap);
    (void)carve<Pt2>(p, pool_cap_for(cap));
    return carve<double>(p, cap * kTangentsHost);
}

// The hexadecapole queue of stage 2 and the per-curve queue of stage 1,
// after the tangents.
int64_t* hex_view(FsState* fs)
{
    char* p = reinterpret_cast<char*>(dint_view(fs));
    (void)carve<double>(p, fs->epoch_cap * kTangentsHost);
    return carve<int64_t>(p, fs->epoch_cap);
}

int64_t* near_curve_view(FsState* fs)
{
    char* p = reinterpret_cast<char*>(hex_view(fs));
    (void)carve<int64_t>(p, fs->epoch_cap);
    return carve<int64_t>(p, fs->epoch_cap);
}

// Work items of the stage-2 distances (at most one per queued epoch).
int64_t* near_chunks_view(FsState* fs)
{
    char* p = reinterpret_cast<char*>(near_curve_view(fs));
    (void)carve<int64_t>(p, fs->epoch_cap);
    return carve<int64_t>(p, fs->epoch_cap);
}

// Per-epoch error estimates, after the stage-2 work items.
double* err_view(FsState* fs)
{
    char* p = reinterpret_cast<char*>(near_chunks_view(fs));
    (void)carve<int64_t>(p, fs->epoch_cap);
    return carve<double>(p, fs->epoch_cap);
}

// Per-item fields of the gradient's quadrature work, after the error estimates.
void grad_item_view(FsState* fs, GradWork& w)
{
    char* p = reinterpret_cast<char*>(err_view(fs));
    const int64_t cap = fs->epoch_cap;
    (void)carve<double>(p, cap);
    w.run_off = carve<int>(p, cap);
    w.run_n = carve<int>(p, cap);
    w.chunk_off = carve<int>(p, cap);
    w.cells = carve<int>(p, cap);
    w.h = carve<double>(p, cap);
    w.points = carve<double>(p, cap * kGradSlots);
}

// Per-epoch refinement levels of adaptive runs, after the gradient item fields.
int32_t* level_view(FsState* fs)
{
    GradWork w{};
    grad_item_view(fs, w);
    char* p = reinterpret_cast<char*>(w.points);
    (void)carve<double>(p, fs->epoch_cap * kGradSlots);
    return carve<int32_t>(p, fs->epoch_cap);
}

// The status a run's outcome maps to: a method outside this library is
// unsupported, a team wait that ran out of time a device failure, and any
// other abandoned epoch (its magnification is NaN) a numerical error, as a
// non-finite magnification is in the reference (src/lcbinint/lcbinint.cpp:113-118).
lcbg_status status_of_run(unsigned flags, int64_t first_failed)
{
    if (flags & kErrUnsupported) return LCBG_UNSUPPORTED;
    if (flags & kErrTeam) return LCBG_DEVICE_ERROR;
    if (flags || first_failed >= 0) return LCBG_NUMERICAL_ERROR;
    return LCBG_OK;
}

// Inverse-ray scratch, one pool per device shared by every batch: its use is
// serialised by a mutex held while a run is enqueued and by an event that the
// next run waits for.
struct ScratchPool {
    std::mutex mu;
    IrScratch scr{};
    int cus = 0;
    void* mem = nullptr;
    hipEvent_t last_use = nullptr;
    GradWork grad{};       // pools of the gradient quadrature (first gradient run)
    PolarScratch grad_slots{}, grad_large{};  // its enumeration slots
    void* grad_mem = nullptr;
    void* adapt_mem = nullptr;  // adaptive Cartesian slots (first adaptive run)
};

// Runs and 64-cell chunks the gradient's inverse-ray quadrature may hold in one
// run over all its items: 2^24 each (1.3 GB).
constexpr unsigned kGradRunCap = 1u << 24;
constexpr unsigned kGradChunkCap = 1u << 24;
constexpr unsigned kGradTaskCap = 1u << 21;  // chunk tasks per run (more go to k_grad_cells)

lcbg_status ensure_grad_pool(ScratchPool& pool)
{
    if (pool.grad_mem) return LCBG_OK;
    const std::size_t r = kGradRunCap, c = kGradChunkCap;
    // enumeration slots: one per CU with 2^16 runs, 16 with 2^20
    pool.grad_slots.n_slots = pool.cus;
    pool.grad_slots.run_cap = kRunCapSmall;
    pool.grad_large.n_slots = 16;
    pool.grad_large.run_cap = kRunCapLarge;
    auto slot_bytes = [](const PolarScratch& ps) {
        const std::size_t n = static_cast<std::size_t>(ps.n_slots);
        return carve_size(n * kPhiHash * sizeof(int)) + carve_size(n * ps.run_cap * 4 * sizeof(int));
    };
    const std::size_t tc = kGradTaskCap;
    const std::size_t bytes = carve_size(r * 4 * sizeof(int)) + carve_size(r * sizeof(int)) +
                              carve_size(c * sizeof(int)) + carve_size(c * kGradSlots * sizeof(double)) +
                              carve_size(tc * 2 * sizeof(int)) + carve_size(tc * sizeof(int)) + carve_size(c) +
                              slot_bytes(pool.grad_slots) + slot_bytes(pool.grad_large);
    lcbg_status st = map_hip(hipMalloc(&pool.grad_mem, bytes));
    if (st != LCBG_OK) {
        pool.grad_mem = nullptr;
        return st;
    }
    char* p = static_cast<char*>(pool.grad_mem);
    GradWork& w = pool.grad;
    w.runs = carve<int>(p, r * 4);
    w.run_cells = carve<int>(p, r);
    w.chunk_item = carve<int>(p, c);
    w.partial = carve<double>(p, c * kGradSlots);
    w.tasks = carve<int>(p, tc * 2);
    w.task_ready = carve<int>(p, tc);
    w.chunk_done = carve<uint8_t>(p, c);
    w.run_cap = kGradRunCap;
    w.chunk_cap = kGradChunkCap;
    w.task_cap = kGradTaskCap;
    static_assert(kGradChunkCap % 16 == 0, "chunk flags are cleared 16 at a time");
    for (PolarScratch* ps : {&pool.grad_slots, &pool.grad_large}) {
        const std::size_t n = static_cast<std::size_t>(ps->n_slots);
        ps->phi_head = carve<int>(p, n * kPhiHash);
        ps->runs = carve<int>(p, n * ps->run_cap * 4);
        // empty run chains; the kernel restores them per item
        st = map_hip(hipMemset(ps->phi_head, 0xff, n * kPhiHash * sizeof(int)));
        if (st != LCBG_OK) return st;
    }
    return map_hip(hipDeviceSynchronize());
}

// The adaptive Cartesian kernel's workspace for refinement budgets up to
// `max_bins` (max of the curves' source_bins and max_source_bins) and tapes of
// up to max(64, max_source_bins) cells; grown (after a device-wide wait) when
// a run needs more.  Its queue lives in LDS when it fits in 64 KiB.
lcbg_status ensure_adapt_pool(ScratchPool& pool, int max_source_bins, int max_bins)
{
    const int tape_cap = max_source_bins > 64 ? max_source_bins : 64;
    const int splits = 4 * max_bins > 64 ? 4 * max_bins : 64;
    const int cell_cap = tape_cap + 4 * splits;
    const int heap_cap = tape_cap + 3 * splits + 1;
    AdaptScratch& a = pool.scr.adapt;
    if (pool.adapt_mem && a.tape_cap >= tape_cap && a.cell_cap >= cell_cap && a.heap_cap >= heap_cap)
        return LCBG_OK;
    if (pool.adapt_mem) {
        (void)hipDeviceSynchronize();
        (void)hipFree(pool.adapt_mem);
        pool.adapt_mem = nullptr;
    }
    const bool lds = static_cast<std::size_t>(heap_cap) * 12 <= 65536;
    const int large = pool.scr.cart_large.n_slots < 128 ? pool.scr.cart_large.n_slots : 128;
    int small = pool.cus * adapt_resident_per_cu(lds ? static_cast<unsigned>(heap_cap) * 12u : 0u);
    if (small > pool.scr.cart.n_slots) small = pool.scr.cart.n_slots;
    const std::size_t n = static_cast<std::size_t>(large + small);
    const std::size_t tc = static_cast<std::size_t>(tape_cap), cc = static_cast<std::size_t>(cell_cap),
                      hc = lds ? 0 : static_cast<std::size_t>(heap_cap);
    const std::size_t bytes = carve_size(n * tc * 3 * sizeof(double)) + carve_size(n * tc * 2 * sizeof(long long)) +
                              carve_size(n * cc * 4 * sizeof(double)) + carve_size(n * cc * sizeof(int)) +
                              carve_size(n * hc * sizeof(double)) + carve_size(n * hc * sizeof(int));
    lcbg_status st = map_hip(hipMalloc(&pool.adapt_mem, bytes));
    if (st != LCBG_OK) {
        pool.adapt_mem = nullptr;
        return st;
    }
    char* p = static_cast<char*>(pool.adapt_mem);
    a.tape_cap = tape_cap;
    a.cell_cap = cell_cap;
    a.heap_cap = heap_cap;
    a.heap_lds = lds ? 1 : 0;
    a.tape = carve<double>(p, n * tc * 3);
    a.key = carve<long long>(p, n * tc * 2);
    a.cell = carve<double>(p, n * cc * 4);
    a.depth = carve<int>(p, n * cc);
    a.heap_key = carve<double>(p, n * hc);
    a.heap_val = carve<int>(p, n * hc);
    pool.scr.adapt_large = large;
    pool.scr.adapt_small = small;
    return LCBG_OK;
}

constexpr int kMaxDevices = 64;
ScratchPool g_pool[kMaxDevices];

lcbg_status ensure_scratch(ScratchPool& pool)
{
    if (pool.mem) return LCBG_OK;
    int cus = 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    // as many slots as the kernel keeps resident: its occupancy per CU
    const int seed_slots = cus * ir_resident_per_cu(0);
    // two Cartesian waves per SIMD: the largest epochs bound the stage, and a
    // slot of 2^17 rows holds them without a second pass
    const int cart_occ = ir_resident_per_cu(1);
    const int cart_slots = cus * (cart_occ < 8 ? cart_occ : 8);
    const int polar_slots = cus;  // few polar epochs; they share the CUs with the Cartesian pass
    const int large_slots = 128;  // 32 Cartesian teams for the leading items
    IrScratch& s = pool.scr;
    pool.cus = cus;
    s.seed.n_slots = seed_slots;
    s.farm_slots = cus * ir_resident_per_cu(3);
    s.cart.n_slots = cart_slots;
    s.cart.row_cap = kRowCapSmall;
    s.cart_large.n_slots = large_slots;
    s.cart_large.row_cap = kRowCapLarge;
    s.polar.n_slots = polar_slots;
    s.polar.run_cap = kRunCapSmall;
    s.polar_large.n_slots = large_slots;
    s.polar_large.run_cap = kRunCapLarge;
    auto cart_bytes = [](const CartScratch& c) {
        const std::size_t n = static_cast<std::size_t>(c.n_slots);
        return carve_size(n * kSeedCap * sizeof(double)) + carve_size(n * kSeedCap * sizeof(int)) +
               carve_size(n * c.row_cap * 5 * sizeof(double)) +
               carve_size(n * static_cast<std::size_t>(row_ints_per_slot(c.row_cap)) * sizeof(int));
    };
    auto polar_bytes = [](const PolarScratch& c) {
        const std::size_t n = static_cast<std::size_t>(c.n_slots);
        return carve_size(n * kPhiHash * sizeof(int)) + carve_size(n * c.run_cap * 4 * sizeof(int));
    };
    const std::size_t heavy_bytes =
        carve_size(kHeavyCap * sizeof(int)) * 3 +
        carve_size(static_cast<std::size_t>(kHeavyCap) * kScanPositions * sizeof(uint16_t)) +
        carve_size(static_cast<std::size_t>(kHeavyCap) * kScanPositions * kProbesPerPosition);
    const std::size_t bytes = carve_size(static_cast<std::size_t>(seed_slots) * kSeedCap * sizeof(Pt2)) +
                              heavy_bytes + cart_bytes(s.cart) + cart_bytes(s.cart_large) +
                              polar_bytes(s.polar) + polar_bytes(s.polar_large);
    lcbg_status st = map_hip(hipMalloc(&pool.mem, bytes));
    if (st != LCBG_OK) {
        pool.mem = nullptr;
        return st;
    }
    char* p = static_cast<char*>(pool.mem);
    s.seed.seeds = carve<Pt2>(p, static_cast<std::size_t>(seed_slots) * kSeedCap);
    s.seed.heavy.item = carve<int>(p, kHeavyCap);
    s.seed.heavy.start = carve<int>(p, kHeavyCap);
    s.seed.heavy.n_list = carve<int>(p, kHeavyCap);
    s.seed.heavy.list = carve<uint16_t>(p, static_cast<std::size_t>(kHeavyCap) * kScanPositions);
    s.seed.heavy.flags = carve<uint8_t>(p, static_cast<std::size_t>(kHeavyCap) * kScanPositions *
                                               kProbesPerPosition);
    for (CartScratch* c : {&s.cart, &s.cart_large}) {
        const std::size_t n = static_cast<std::size_t>(c->n_slots);
        c->areaimage = carve<double>(p, n * kSeedCap);
        c->overlap = carve<int>(p, n * kSeedCap);
        c->rows = carve<double>(p, n * c->row_cap * 5);
        c->row_key = carve<int>(p, n * static_cast<std::size_t>(row_ints_per_slot(c->row_cap)));
        // empty row-key chains: -1 heads, dirty count -1 (nothing to clear);
        // the kernel restores them per seed
        st = map_hip(hipMemset(c->row_key, 0xff,
                               n * static_cast<std::size_t>(row_ints_per_slot(c->row_cap)) * sizeof(int)));
        if (st != LCBG_OK) return st;
    }
    for (PolarScratch* c : {&s.polar, &s.polar_large}) {
        const std::size_t n = static_cast<std::size_t>(c->n_slots);
        c->phi_head = carve<int>(p, n * kPhiHash);
        c->runs = carve<int>(p, n * c->run_cap * 4);
        // empty visited-run lists; the kernel restores them per epoch
        st = map_hip(hipMemset(c->phi_head, 0xff, n * kPhiHash * sizeof(int)));
        if (st != LCBG_OK) return st;
    }
    st = map_hip(hipEventCreateWithFlags(&pool.last_use, hipEventDisableTiming));
    if (st != LCBG_OK) return st;
    // record once so that the first run's stream wait has a completed event
    st = map_hip(hipEventRecord(pool.last_use, nullptr));
    if (st != LCBG_OK) return st;
    return map_hip(hipDeviceSynchronize());
}

}  // namespace

void angle_tables(int bins, std::vector<Pt2>& host)
{
    // each expression as the reference writes it
    host.clear();
    host.reserve(static_cast<std::size_t>(angle_table_count(bins)));
    for (int i = 0; i < bins; ++i) {
        const double phase = 2.0 * kPiHost * static_cast<double>(i) / static_cast<double>(bins);
        host.push_back({std::cos(phase), std::sin(phase)});
    }
    const double phase_step = 2.0 * kPiHost / static_cast<double>(kSeedPhases);
    for (int k = 0; k < kSeedPhases; ++k) {
        const double phi = phase_step * static_cast<double>(k);
        host.push_back({std::cos(phi), std::sin(phi)});
    }
    for (int i = 0; i < kBoundaryProbes; ++i) {
        const double phi = 2.0 * kPiHost * static_cast<double>(i) / static_cast<double>(kBoundaryProbes);
        host.push_back({std::cos(phi), std::sin(phi)});
    }
    for (int i = 0; i < kAngularProbes; ++i) {
        const double th = 2.0 * kPiHost * static_cast<double>(i) / static_cast<double>(kAngularProbes);
        host.push_back({std::cos(th), std::sin(th)});
    }
    for (int i = 0; i < kInteriorProbes; ++i) {
        const double phi = 2.0 * kPiHost * static_cast<double>(i) / static_cast<double>(kInteriorProbes);
        host.push_back({std::cos(phi), std::sin(phi)});
    }
}

FsTables table_view(const Pt2* t, int bins)
{
    FsTables tab;
    tab.caus_trig = t;
    tab.seed_trig = t + bins;
    tab.boundary_trig = tab.seed_trig + kSeedPhases;
    tab.angular_trig = tab.boundary_trig + kBoundaryProbes;
    tab.interior_trig = tab.angular_trig + kAngularProbes;
    return tab;
}

void fs_destroy(FsState* fs)
{
    if (!fs) return;
    if (fs->table_mem) (void)hipFree(fs->table_mem);
    if (fs->cache_mem) (void)hipFree(fs->cache_mem);
    if (fs->epoch_mem) (void)hipFree(fs->epoch_mem);
    if (fs->sched) (void)hipFree(fs->sched);
    if (fs->ir.polar) (void)hipStreamDestroy(fs->ir.polar);
    if (fs->ir.polar_small) (void)hipStreamDestroy(fs->ir.polar_small);
    if (fs->ir.big) (void)hipStreamDestroy(fs->ir.big);
    if (fs->ir.grad) (void)hipStreamDestroy(fs->ir.grad);
    for (hipEvent_t e : {fs->ir.fork, fs->ir.fork_lead, fs->ir.join_polar, fs->ir.join_polar_small, fs->ir.join_big,
                         fs->ir.cache_fork, fs->ir.cache_join, fs->ir.join_grad})
        if (e) (void)hipEventDestroy(e);
    delete fs;
}

// Allocates the trig tables and the per-curve cache storage of a batch; the
// caches themselves are built in-stream by fs_refresh.
lcbg_status fs_prepare(lcbg_batch* b)
{
    if (b->fs) return LCBG_OK;
    FsState* fs = new (std::nothrow) FsState();
    if (!fs) return LCBG_DEVICE_ERROR;
    b->fs = fs;
    const lcbg_options& o = b->options;
    fs->opt.source_bins = o.source_bins;
    fs->opt.caustic_bins = o.caustic_bins;
    fs->opt.mode = o.mode;
    fs->opt.polar_source_bins = o.polar_source_bins;
    fs->opt.grid_ratio = o.grid_ratio;
    fs->opt.polar_grid_ratio = o.polar_grid_ratio;
    fs->opt.kinji = o.point_source_threshold;
    fs->opt.hex_threshold = o.hexadecapole_threshold;
    fs->opt.adaptive_hex = o.adaptive_hex_threshold;
    fs->opt.adaptive_source_bins = o.adaptive_source_bins;
    fs->opt.max_source_bins = o.max_source_bins;
    fs->opt.tol = o.finite_source_tol;
    fs->opt.reltol = o.finite_source_reltol;
    fs->opt.adaptive = (o.adaptive_source_bins != 0 && (o.finite_source_tol > 0.0 || o.finite_source_reltol > 0.0))
                           ? 1 : 0;
    fs->opt.ps_images = b->dynamic_images ? 1 : 0;
    fs->opt.heavy_budget = kHeavyBudget;
    fs->opt.lead_items = kLeadItems;
    fs->opt.step_budget = kStepBudget;
    const int bins = o.caustic_bins > 32 ? o.caustic_bins : 32;

    std::vector<Pt2> host;
    angle_tables(bins, host);
    lcbg_status st = map_hip(hipMalloc(&fs->table_mem, host.size() * sizeof(Pt2)));
    if (st != LCBG_OK) return st;
    st = map_hip(hipMemcpy(fs->table_mem, host.data(), host.size() * sizeof(Pt2), hipMemcpyHostToDevice));
    if (st != LCBG_OK) return st;
    fs->tab = table_view(static_cast<const Pt2*>(fs->table_mem), bins);

    const int nc = b->n_curves > 0 ? b->n_curves : 1;
    const std::size_t per = static_cast<std::size_t>(nc);
    const std::size_t bytes = carve_size(per * bins * 4 * sizeof(Pt2)) * 2 +
                              carve_size(per * bins * 4 * sizeof(int)) * 3 +
                              carve_size(per * (kGrid + 1) * sizeof(int)) +
                              carve_size(per * kSeedPhases * 4 * sizeof(Pt2)) +
                              carve_size(per * kSeedPhases * 4) + carve_size(per * sizeof(CausticScalars)) +
                              carve_size(per * sizeof(int));
    st = map_hip(hipMalloc(&fs->cache_mem, bytes));
    if (st != LCBG_OK) return st;
    char* p = static_cast<char*>(fs->cache_mem);
    fs->cache.bins = bins;
    fs->cache.pts = carve<Pt2>(p, per * bins * 4);
    fs->cache.br = carve<Pt2>(p, per * bins * 4);
    fs->cache.pt_cell = carve<int>(p, per * bins * 4);
    fs->cache.br_cell = carve<int>(p, per * bins * 4);
    fs->cache.br_order = carve<int>(p, per * bins * 4);
    fs->cache.br_row = carve<int>(p, per * (kGrid + 1));
    fs->cache.crit = carve<Pt2>(p, per * kSeedPhases * 4);
    fs->cache.assign = carve<uint8_t>(p, per * kSeedPhases * 4);
    fs->cache.sc = carve<CausticScalars>(p, per);
    fs->cache.near_n = carve<int>(p, per);
    b->caches_stale = true;
    st = map_hip(hipStreamCreateWithFlags(&fs->ir.polar, hipStreamNonBlocking));
    if (st == LCBG_OK) st = map_hip(hipStreamCreateWithFlags(&fs->ir.polar_small, hipStreamNonBlocking));
    if (st == LCBG_OK) st = map_hip(hipStreamCreateWithFlags(&fs->ir.big, hipStreamNonBlocking));
    if (st == LCBG_OK) st = map_hip(hipStreamCreateWithFlags(&fs->ir.grad, hipStreamNonBlocking));
    for (hipEvent_t* e : {&fs->ir.fork, &fs->ir.fork_lead, &fs->ir.join_polar, &fs->ir.join_polar_small,
                          &fs->ir.join_big, &fs->ir.cache_fork, &fs->ir.cache_join, &fs->ir.join_grad})
        if (st == LCBG_OK) st = map_hip(hipEventCreateWithFlags(e, hipEventDisableTiming));
    return st;
}

// Rebuilds the per-curve caches on `stream` if the curves changed since.
// With `beside` (finite-source runs) the branch part is forked onto the large
// pass's stream, and *joined is set: the caller makes `stream` wait for
// ir.cache_join before anything reads the branches.
lcbg_status fs_refresh(lcbg_batch* b, hipStream_t stream, bool beside = false, bool* joined = nullptr)
{
    if (joined) *joined = false;
    if (!b->caches_stale) return LCBG_OK;
    FsState* fs = b->fs;
    lcbg_status st = launch_curve_prep(b->d_curves, b->n_curves, fs->cache, fs->tab, stream);
    if (st == LCBG_OK && beside) {
        st = map_hip(hipEventRecord(fs->ir.cache_fork, stream));
        if (st == LCBG_OK) st = map_hip(hipStreamWaitEvent(fs->ir.big, fs->ir.cache_fork, 0));
        if (st == LCBG_OK) st = launch_branch_cache(b->d_curves, b->n_curves, fs->cache, fs->ir.big);
        if (st == LCBG_OK) st = map_hip(hipEventRecord(fs->ir.cache_join, fs->ir.big));
        if (st == LCBG_OK && joined) *joined = true;
    } else if (st == LCBG_OK) {
        st = launch_branch_cache(b->d_curves, b->n_curves, fs->cache, stream);
    }
    if (st == LCBG_OK) b->caches_stale = false;
    return st;
}

// Outputs of a gradient run (any pointer may be null).
struct GradArgs {
    double* mag_aa;
    double* dmag;
    const double* flux;
    const double* sigma;
    double fs, fb;
    double* chi2;
    double* grad;
};

// One finite-source run enqueued on `stream`.  capture: the run is being
// recorded into a HIP graph, so no host waits, no timing events, the caustic
// caches are always rebuilt, and the workspace hand-over events are left to
// the graph launch.
lcbg_status fs_run(lcbg_batch* b, const int64_t* offsets, const double* times, int64_t n_epochs,
                   double* mag, int32_t* method, hipStream_t stream, float* stage_ms,
                   const GradArgs* ga = nullptr, bool capture = false)
{
    // the gradient differentiates the fixed-lattice quadrature only
    if (ga && b->options.adaptive_source_bins != 0) return LCBG_UNSUPPORTED;
    lcbg_status st = fs_prepare(b);
    if (st != LCBG_OK) return st;
    FsState* fs = b->fs;
    const void* epoch_mem = fs->epoch_mem;
    const float* sched = fs->sched;
    st = ensure_epochs(fs, n_epochs > 0 ? n_epochs : 1);
    if (st != LCBG_OK) return st;
    FsEpochs ep = epoch_view(fs, n_epochs, offsets, times, mag, method);
    ep.cost = b->d_cost;
    ep.dint = dint_view(fs);
    ep.hex_list = hex_view(fs);
    ep.near_curve = near_curve_view(fs);
    ep.near_chunks = near_chunks_view(fs);
    ep.err = err_view(fs);
    ep.level = level_view(fs);
    ep.src_y = b->src_y;
    // the pseudo-curve batch of the dynamic path runs different epochs under
    // the same count, so it keeps no keys
    if (!capture && !b->dynamic_images) {
        st = ensure_sched(fs, n_epochs, stream, false);
        if (st != LCBG_OK) return st;
    }
    ep.sched = (!b->dynamic_images && fs->sched_n == n_epochs) ? fs->sched : nullptr;
    // a graph recorded over workspace that has just been reallocated is void
    if (b->graph && !capture && (fs->epoch_mem != epoch_mem || fs->sched != sched)) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    fs->last_n = n_epochs;
    fs->last_mag = mag;
    fs->last_method = method;
    // timing events of a profiled run, released on every return path
    struct RunEvents {
        hipEvent_t ev[5] = {};
        hipEvent_t kev[2 * kIrLaunches] = {};
        IrStreams* ir = nullptr;
        ~RunEvents()
        {
            if (!ir) return;
            ir->kev = nullptr;
            for (auto& e : ev) (void)hipEventDestroy(e);
            for (auto& e : kev) (void)hipEventDestroy(e);
        }
    } events;
    hipEvent_t* ev = events.ev;
    hipEvent_t* kev = events.kev;
    if (stage_ms) {
        for (auto& e : events.ev) (void)hipEventCreate(&e);
        for (auto& e : events.kev) (void)hipEventCreate(&e);
        events.ir = &fs->ir;
        fs->ir.kev = kev;
        (void)hipEventRecord(ev[0], stream);
    }
    bool cache_joined = false;
    // curves of one epoch each (the pseudo-curves of orbital motion) build
    // their branches after stage 1, for the curves it queued only; such
    // partial caches are never reused
    const bool lazy_branches = b->one_epoch_curves && b->caches_stale;
    if (lazy_branches)
        st = launch_curve_prep(b->d_curves, b->n_curves, fs->cache, fs->tab, stream);
    else
        st = fs_refresh(b, stream, true, &cache_joined);
    // every return path leaves `stream` ordered after the branch caches
    struct CacheJoin {
        hipStream_t stream;
        hipEvent_t join;
        bool armed;
        ~CacheJoin()
        {
            if (armed) (void)hipStreamWaitEvent(stream, join, 0);
        }
    } cache_join{stream, fs->ir.cache_join, cache_joined};
    if (st != LCBG_OK) return st;
    if (stage_ms) (void)hipEventRecord(ev[1], stream);
    ScratchPool& pool = g_pool[b->device % kMaxDevices];
    std::lock_guard<std::mutex> lock(pool.mu);
    st = ensure_scratch(pool);
    if (st == LCBG_OK && ga) st = ensure_grad_pool(pool);
    if (st == LCBG_OK && b->options.adaptive_source_bins != 0) {
        int max_bins = b->options.max_source_bins > b->options.source_bins ? b->options.max_source_bins
                                                                             : b->options.source_bins;
        for (const CurveRec& r : b->h_curves)
            if (r.source_bins > max_bins) max_bins = r.source_bins;
        st = ensure_adapt_pool(pool, b->options.max_source_bins, max_bins);
    }
    if (st != LCBG_OK) return st;
    // LCBG_STAGE_SYNC=1: wait after every stage and report it on stderr
    // (a diagnostic in the spirit of the reference's LCBININT_AREA_DIAGNOSTICS)
    const bool stage_sync = !capture && std::getenv("LCBG_STAGE_SYNC") != nullptr;
    auto stage_done = [&](const char* name) {
        if (!stage_sync || st != LCBG_OK) return;
        st = map_hip(hipStreamSynchronize(stream));
        unsigned h[8] = {0};
        (void)hipMemcpy(h, ep.counters, sizeof(h), hipMemcpyDeviceToHost);
        std::fprintf(stderr, "[lcbg] %s done (status %d, queued %u / %u, flags %u)\n", name,
                     static_cast<int>(st), h[0], h[2], h[4]);
        std::fflush(stderr);
    };
    if (!capture) st = map_hip(hipStreamWaitEvent(stream, pool.last_use, 0));
    if (st == LCBG_OK) st = launch_reset_counters(ep.counters, kCounterWords, stream);
    if (st == LCBG_OK)
        st = launch_fs_stage1(b->d_curves, b->n_curves, fs->cache, fs->opt, ep, stream, b->one_epoch_curves);
    if (st == LCBG_OK && lazy_branches)
        st = launch_branch_cache(b->d_curves, b->n_curves, fs->cache, stream, true);
    stage_done("stage 1 (point source, bounding box)");
    if (stage_ms && st == LCBG_OK) (void)hipEventRecord(ev[2], stream);
    // stage 2 reads the branch caches built beside stage 1
    if (st == LCBG_OK && cache_join.armed) {
        cache_join.armed = false;
        st = map_hip(hipStreamWaitEvent(stream, fs->ir.cache_join, 0));
    }
    if (st == LCBG_OK) st = launch_fs_stage2(b->d_curves, pool.cus, fs->cache, fs->opt, ep, fs->stage2_waves, stream);
    stage_done("stage 2 (caustic distance, hexadecapole)");
    if (stage_ms && st == LCBG_OK) (void)hipEventRecord(ev[3], stream);
    // the batch's schedule: fewer resident waves than slots, the deferral
    // budget and the large-slot threshold
    IrScratch scr = pool.scr;
    const lcbg_schedule& sc = b->schedule;
    auto waves = [&](int per_cu, int alloc) {
        const int n = per_cu > 0 ? per_cu * pool.cus : alloc;
        return n < alloc ? n : alloc;
    };
    scr.seed.n_slots = waves(sc.seed_waves_per_cu, scr.seed.n_slots);
    scr.cart.n_slots = waves(sc.cart_waves_per_cu, scr.cart.n_slots);
    scr.polar.n_slots = waves(sc.polar_waves_per_cu, scr.polar.n_slots);
    fs->opt.heavy_budget = sc.heavy_budget > 0 ? sc.heavy_budget : kHeavyBudget;
    fs->opt.lead_items = sc.lead_items > 0 ? sc.lead_items : kLeadItems;
    fs->opt.step_budget = b->work_budget > 0 ? b->work_budget : kStepBudget;
    if (st == LCBG_OK)
        st = launch_fs_ir(b->d_curves, fs->cache, fs->tab, fs->opt, ep, lists_view(fs), scr, stream, fs->ir);
    stage_done("stage 3 (inverse ray)");
    if (ga && st == LCBG_OK) {
        // the cell enumeration of the gradient quadrature needs only the seeds:
        // it runs on its own stream from the end of the seed search on (the
        // ir.fork event), beside the area passes, and the rest joins after both
        GradWork work = pool.grad;
        grad_item_view(fs, work);
        st = map_hip(hipStreamWaitEvent(fs->ir.grad, fs->ir.fork, 0));
        if (st == LCBG_OK)
            st = launch_grad_enumerate(b->d_curves, fs->opt, ep, lists_view(fs), pool.grad_slots, pool.grad_large,
                                       work, fs->ir.grad);
        if (st == LCBG_OK) st = map_hip(hipEventRecord(fs->ir.join_grad, fs->ir.grad));
        if (st == LCBG_OK) st = map_hip(hipStreamWaitEvent(stream, fs->ir.join_grad, 0));
        if (st == LCBG_OK)
            st = launch_grad_rest(b->d_curves, b->n_curves, fs->opt, ep, lists_view(fs), work, pool.cus, ga->mag_aa,
                                  ga->dmag, ga->flux, ga->sigma, ga->fs, ga->fb, ga->chi2, ga->grad, stream);
    }
    if (st == LCBG_OK) st = launch_first_nonfinite(mag, n_epochs, ep.counters, stream);
    if (st == LCBG_OK && !capture) st = map_hip(hipEventRecord(pool.last_use, stream));
    if (stage_ms && st == LCBG_OK) {
        (void)hipEventRecord(ev[4], stream);
        (void)hipEventSynchronize(ev[4]);
        for (int k = 0; k < 4; ++k) (void)hipEventElapsedTime(&stage_ms[k], ev[k], ev[k + 1]);
        for (int k = 0; k < kIrLaunches; ++k)
            (void)hipEventElapsedTime(&fs->ir_kernel_ms[k], kev[2 * k], kev[2 * k + 1]);
    }
    return st;
}

lcbg_status fs_error_flags(lcbg_batch* b, unsigned* flags, unsigned* counts, int64_t* first_failed = nullptr)
{
    if (!b->fs || !b->fs->epoch_mem) return LCBG_INVALID_ARGUMENT;
    const FsEpochs ep = epoch_view(b->fs, 0, nullptr, nullptr, nullptr, nullptr);
    unsigned h[kCounterWords];
    lcbg_status st = map_hip(hipMemcpy(h, ep.counters, sizeof(h), hipMemcpyDeviceToHost));
    if (st != LCBG_OK) return st;
    if (flags) *flags = h[kQFlags];
    if (counts) {
        counts[0] = h[kQNearCount];
        counts[1] = h[kQIrCount];
    }
    if (first_failed) {
        const unsigned w = h[kFirstFailWord];
        *first_failed = w == 0 ? -1 : b->fs->last_n - static_cast<int64_t>(w);
    }
    return LCBG_OK;
}

}  // namespace lcbg

using namespace lcbg;

extern "C" lcbg_status lcbg_batch_finite_source_device(lcbg_batch* batch, const int64_t* offsets,
                                                       const double* times, int64_t n_epochs,
                                                       double* mag, int32_t* method, void* stream)
{
    if (!batch || n_epochs < 0) return LCBG_INVALID_ARGUMENT;
    if (batch->n_curves == 0 || n_epochs == 0) return LCBG_OK;
    if (!offsets || !times || !mag || !method) return LCBG_INVALID_ARGUMENT;
    return fs_run(batch, offsets, times, n_epochs, mag, method, static_cast<hipStream_t>(stream),
                  nullptr);
}

extern "C" lcbg_status lcbg_batch_set_schedule(lcbg_batch* batch, const lcbg_schedule* schedule)
{
    if (!batch || !schedule) return LCBG_INVALID_ARGUMENT;
    if (schedule->seed_waves_per_cu < 0 || schedule->cart_waves_per_cu < 0 ||
        schedule->polar_waves_per_cu < 0 || schedule->heavy_budget < 0 || schedule->lead_items < 0)
        return LCBG_INVALID_ARGUMENT;
    batch->schedule = *schedule;
    return LCBG_OK;
}

extern "C" lcbg_status lcbg_batch_set_work_budget(lcbg_batch* batch, int64_t steps)
{
    if (!batch || steps < 0 || steps > kStepBudget) return LCBG_INVALID_ARGUMENT;
    batch->work_budget = steps;
    return LCBG_OK;
}

extern "C" lcbg_status lcbg_batch_get_schedule(const lcbg_batch* batch, lcbg_schedule* schedule)
{
    if (!batch || !schedule) return LCBG_INVALID_ARGUMENT;
    *schedule = batch->schedule;
    return LCBG_OK;
}

extern "C" lcbg_status lcbg_batch_autotune(lcbg_batch* batch, const int64_t* offsets, const double* times,
                                           int64_t n_epochs, double* mag, int32_t* method, void* stream,
                                           int32_t max_runs, lcbg_schedule* best, float* best_ms)
{
    if (!batch || !best || !best_ms || max_runs < 1) return LCBG_INVALID_ARGUMENT;
    const hipStream_t stream_ = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    lcbg_status st = map_hip(hipEventCreate(&e0));
    if (st == LCBG_OK) st = map_hip(hipEventCreate(&e1));
    int runs = 0;
    // one run of the workload under the batch's current schedule
    auto timed = [&](float* ms) -> lcbg_status {
        lcbg_status r = map_hip(hipEventRecord(e0, stream_));
        if (r == LCBG_OK) r = lcbg_batch_finite_source_device(batch, offsets, times, n_epochs, mag, method, stream);
        if (r == LCBG_OK) r = map_hip(hipEventRecord(e1, stream_));
        if (r == LCBG_OK) r = map_hip(hipEventSynchronize(e1));
        if (r == LCBG_OK) r = map_hip(hipEventElapsedTime(ms, e0, e1));
        ++runs;
        return r;
    };
    if (st == LCBG_OK) st = timed(best_ms);  // warm-up: caches, workspace
    lcbg_schedule cur = batch->schedule;
    if (st == LCBG_OK && runs < max_runs) st = timed(best_ms);
    // candidates per field, tried one field at a time with the others fixed
    const int32_t waves[] = {1, 2, 3, 4};
    const int32_t budgets[] = {4, 8, 16, 32, 64};
    const int32_t leads[] = {8, 16, 32, 48, 64};
    for (int field = 0; field < 5 && st == LCBG_OK; ++field) {
        const int n_cand = field == 3 ? 5 : (field == 4 ? 5 : 4);
        for (int k = 0; k < n_cand && st == LCBG_OK && runs < max_runs; ++k) {
            lcbg_schedule trial = cur;
            switch (field) {
            case 0: trial.seed_waves_per_cu = waves[k]; break;
            case 1: trial.cart_waves_per_cu = waves[k]; break;
            case 2: trial.polar_waves_per_cu = waves[k]; break;
            case 3: trial.heavy_budget = budgets[k]; break;
            default: trial.lead_items = leads[k]; break;
            }
            batch->schedule = trial;
            float ms = 0.0f;
            st = timed(&ms);
            if (st == LCBG_OK && ms < *best_ms) {
                *best_ms = ms;
                cur = trial;
            }
        }
    }
    batch->schedule = cur;
    *best = cur;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return st;
}

extern "C" lcbg_status lcbg_batch_capture_finite_source(lcbg_batch* batch, const int64_t* offsets,
                                                         const double* times, int64_t n_epochs,
                                                         double* mag, int32_t* method, void* stream)
{
    (void)stream;
    if (!batch || n_epochs <= 0 || batch->n_curves == 0) return LCBG_INVALID_ARGUMENT;
    if (!offsets || !times || !mag || !method) return LCBG_INVALID_ARGUMENT;
    // everything the run allocates, allocated before the capture
    lcbg_status st = fs_prepare(batch);
    if (st == LCBG_OK) st = ensure_epochs(batch->fs, n_epochs);
    if (st == LCBG_OK) st = ensure_sched(batch->fs, n_epochs, nullptr, true);
    if (st == LCBG_OK) {
        ScratchPool& pool = g_pool[batch->device % kMaxDevices];
        std::lock_guard<std::mutex> lock(pool.mu);
        st = ensure_scratch(pool);
    }
    if (st != LCBG_OK) return st;
    if (batch->graph) {
        (void)hipGraphExecDestroy(batch->graph);
        batch->graph = nullptr;
    }
    if (!batch->capture_stream) {
        st = map_hip(hipStreamCreateWithFlags(&batch->capture_stream, hipStreamNonBlocking));
        if (st != LCBG_OK) return st;
    }
    hipStream_t cs = batch->capture_stream;
    st = map_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    if (st != LCBG_OK) return st;
    batch->caches_stale = true;
    const lcbg_status run = fs_run(batch, offsets, times, n_epochs, mag, method, cs, nullptr, nullptr, true);
    hipGraph_t g = nullptr;
    st = map_hip(hipStreamEndCapture(cs, &g));  // ends the capture on every path
    if (st == LCBG_OK) st = run;
    if (st == LCBG_OK) st = map_hip(hipGraphInstantiate(&batch->graph, g, nullptr, nullptr, 0));
    if (g) (void)hipGraphDestroy(g);
    batch->caches_stale = true;  // nothing has run yet
    return st;
}

extern "C" lcbg_status lcbg_batch_launch_graph(lcbg_batch* batch, void* stream)
{
    if (!batch || !batch->graph) return LCBG_INVALID_ARGUMENT;
    const hipStream_t st_ = static_cast<hipStream_t>(stream);
    ScratchPool& pool = g_pool[batch->device % kMaxDevices];
    std::lock_guard<std::mutex> lock(pool.mu);
    lcbg_status st = map_hip(hipStreamWaitEvent(st_, pool.last_use, 0));
    if (st == LCBG_OK) st = map_hip(hipGraphLaunch(batch->graph, st_));
    if (st == LCBG_OK) st = map_hip(hipEventRecord(pool.last_use, st_));
    if (st == LCBG_OK) batch->caches_stale = false;  // the graph rebuilt them
    return st;
}

namespace lcbg {
void trajectory_constants(const lcbg_curve& cv, const lcbg_dynamics& d, int orbit_mode, TrajConst<double>* val_out,
                          TrajConst<TDual<kTrajParams>>* dual_out);
std::vector<EarthRowD> earth_rows_device(int* generation);

// Device buffers of the dynamic path, kept across calls and grown on demand
// (a call allocates only when it holds more epochs or curves than any before).
struct DynState {
    void* mem = nullptr;
    int64_t cap_epochs = 0;
    int32_t cap_curves = 0;
    void* consts = nullptr;      // [curve] TrajConst<double> or TrajConst<TDual<9>>
    int64_t* off = nullptr;      // [curve + 1] offsets (host calls)
    double* t = nullptr;         // [epoch] times (host calls)
    double *x = nullptr, *y = nullptr, *s = nullptr;  // [epoch] trajectory
    double* jac = nullptr;       // [epoch][3][9]
    double* mag = nullptr;       // [epoch] (host calls)
    int32_t* meth = nullptr;     // [epoch] (host calls)
    double* aa = nullptr;        // [epoch] (host calls)
    double* gp = nullptr;        // [epoch][8] pseudo-curve derivatives
    double* dmag = nullptr;      // [epoch][12] (host calls)
    lcbg_result* res = nullptr;  // [epoch]
    unsigned* bad = nullptr;     // non-finite orbit flag
    int64_t* child_off = nullptr;  // [chunk + 1]
    double* child_t = nullptr;     // [chunk]
    EarthRowD* earth = nullptr;
    int n_earth = 0;
    int earth_gen = -1;
    lcbg_batch* static_child = nullptr;  // one pseudo-curve per curve (no orbital motion)
};

void dyn_destroy(DynState* d)
{
    if (!d) return;
    if (d->static_child) lcbg_batch_destroy(d->static_child);
    if (d->mem) (void)hipFree(d->mem);
    if (d->earth) (void)hipFree(d->earth);
    delete d;
}
}  // namespace lcbg

namespace {
// Epochs per device pass of the dynamic path with orbital motion: each epoch
// is a curve of its own (its own separation, its own caustic cache of about
// 340 KB), so 524288 of them hold about 178 GB; fewer, larger passes share
// the inverse-ray stage's tail (its longest epoch) among more epochs.
constexpr int32_t kDynChunk = 524288;
constexpr int kDynGradParams = 12;

lcbg_status dyn_reserve(lcbg_batch* b, int64_t n_epochs)
{
    if (!b->dyn) b->dyn = new (std::nothrow) DynState();
    if (!b->dyn) return LCBG_DEVICE_ERROR;
    DynState& d = *b->dyn;
    const int32_t nc = b->n_curves;
    if (d.mem && d.cap_epochs >= n_epochs && d.cap_curves >= nc) return LCBG_OK;
    int64_t cap = 1024;
    while (cap < n_epochs) cap *= 2;
    const std::size_t ce = static_cast<std::size_t>(cap), cc = static_cast<std::size_t>(nc);
    const std::size_t ck = static_cast<std::size_t>(kDynChunk);
    const std::size_t bytes = carve_size(cc * sizeof(TrajConst<TDual<kTrajParams>>)) +
                              carve_size((cc + 1) * sizeof(int64_t)) + carve_size(ce * sizeof(double)) * 7 +
                              carve_size(ce * sizeof(int32_t)) +
                              carve_size(ce * 3 * kTrajParams * sizeof(double)) +
                              carve_size(ce * kGradParams * sizeof(double)) +
                              carve_size(ce * kDynGradParams * sizeof(double)) +
                              carve_size(ce * sizeof(lcbg_result)) + carve_size(64) +
                              carve_size((ck + 1) * sizeof(int64_t)) + carve_size(ck * sizeof(double));
    if (d.mem) {
        (void)hipFree(d.mem);
        d.mem = nullptr;
        d.cap_epochs = 0;
        d.cap_curves = 0;
    }
    lcbg_status st = map_hip(hipMalloc(&d.mem, bytes));
    if (st != LCBG_OK) return st;
    char* p = static_cast<char*>(d.mem);
    d.consts = carve<unsigned char>(p, cc * sizeof(TrajConst<TDual<kTrajParams>>));
    d.off = carve<int64_t>(p, cc + 1);
    d.t = carve<double>(p, ce);
    d.x = carve<double>(p, ce);
    d.y = carve<double>(p, ce);
    d.s = carve<double>(p, ce);
    d.mag = carve<double>(p, ce);
    d.aa = carve<double>(p, ce);
    (void)carve<double>(p, ce);  // spare
    d.meth = carve<int32_t>(p, ce);
    d.jac = carve<double>(p, ce * 3 * kTrajParams);
    d.gp = carve<double>(p, ce * kGradParams);
    d.dmag = carve<double>(p, ce * kDynGradParams);
    d.res = carve<lcbg_result>(p, ce);
    d.bad = carve<unsigned>(p, 16);
    d.child_off = carve<int64_t>(p, ck + 1);
    d.child_t = carve<double>(p, ck);
    d.cap_epochs = cap;
    d.cap_curves = nc;
    return LCBG_OK;
}

// The Earth table on the device, uploaded again after a new table was read.
lcbg_status dyn_earth(DynState& d)
{
    int gen = 0;
    std::vector<EarthRowD> rows = earth_rows_device(&gen);
    if (gen == d.earth_gen) return LCBG_OK;
    if (d.earth) (void)hipFree(d.earth);
    d.earth = nullptr;
    d.n_earth = 0;
    if (!rows.empty()) {
        lcbg_status st = map_hip(hipMalloc(reinterpret_cast<void**>(&d.earth), rows.size() * sizeof(EarthRowD)));
        if (st == LCBG_OK)
            st = map_hip(hipMemcpy(d.earth, rows.data(), rows.size() * sizeof(EarthRowD), hipMemcpyHostToDevice));
        if (st != LCBG_OK) return st;
        d.n_earth = static_cast<int>(rows.size());
    }
    d.earth_gen = gen;
    return LCBG_OK;
}

// The trajectories of every epoch of the batch's curves on the device: per
// epoch x, y, s (and with jac the [3][9] Jacobian), from per-curve constants
// computed on the host (its libm, as the host trajectory).  Ends with the
// stream synchronised: LCBG_NUMERICAL_ERROR if an orbit is not finite
// (lens_model.cpp:117-120).
lcbg_status dyn_trajectory(lcbg_batch* batch, const lcbg_dynamics* dyn, int32_t orbit_mode, const int64_t* off,
                           const double* times, bool with_jac, double* x, double* y, double* s, double* jac,
                           hipStream_t stream, bool sky_rotation = false)
{
    DynState& d = *batch->dyn;
    const int32_t nc = batch->n_curves;
    // xallarap amplitudes of the form the options do not name are
    // unsupported (has_unsupported_dynamic_effects, lens_model.cpp:15-26)
    for (int32_t c = 0; c < nc; ++c) {
        const lcbg_dynamics& q = dyn[c];
        if ((q.xallarap_type != 1 && (q.xi_1 != 0.0 || q.xi_2 != 0.0)) ||
            (q.xallarap_type != 2 && (q.piEN_xa != 0.0 || q.piEE_xa != 0.0)) || q.xallarap_type < 0 ||
            q.xallarap_type > 2)
            return LCBG_UNSUPPORTED;
    }
    lcbg_status st = LCBG_OK;
    bool any_parallax = false;
    for (int32_t c = 0; c < nc; ++c) any_parallax |= dyn[c].piEN != 0.0 || dyn[c].piEE != 0.0;
    if (any_parallax) st = dyn_earth(d);
    if (st != LCBG_OK) return st;
    std::vector<TrajConst<double>> kv(with_jac ? 0 : static_cast<std::size_t>(nc));
    std::vector<TrajConst<TDual<kTrajParams>>> kd(with_jac ? static_cast<std::size_t>(nc) : 0);
    for (int32_t c = 0; c < nc; ++c)
        trajectory_constants(batch->h_params[static_cast<std::size_t>(c)], dyn[c], orbit_mode,
                             with_jac ? nullptr : &kv[static_cast<std::size_t>(c)],
                             with_jac ? &kd[static_cast<std::size_t>(c)] : nullptr);
    for (auto& k : kv) k.sky_rotation = sky_rotation ? 1 : 0;
    for (auto& k : kd) k.sky_rotation = sky_rotation ? 1 : 0;
    const void* kh = with_jac ? static_cast<const void*>(kd.data()) : static_cast<const void*>(kv.data());
    const std::size_t kbytes = with_jac ? kd.size() * sizeof(kd[0]) : kv.size() * sizeof(kv[0]);
    st = map_hip(hipMemcpyAsync(d.consts, kh, kbytes, hipMemcpyHostToDevice, stream));
    if (st == LCBG_OK) st = map_hip(hipMemsetAsync(d.bad, 0, sizeof(unsigned), stream));
    if (st == LCBG_OK)
        st = launch_traj(d.consts, with_jac, nc, off, times, d.earth, d.n_earth, x, y, s, jac, d.bad, stream);
    unsigned h_bad = 0;
    if (st == LCBG_OK) st = map_hip(hipMemcpyAsync(&h_bad, d.bad, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    // the host copies of the constants and of h_bad live until here
    if (st == LCBG_OK) st = map_hip(hipStreamSynchronize(stream));
    if (st != LCBG_OK) return st;
    return h_bad ? LCBG_NUMERICAL_ERROR : LCBG_OK;
}

// Inputs and outputs of one dynamic run: host arrays (device == false, the
// run ends synchronised) or device arrays enqueued on `stream`.
struct DynIO {
    const int64_t* offsets;
    const double* times;
    int64_t n_epochs;
    double* mag;
    int32_t* method;
    double* mag_aa;         // may be null
    double* dmag;           // [epoch][12], null: values only
    lcbg_result* results;   // host calls only; may be null
    bool device;
    hipStream_t stream;
};

// The dynamic path (LensModel with parallax and orbital motion,
// src/lcbinint/model/lens_model.cpp:106-231): the trajectory of every epoch
// on the device (k_traj), then the batched finite-source run on pseudo-curves
// that place each epoch's source and separation; optionally the derivatives of
// each epoch's magnification with respect to the 12 parameters (t0, tE, u0,
// alpha, s, q, rho, u_LD, piEN, piEE, g1, g2).  Without orbital motion the
// separation is constant: one pseudo-curve per curve, with the impact
// parameter per epoch, keeps one caustic cache per curve, reused across calls
// while s and q stay; with it, every epoch is a curve of its own, in chunks.
lcbg_status dynamic_run(lcbg_batch* batch, const lcbg_dynamics* dyn, int32_t orbital_motion_mode, const DynIO& io)
{
    if (!batch || io.n_epochs < 0) return LCBG_INVALID_ARGUMENT;
    if (orbital_motion_mode < LCBG_ORBIT_STATIC || orbital_motion_mode > LCBG_ORBIT_KEPLER)
        return LCBG_UNSUPPORTED;
    const int32_t nc = batch->n_curves;
    const int64_t n = io.n_epochs;
    if (nc == 0 || n == 0) return LCBG_OK;
    if (!dyn || !io.offsets || !io.times || !io.mag || !io.method) return LCBG_INVALID_ARGUMENT;
    if (n > INT32_MAX) return LCBG_INVALID_ARGUMENT;
    if (!io.device) {
        if (io.offsets[0] != 0 || io.offsets[nc] != n) return LCBG_INVALID_ARGUMENT;
        for (int32_t c = 0; c < nc; ++c)
            if (io.offsets[c + 1] < io.offsets[c]) return LCBG_INVALID_ARGUMENT;
    }
    const bool grad = io.dmag != nullptr;
    const hipStream_t stream = io.stream;
    lcbg_status st = dyn_reserve(batch, n);
    if (st != LCBG_OK) return st;
    DynState& d = *batch->dyn;
    const int64_t* off = io.offsets;
    const double* times = io.times;
    if (!io.device) {
        st = map_hip(hipMemcpyAsync(d.off, io.offsets, (nc + 1) * sizeof(int64_t), hipMemcpyHostToDevice, stream));
        if (st == LCBG_OK)
            st = map_hip(hipMemcpyAsync(d.t, io.times, n * sizeof(double), hipMemcpyHostToDevice, stream));
        off = d.off;
        times = d.t;
    }
    if (st == LCBG_OK)
        st = dyn_trajectory(batch, dyn, orbital_motion_mode, off, times, grad, d.x, d.y, d.s, grad ? d.jac : nullptr,
                            stream);
    if (st != LCBG_OK) return st;
    double* mag = io.device ? io.mag : d.mag;
    int32_t* meth = io.device ? io.method : d.meth;
    double* aa = io.device ? io.mag_aa : d.aa;
    double* dm = io.device ? io.dmag : d.dmag;
    if (orbital_motion_mode == LCBG_ORBIT_STATIC) {
        // one pseudo-curve per curve: t0 = 0, tE = 1, alpha = 0, the curve's
        // s, q, rho, LD and lattice; per epoch "time" x and impact parameter y
        if (d.static_child && d.static_child->n_curves != nc) {
            lcbg_batch_destroy(d.static_child);
            d.static_child = nullptr;
        }
        if (!d.static_child) {
            st = lcbg_batch_create(batch->h_params.data(), nc, &batch->options, &d.static_child);
            if (st != LCBG_OK) return st;
            d.static_child->dynamic_images = true;
            d.static_child->h_curves.assign(static_cast<std::size_t>(nc), CurveRec{});
        }
        lcbg_batch* ch = d.static_child;
        std::vector<CurveRec> recs(batch->h_curves);
        for (CurveRec& r : recs) {
            r.t0 = 0.0;
            r.inverse_tE = 1.0;
            r.u0 = 0.0;
            r.cos_alpha = 1.0;
            r.sin_alpha = 0.0;
        }
        if (std::memcmp(recs.data(), ch->h_curves.data(), recs.size() * sizeof(CurveRec)) != 0) {
            ch->h_curves = recs;
            st = map_hip(hipMemcpy(ch->d_curves, recs.data(), recs.size() * sizeof(CurveRec), hipMemcpyHostToDevice));
            ch->caches_stale = true;
        }
        ch->src_y = d.y;
        if (st == LCBG_OK) {
            if (grad) {
                const GradArgs ga{aa, d.gp, nullptr, nullptr, 0.0, 0.0, nullptr, nullptr};
                st = fs_run(ch, off, d.x, n, mag, meth, stream, nullptr, &ga);
            } else {
                st = fs_run(ch, off, d.x, n, mag, meth, stream, nullptr);
            }
        }
        ch->src_y = nullptr;
        if (st == LCBG_OK && grad) st = launch_dyn_chain(d.gp, d.jac, 0, static_cast<int>(n), dm, stream);
        if (st == LCBG_OK && io.results) st = lcbg_batch_epoch_results_device(ch, n, d.res, stream);
    } else {
        // pseudo-curve capacity: the run's epoch count rounded up to a power of
        // two, at least 64 and at most kDynChunk; a larger later run replaces it
        int32_t want = 64;
        while (want < kDynChunk && want < n) want *= 2;
        if (batch->dyn_child && batch->dyn_child->n_curves < want) {
            lcbg_batch_destroy(batch->dyn_child);
            batch->dyn_child = nullptr;
        }
        if (!batch->dyn_child) {
            std::vector<lcbg_curve> idle(static_cast<std::size_t>(want),
                                         lcbg_curve{0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0});
            st = lcbg_batch_create(idle.data(), want, &batch->options, &batch->dyn_child);
            if (st != LCBG_OK) return st;
            batch->dyn_child->dynamic_images = true;
            batch->dyn_child->one_epoch_curves = true;
        }
        lcbg_batch* ch = batch->dyn_child;
        const int32_t chunk = ch->n_curves;
        for (int64_t base = 0; st == LCBG_OK && base < n; base += chunk) {
            const int32_t m = static_cast<int32_t>(std::min<int64_t>(chunk, n - base));
            st = launch_pseudo_curves(batch->d_curves, nc, off, d.x, d.y, d.s, base, m, chunk, ch->d_curves,
                                      d.child_off, d.child_t, stream);
            ch->caches_stale = true;
            if (st == LCBG_OK) {
                if (grad) {
                    const GradArgs ga{aa ? aa + base : nullptr, d.gp + base * kGradParams, nullptr, nullptr, 0.0,
                                      0.0, nullptr, nullptr};
                    st = fs_run(ch, d.child_off, d.child_t, m, mag + base, meth + base, stream, nullptr, &ga);
                } else {
                    st = fs_run(ch, d.child_off, d.child_t, m, mag + base, meth + base, stream, nullptr);
                }
            }
            // the pseudo-curve's source is the magnifier's source position, its
            // separation the epoch's s(t): the records read as LensModel fills them
            if (st == LCBG_OK && io.results) st = lcbg_batch_epoch_results_device(ch, m, d.res + base, stream);
            if (st == LCBG_OK && grad)
                st = launch_dyn_chain(d.gp + base * kGradParams, d.jac, base, m, dm, stream);
        }
    }
    if (st != LCBG_OK || io.device) return st;
    st = map_hip(hipStreamSynchronize(stream));
    if (st == LCBG_OK) st = map_hip(hipMemcpy(io.mag, d.mag, n * sizeof(double), hipMemcpyDeviceToHost));
    if (st == LCBG_OK) st = map_hip(hipMemcpy(io.method, d.meth, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (st == LCBG_OK && grad && io.mag_aa)
        st = map_hip(hipMemcpy(io.mag_aa, d.aa, n * sizeof(double), hipMemcpyDeviceToHost));
    if (st == LCBG_OK && grad)
        st = map_hip(hipMemcpy(io.dmag, d.dmag, n * kDynGradParams * sizeof(double), hipMemcpyDeviceToHost));
    if (st == LCBG_OK && io.results)
        st = map_hip(hipMemcpy(io.results, d.res, n * sizeof(lcbg_result), hipMemcpyDeviceToHost));
    // a non-finite magnification is a numerical error (lens_model.cpp:228-231)
    for (int64_t i = 0; st == LCBG_OK && i < n; ++i)
        if (!std::isfinite(io.mag[i])) st = LCBG_NUMERICAL_ERROR;
    return st;
}

DynIO host_io(const int64_t* offsets, const double* times, int64_t n, double* mag, int32_t* method,
              double* mag_aa = nullptr, double* dmag = nullptr, lcbg_result* results = nullptr)
{
    return DynIO{offsets, times, n, mag, method, mag_aa, dmag, results, false, nullptr};
}

}  // namespace

extern "C" lcbg_status lcbg_batch_light_curves_dynamic(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                                       int32_t orbital_motion_mode,
                                                       const int64_t* offsets, const double* times,
                                                       int64_t n_epochs, double* mag, int32_t* method)
{
    return dynamic_run(batch, dyn, orbital_motion_mode, host_io(offsets, times, n_epochs, mag, method));
}

extern "C" lcbg_status lcbg_batch_light_curves_dynamic_results(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                                               int32_t orbital_motion_mode,
                                                               const int64_t* offsets, const double* times,
                                                               int64_t n_epochs, lcbg_result* results)
{
    if (n_epochs > 0 && !results) return LCBG_INVALID_ARGUMENT;
    std::vector<double> mag(static_cast<std::size_t>(n_epochs > 0 ? n_epochs : 0));
    std::vector<int32_t> meth(mag.size());
    return dynamic_run(batch, dyn, orbital_motion_mode,
                       host_io(offsets, times, n_epochs, mag.data(), meth.data(), nullptr, nullptr, results));
}

extern "C" lcbg_status lcbg_batch_light_curves_dynamic_grad(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                                            int32_t orbital_motion_mode,
                                                            const int64_t* offsets, const double* times,
                                                            int64_t n_epochs, double* mag, int32_t* method,
                                                            double* mag_aa, double* dmag)
{
    if (n_epochs > 0 && !dmag) return LCBG_INVALID_ARGUMENT;
    return dynamic_run(batch, dyn, orbital_motion_mode, host_io(offsets, times, n_epochs, mag, method, mag_aa, dmag));
}

extern "C" lcbg_status lcbg_batch_light_curves_dynamic_device(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                                              int32_t orbital_motion_mode,
                                                              const int64_t* offsets, const double* times,
                                                              int64_t n_epochs, double* mag, int32_t* method,
                                                              double* mag_aa, double* dmag, void* stream)
{
    return dynamic_run(batch, dyn, orbital_motion_mode,
                       DynIO{offsets, times, n_epochs, mag, method, mag_aa, dmag, nullptr, true,
                             static_cast<hipStream_t>(stream)});
}

extern "C" lcbg_status lcbg_batch_dynamic_sources_device(lcbg_batch* batch, const lcbg_dynamics* dyn,
                                                         int32_t orbital_motion_mode, const int64_t* offsets,
                                                         const double* times, int64_t n_epochs, double* x,
                                                         double* y, double* s, double* jac, void* stream)
{
    if (!batch || n_epochs < 0) return LCBG_INVALID_ARGUMENT;
    if (orbital_motion_mode < LCBG_ORBIT_STATIC || orbital_motion_mode > LCBG_ORBIT_KEPLER)
        return LCBG_UNSUPPORTED;
    if (batch->n_curves == 0 || n_epochs == 0) return LCBG_OK;
    if (!dyn || !offsets || !times || !x || !y || !s) return LCBG_INVALID_ARGUMENT;
    lcbg_status st = dyn_reserve(batch, 0);
    if (st != LCBG_OK) return st;
    return dyn_trajectory(batch, dyn, orbital_motion_mode, offsets, times, jac != nullptr, x, y, s, jac,
                          static_cast<hipStream_t>(stream));
}

extern "C" lcbg_status lcbg_batch_finite_source_grad_device(lcbg_batch* batch, const int64_t* offsets,
                                                            const double* times, int64_t n_epochs,
                                                            double* mag, int32_t* method,
                                                            double* mag_aa, double* dmag,
                                                            void* stream)
{
    if (!batch || n_epochs < 0) return LCBG_INVALID_ARGUMENT;
    if (batch->n_curves == 0 || n_epochs == 0) return LCBG_OK;
    if (!offsets || !times || !mag || !method || !dmag) return LCBG_INVALID_ARGUMENT;
    const GradArgs ga{mag_aa, dmag, nullptr, nullptr, 0.0, 0.0, nullptr, nullptr};
    return fs_run(batch, offsets, times, n_epochs, mag, method, static_cast<hipStream_t>(stream),
                  nullptr, &ga);
}

extern "C" lcbg_status lcbg_batch_chi2_grad_device(lcbg_batch* batch, const int64_t* offsets,
                                                   const double* times, int64_t n_epochs,
                                                   const double* flux, const double* sigma,
                                                   double fs, double fb, double* mag,
                                                   int32_t* method, double* mag_aa, double* dmag,
                                                   double* chi2, double* grad, void* stream)
{
    if (!batch || n_epochs < 0) return LCBG_INVALID_ARGUMENT;
    if (batch->n_curves == 0 || n_epochs == 0) return LCBG_OK;
    if (!offsets || !times || !mag || !method || !flux || !sigma || !chi2 || !grad)
        return LCBG_INVALID_ARGUMENT;
    const GradArgs ga{mag_aa, dmag, flux, sigma, fs, fb, chi2, grad};
    return fs_run(batch, offsets, times, n_epochs, mag, method, static_cast<hipStream_t>(stream),
                  nullptr, &ga);
}

extern "C" lcbg_status lcbg_batch_finite_source_profile(lcbg_batch* batch, const int64_t* offsets,
                                                        const double* times, int64_t n_epochs,
                                                        double* mag, int32_t* method, void* stream,
                                                        float* stage_ms, uint32_t* queue_counts,
                                                        uint32_t* error_flags)
{
    if (!batch || n_epochs < 0 || !stage_ms) return LCBG_INVALID_ARGUMENT;
    if (batch->n_curves == 0 || n_epochs == 0) return LCBG_OK;
    lcbg_status st = fs_run(batch, offsets, times, n_epochs, mag, method,
                            static_cast<hipStream_t>(stream), stage_ms);
    if (st != LCBG_OK) return st;
    return fs_error_flags(batch, error_flags, queue_counts);
}

extern "C" lcbg_status lcbg_batch_debug_epoch(lcbg_batch* batch, int64_t epoch, int32_t method,
                                              int32_t last_phase, double* seeds, int32_t capacity,
                                              int32_t n_in, int32_t* n_out, double* mag)
{
    if (!batch || !batch->fs || !batch->fs->epoch_mem || !seeds || !n_out || !mag)
        return LCBG_INVALID_ARGUMENT;
    FsState* fs = batch->fs;
    if (epoch < 0 || epoch >= fs->last_n || capacity < 0 || n_in > capacity || n_in > kSeedCap)
        return LCBG_INVALID_ARGUMENT;
    if (method != LCBG_METHOD_INVERSE_RAY_CARTESIAN && method != LCBG_METHOD_INVERSE_RAY_POLAR)
        return LCBG_INVALID_ARGUMENT;
    ScratchPool& pool = g_pool[batch->device % kMaxDevices];
    std::lock_guard<std::mutex> lock(pool.mu);
    lcbg_status st = ensure_scratch(pool);
    if (st != LCBG_OK) return st;
    st = map_hip(hipDeviceSynchronize());
    if (st != LCBG_OK) return st;
    FsEpochs ep = epoch_view(fs, fs->last_n, nullptr, nullptr, nullptr, nullptr);
    char* mem = nullptr;
    const std::size_t seed_bytes = static_cast<std::size_t>(kSeedCap) * sizeof(Pt2);
    st = map_hip(hipMalloc(reinterpret_cast<void**>(&mem), seed_bytes + 256));
    if (st != LCBG_OK) return st;
    Pt2* d_seeds = reinterpret_cast<Pt2*>(mem);
    int* d_n = reinterpret_cast<int*>(mem + seed_bytes);
    double* d_mag = reinterpret_cast<double*>(mem + seed_bytes + 64);
    if (n_in > 0) st = map_hip(hipMemcpy(d_seeds, seeds, static_cast<std::size_t>(n_in) * sizeof(Pt2),
                                         hipMemcpyHostToDevice));
    if (st == LCBG_OK)
        st = launch_ir_epoch_debug(batch->d_curves, fs->cache, fs->tab, fs->opt, ep, pool.scr, epoch,
                                   method, last_phase, d_seeds, n_in, d_n, d_mag, nullptr);
    if (st == LCBG_OK) st = map_hip(hipDeviceSynchronize());
    int n = 0;
    if (st == LCBG_OK) st = map_hip(hipMemcpy(&n, d_n, sizeof(int), hipMemcpyDeviceToHost));
    if (st == LCBG_OK) st = map_hip(hipMemcpy(mag, d_mag, sizeof(double), hipMemcpyDeviceToHost));
    if (st == LCBG_OK && n_in < 0 && n > 0) {
        const int k = n < capacity ? n : capacity;
        st = map_hip(hipMemcpy(seeds, d_seeds, static_cast<std::size_t>(k) * sizeof(Pt2),
                               hipMemcpyDeviceToHost));
    }
    *n_out = n;
    (void)hipFree(mem);
    return st;
}

extern "C" lcbg_status lcbg_batch_debug_critical_sources(lcbg_batch* batch, int32_t curve,
                                                         double* out)
{
    if (!batch || !out || curve < 0 || curve >= batch->n_curves) return LCBG_INVALID_ARGUMENT;
    lcbg_status st = fs_prepare(batch);
    if (st == LCBG_OK) st = fs_refresh(batch, nullptr);
    if (st == LCBG_OK) st = map_hip(hipDeviceSynchronize());
    if (st != LCBG_OK) return st;
    const std::size_t n = static_cast<std::size_t>(kSeedPhases) * 4;
    return map_hip(hipMemcpy(out, batch->fs->cache.crit + static_cast<std::size_t>(curve) * n,
                             n * sizeof(Pt2), hipMemcpyDeviceToHost));
}

extern "C" lcbg_status lcbg_batch_work_counts(lcbg_batch* batch, uint64_t* counts)
{
    if (!batch || !counts || !batch->fs || !batch->fs->epoch_mem) return LCBG_INVALID_ARGUMENT;
    const FsEpochs ep = epoch_view(batch->fs, 0, nullptr, nullptr, nullptr, nullptr);
    lcbg_status st = map_hip(hipDeviceSynchronize());
    if (st != LCBG_OK) return st;
    return map_hip(hipMemcpy(counts, ep.counters + kWorkOffset, 8 * sizeof(uint64_t),
                             hipMemcpyDeviceToHost));
}

extern "C" lcbg_status lcbg_batch_ir_kernel_ms(lcbg_batch* batch, float* ms)
{
    if (!batch || !ms || !batch->fs) return LCBG_INVALID_ARGUMENT;
    for (int k = 0; k < kIrLaunches; ++k) ms[k] = batch->fs->ir_kernel_ms[k];
    return LCBG_OK;
}

extern "C" lcbg_status lcbg_batch_set_epoch_cost(lcbg_batch* batch, int64_t* cost)
{
    if (!batch) return LCBG_INVALID_ARGUMENT;
    batch->d_cost = cost;
    return LCBG_OK;
}

extern "C" int64_t lcbg_wall_clock_rate_khz(void)
{
    int dev = 0;
    int khz = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return -1;
    return khz;
}

namespace {

// The single-curve host path keeps one batch and its device buffers per host
// thread and device, so that a loop over curves (a fit calling
// LightCurve(...)(times, params) again and again) reuses the workspace: a
// call with the options of the previous call only uploads the curve and the
// times.  A call with other options, or more epochs than any before, replaces
// them.  The objects live until the process ends.
struct SingleCurve {
    lcbg_batch* batch = nullptr;
    lcbg_options options{};
    int device = -1;
    int64_t cap = 0;
    char* mem = nullptr;  // times, mag, method, offsets, results
};

thread_local SingleCurve* g_single = nullptr;

bool same_options(const lcbg_options& a, const lcbg_options& b)
{
    return std::memcmp(&a, &b, sizeof(lcbg_options)) == 0;
}

lcbg_status single_curve_run(const lcbg_curve* curve, const lcbg_options* options, const double* times,
                             int64_t n, double* mag, int32_t* method, lcbg_result* results,
                             int64_t* first_failed)
{
    if (first_failed) *first_failed = -1;
    if (!curve || n < 0 || (n > 0 && !times) || (n > 0 && !mag && !results)) return LCBG_INVALID_ARGUMENT;
    if (n == 0) return LCBG_OK;
    lcbg_options opt = options ? *options : lcbg_default_options();
    int dev = 0;
    lcbg_status st = map_hip(hipGetDevice(&dev));
    if (st != LCBG_OK) return st;
    if (!g_single) g_single = new (std::nothrow) SingleCurve();
    if (!g_single) return LCBG_DEVICE_ERROR;
    SingleCurve& sc = *g_single;
    if (sc.batch && (sc.device != dev || !same_options(sc.options, opt))) {
        lcbg_batch_destroy(sc.batch);
        sc.batch = nullptr;
        if (sc.device != dev && sc.mem) {
            int cur = dev;
            (void)hipSetDevice(sc.device);
            (void)hipFree(sc.mem);
            (void)hipSetDevice(cur);
            sc.mem = nullptr;
            sc.cap = 0;
        }
    }
    if (!sc.batch) {
        st = lcbg_batch_create(curve, 1, &opt, &sc.batch);
        if (st != LCBG_OK) {
            sc.batch = nullptr;
            return st;
        }
        sc.options = opt;
        sc.device = dev;
    } else {
        st = lcbg_batch_update_curves(sc.batch, curve, nullptr);
        if (st != LCBG_OK) return st;
    }
    const std::size_t per = sizeof(double) * 2 + sizeof(int32_t) + sizeof(lcbg_result) + 3 * 256;
    if (n > sc.cap) {
        if (sc.mem) (void)hipFree(sc.mem);
        sc.mem = nullptr;
        sc.cap = 0;
        int64_t cap = 256;
        while (cap < n) cap *= 2;
        st = map_hip(hipMalloc(reinterpret_cast<void**>(&sc.mem), static_cast<std::size_t>(cap) * per + 1024));
        if (st != LCBG_OK) {
            sc.mem = nullptr;
            return st;
        }
        sc.cap = cap;
    }
    char* p = sc.mem;
    double* d_times = carve<double>(p, sc.cap);
    double* d_mag = carve<double>(p, sc.cap);
    int32_t* d_method = carve<int32_t>(p, sc.cap);
    lcbg_result* d_res = carve<lcbg_result>(p, sc.cap);
    int64_t* d_off = carve<int64_t>(p, 2);
    const int64_t off[2] = {0, n};
    const std::size_t bytes = static_cast<std::size_t>(n) * sizeof(double);
    st = map_hip(hipMemcpy(d_times, times, bytes, hipMemcpyHostToDevice));
    if (st == LCBG_OK) st = map_hip(hipMemcpy(d_off, off, sizeof(off), hipMemcpyHostToDevice));
    if (st == LCBG_OK)
        st = lcbg_batch_finite_source_device(sc.batch, d_off, d_times, n, d_mag, d_method, nullptr);
    if (st == LCBG_OK && results) st = lcbg_batch_epoch_results_device(sc.batch, n, d_res, nullptr);
    if (st == LCBG_OK && mag) st = map_hip(hipMemcpy(mag, d_mag, bytes, hipMemcpyDeviceToHost));
    if (st == LCBG_OK && method)
        st = map_hip(hipMemcpy(method, d_method, static_cast<std::size_t>(n) * sizeof(int32_t),
                               hipMemcpyDeviceToHost));
    if (st == LCBG_OK && results)
        st = map_hip(hipMemcpy(results, d_res, static_cast<std::size_t>(n) * sizeof(lcbg_result),
                               hipMemcpyDeviceToHost));
    unsigned flags = 0;
    int64_t first = -1;
    if (st == LCBG_OK) st = fs_error_flags(sc.batch, &flags, nullptr, &first);
    if (st == LCBG_OK) {
        if (first_failed) *first_failed = first;
        st = status_of_run(flags, first);
    }
    return st;
}

}  // namespace

extern "C" lcbg_status lcbg_light_curve(const lcbg_curve* curve, const lcbg_options* options,
                                        const double* times, int64_t n, double* mag,
                                        int32_t* method)
{
    if (n > 0 && !mag) return LCBG_INVALID_ARGUMENT;
    return single_curve_run(curve, options, times, n, mag, method, nullptr, nullptr);
}

extern "C" lcbg_status lcbg_light_curve_results(const lcbg_curve* curve, const lcbg_options* options,
                                                const double* times, int64_t n, lcbg_result* results,
                                                int64_t* first_failed_epoch)
{
    if (n > 0 && !results) return LCBG_INVALID_ARGUMENT;
    return single_curve_run(curve, options, times, n, nullptr, nullptr, results, first_failed_epoch);
}

extern "C" lcbg_status lcbg_batch_last_run_status(lcbg_batch* batch, uint32_t* error_flags,
                                                  int64_t* first_failed_epoch)
{
    if (!batch || !batch->fs || !batch->fs->epoch_mem) return LCBG_INVALID_ARGUMENT;
    {
        ScratchPool& pool = g_pool[batch->device % kMaxDevices];
        std::lock_guard<std::mutex> lock(pool.mu);
        if (pool.last_use) {
            const lcbg_status st = map_hip(hipEventSynchronize(pool.last_use));
            if (st != LCBG_OK) return st;
        }
    }
    unsigned flags = 0;
    int64_t first = -1;
    const lcbg_status st = fs_error_flags(batch, &flags, nullptr, &first);
    if (st != LCBG_OK) return st;
    if (error_flags) *error_flags = flags;
    if (first_failed_epoch) *first_failed_epoch = first;
    return status_of_run(flags, first);
}

extern "C" lcbg_status lcbg_batch_epoch_results_device(lcbg_batch* batch, int64_t n_epochs, lcbg_result* results,
                                                       void* stream)
{
    if (!batch || !results || n_epochs < 0) return LCBG_INVALID_ARGUMENT;
    if (n_epochs == 0) return LCBG_OK;
    FsState* fs = batch->fs;
    if (!fs || !fs->epoch_mem || fs->last_n != n_epochs) return LCBG_INVALID_ARGUMENT;
    FsEpochs ep = epoch_view(fs, n_epochs, nullptr, nullptr, fs->last_mag, fs->last_method);
    ep.err = err_view(fs);
    ep.level = level_view(fs);
    return launch_epoch_results(batch->d_curves, fs->opt, ep, batch->options.finite_source_tol,
                                batch->options.finite_source_reltol, results, static_cast<hipStream_t>(stream));
}

namespace {
// One-curve batch of the source-trajectory helper, kept per host thread.
thread_local lcbg_batch* g_geometry_batch = nullptr;
thread_local int g_geometry_device = -1;
}  // namespace

extern "C" lcbg_status lcbg_source_trajectory(const lcbg_curve* curve, const lcbg_dynamics* dyn,
                                              int32_t orbital_motion_mode, const double* times, int64_t n,
                                              double* x, double* y)
{
    if (!curve || !dyn || n < 0 || (n > 0 && (!times || !x || !y))) return LCBG_INVALID_ARGUMENT;
    if (orbital_motion_mode < LCBG_ORBIT_STATIC || orbital_motion_mode > LCBG_ORBIT_KEPLER)
        return LCBG_UNSUPPORTED;
    if (n == 0) return LCBG_OK;
    int dev = 0;
    lcbg_status st = map_hip(hipGetDevice(&dev));
    if (st != LCBG_OK) return st;
    if (g_geometry_batch && g_geometry_device != dev) {
        // a handle of another device is left to the process end
        g_geometry_batch = nullptr;
    }
    if (!g_geometry_batch) {
        const lcbg_options opt = lcbg_default_options();
        st = lcbg_batch_create(curve, 1, &opt, &g_geometry_batch);
        if (st != LCBG_OK) return st;
        g_geometry_device = dev;
    } else {
        st = lcbg_batch_update_curves(g_geometry_batch, curve, nullptr);
        if (st != LCBG_OK) return st;
    }
    lcbg_batch* b = g_geometry_batch;
    st = dyn_reserve(b, n);
    if (st != LCBG_OK) return st;
    DynState& d = *b->dyn;
    const int64_t off[2] = {0, n};
    st = map_hip(hipMemcpy(d.off, off, sizeof(off), hipMemcpyHostToDevice));
    if (st == LCBG_OK) st = map_hip(hipMemcpy(d.t, times, n * sizeof(double), hipMemcpyHostToDevice));
    if (st == LCBG_OK)
        st = dyn_trajectory(b, dyn, orbital_motion_mode, d.off, d.t, false, d.x, d.y, d.s, nullptr, nullptr, true);
    if (st == LCBG_OK) st = map_hip(hipMemcpy(x, d.x, n * sizeof(double), hipMemcpyDeviceToHost));
    if (st == LCBG_OK) st = map_hip(hipMemcpy(y, d.y, n * sizeof(double), hipMemcpyDeviceToHost));
    return st;
}

extern "C" lcbg_status lcbg_geometry_branches(double s, double q, int32_t n_points, int32_t map_to_source_plane,
                                              int32_t vbm, double* x, double* y)
{
    if (n_points < 8 || !x || !y) return LCBG_INVALID_ARGUMENT;
    const double q_eff = (vbm != 0 && q != 0.0) ? 1.0 / q : q;
    const std::size_t n = static_cast<std::size_t>(n_points);
    std::vector<Pt2> host(n);
    for (std::size_t i = 0; i < n; ++i) {
        // as geometry_branches writes the phase (python/lcbinint_pybind.cpp:1917-1918)
        const double phase = 2.0 * kPiHost * static_cast<double>(i) / static_cast<double>(n_points);
        host[i] = Pt2{std::cos(phase), std::sin(phase)};
    }
    Pt2* mem = nullptr;  // phases, points [n][4], branches [4][n]
    lcbg_status st = map_hip(hipMalloc(reinterpret_cast<void**>(&mem), 9 * n * sizeof(Pt2)));
    if (st != LCBG_OK) return st;
    st = map_hip(hipMemcpy(mem, host.data(), n * sizeof(Pt2), hipMemcpyHostToDevice));
    if (st == LCBG_OK)
        st = launch_geometry_branches(s, q_eff, mem, n_points, map_to_source_plane != 0, mem + n, mem + 5 * n,
                                      nullptr);
    std::vector<Pt2> br(4 * n);
    if (st == LCBG_OK) st = map_hip(hipMemcpy(br.data(), mem + 5 * n, 4 * n * sizeof(Pt2), hipMemcpyDeviceToHost));
    (void)hipFree(mem);
    if (st != LCBG_OK) return st;
    for (std::size_t k = 0; k < 4 * n; ++k) {
        x[k] = br[k].x;
        y[k] = br[k].y;
    }
    return LCBG_OK;
}
