// lcbg_light_curve_device_noalloc: one curve's finite-source light curve with
// every piece of device workspace in one fixed array of this file, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).  The pipeline is the one the batch
// entry points run (caustic caches, stage 1, stage 2, seeds, Cartesian and
// polar scans), on the caller's stream only, with fewer resident waves and
// smaller slot tiers than the batch path's per-device pool:
//
//   epochs per call          <= kNaMaxEpochs
//   caustic_bins             <= kNaMaxBins
//   Cartesian rows per image <= kNaRowCapLarge (an epoch past it gets a NaN
//                               magnification, as one past the work budget)
//
// The array (`scratch`) holds, in this order: the angle tables, the curve
// record and its offsets, the curve's caustic cache, the epoch workspace, the
// seed and the inverse-ray slot scratch.  Host code reaches it only through
// the address hipGetSymbolAddress returns for the current device.
#include <cmath>
#include <mutex>
#include <vector>

#include "fs_launch.hpp"
#include "fs_wave.hpp"

namespace lcbg {

namespace {

constexpr int64_t kNaMaxEpochs = 8192;
constexpr int kNaMaxBins = 2048;
constexpr int kNaSeedSlots = 64;
constexpr int kNaCartSlots = 4 * kTeamWaves;   // four teams
constexpr int kNaCartLargeSlots = kTeamWaves;  // one team
constexpr int kNaRowCapSmall = 1 << 15;
constexpr int kNaRowCapLarge = 1 << 17;
constexpr int kNaPolarSlots = 16;
constexpr int kNaPolarLargeSlots = 1;
constexpr int kNaRunCapSmall = 1 << 14;
constexpr int kNaRunCapLarge = 1 << 20;
constexpr unsigned kNaPoolCap = 8u * kNaMaxEpochs;  // seeds of all inverse-ray epochs of a call
constexpr int kNaCounterWords = 128;
constexpr int kNaTangents = 6;
constexpr int kNaLeadItems = 4;
constexpr int kNaStage2Waves = 1024;
constexpr std::size_t kScratchBytes = std::size_t{128} << 20;
static_assert(kCounterWords <= kNaCounterWords, "counter words");

}  // namespace

// The workspace of every call on a device (128 MiB; each device that runs the
// entry point holds its own copy until the process ends).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Curve record and offsets {0, n} of the call, written in-stream.
__global__ void k_na_setup(CurveRec rec, int64_t n, CurveRec* d_curve, int64_t* d_off)
{
    if (threadIdx.x == 0) {
        d_curve[0] = rec;
        d_off[0] = 0;
        d_off[1] = n;
    }
}

template <class T>
T* bump(std::uintptr_t& p, std::size_t count)
{
    p = (p + 255) & ~std::uintptr_t{255};
    T* out = reinterpret_cast<T*>(p);
    p += count * sizeof(T);
    return out;
}

// Views of the workspace at `base` (any address: with 0 the pointers are
// offsets and the return value is the size).
struct NaViews {
    Pt2* tables;
    CurveRec* curve;
    int64_t* offsets;
    FsCache cache;
    FsEpochs ep;
    IrLists lists;
    IrScratch scr;
};

std::size_t na_layout(std::uintptr_t base, NaViews& v)
{
    std::uintptr_t p = base;
    const std::size_t bins = kNaMaxBins;
    v.tables = bump<Pt2>(p, static_cast<std::size_t>(angle_table_count(kNaMaxBins)));
    v.curve = bump<CurveRec>(p, 1);
    v.offsets = bump<int64_t>(p, 2);
    FsCache& c = v.cache;
    c.bins = kNaMaxBins;
    c.pts = bump<Pt2>(p, bins * 4);
    c.br = bump<Pt2>(p, bins * 4);
    c.pt_cell = bump<int>(p, bins * 4);
    c.br_cell = bump<int>(p, bins * 4);
    c.br_order = bump<int>(p, bins * 4);
    c.br_row = bump<int>(p, kGrid + 1);
    c.crit = bump<Pt2>(p, static_cast<std::size_t>(kSeedPhases) * 4);
    c.assign = bump<uint8_t>(p, static_cast<std::size_t>(kSeedPhases) * 4);
    c.sc = bump<CausticScalars>(p, 1);
    c.near_n = bump<int>(p, 1);
    const std::size_t cap = kNaMaxEpochs;
    FsEpochs& ep = v.ep;
    ep = FsEpochs{};
    ep.curve_of = bump<int32_t>(p, cap);
    ep.src = bump<Pt2>(p, cap);
    ep.a_ps = bump<double>(p, cap);
    ep.dist = bump<double>(p, cap);
    ep.roots = bump<double>(p, cap * 10);
    ep.near_list = bump<int64_t>(p, cap);
    ep.ir_list = bump<int64_t>(p, cap);
    ep.hex_list = bump<int64_t>(p, cap);
    ep.near_curve = bump<int64_t>(p, cap);
    ep.near_chunks = bump<int64_t>(p, cap);
    ep.counters = bump<unsigned>(p, kNaCounterWords);
    ep.dint = bump<double>(p, cap * kNaTangents);
    ep.err = bump<double>(p, cap);
    IrLists& L = v.lists;
    L.ir_sorted = bump<int64_t>(p, cap);
    L.bucket_of = bump<int>(p, cap);
    L.seed_off = bump<int>(p, cap);
    L.seed_n = bump<int>(p, cap);
    L.cart_q = bump<int>(p, cap);
    L.polar_q = bump<int>(p, cap);
    L.cart_retry = bump<int>(p, cap);
    L.polar_retry = bump<int>(p, cap);
    L.cart_big = bump<int>(p, cap);
    L.polar_big = bump<int>(p, cap);
    L.grad_retry = bump<int>(p, cap);
    L.pool_cap = kNaPoolCap;
    L.seed_pool = bump<Pt2>(p, kNaPoolCap);
    IrScratch& s = v.scr;
    s = IrScratch{};
    s.seed.n_slots = kNaSeedSlots;
    s.seed.seeds = bump<Pt2>(p, static_cast<std::size_t>(kNaSeedSlots) * kSeedCap);
    // no search defers to the probe farm (heavy_budget < 0): its lists stay unused
    s.seed.heavy.item = bump<int>(p, kHeavyCap);
    s.seed.heavy.start = bump<int>(p, kHeavyCap);
    s.seed.heavy.n_list = bump<int>(p, kHeavyCap);
    s.seed.heavy.list = nullptr;
    s.seed.heavy.flags = nullptr;
    s.farm_slots = 0;
    auto cart = [&](CartScratch& cs, int slots, int row_cap) {
        const std::size_t n = static_cast<std::size_t>(slots);
        cs.n_slots = slots;
        cs.row_cap = row_cap;
        cs.areaimage = bump<double>(p, n * kSeedCap);
        cs.overlap = bump<int>(p, n * kSeedCap);
        cs.rows = bump<double>(p, n * static_cast<std::size_t>(row_cap) * 5);
        cs.row_key = bump<int>(p, n * static_cast<std::size_t>(row_ints_per_slot(row_cap)));
    };
    cart(s.cart, kNaCartSlots, kNaRowCapSmall);
    cart(s.cart_large, kNaCartLargeSlots, kNaRowCapLarge);
    auto polar = [&](PolarScratch& ps, int slots, int run_cap) {
        const std::size_t n = static_cast<std::size_t>(slots);
        ps.n_slots = slots;
        ps.run_cap = run_cap;
        ps.phi_head = bump<int>(p, n * kPhiHash);
        ps.runs = bump<int>(p, n * static_cast<std::size_t>(run_cap) * 4);
    };
    polar(s.polar, kNaPolarSlots, kNaRunCapSmall);
    polar(s.polar_large, kNaPolarLargeSlots, kNaRunCapLarge);
    return static_cast<std::size_t>(p - base);
}

struct NaDevice {
    unsigned char* ptr = nullptr;  // this device's copy of `scratch`
    int table_bins = -1;           // caustic_bins of the angle tables held
    int cus = 0;
};

constexpr int kNaDevices = 64;
NaDevice g_na[kNaDevices];
std::mutex g_na_mu;

// Once per device, with that device current: the array's address, the row and
// phi hash heads set to empty (the kernels keep them so from call to call).
lcbg_status na_setup(NaDevice& d, int dev)
{
    void* p = nullptr;
    lcbg_status st = map_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (st != LCBG_OK) return st;
    if (!p) return LCBG_DEVICE_ERROR;
    NaViews v{};
    const std::size_t need = na_layout(reinterpret_cast<std::uintptr_t>(p), v);
    if (need > kScratchBytes) return LCBG_DEVICE_ERROR;
    for (const CartScratch* cs : {&v.scr.cart, &v.scr.cart_large}) {
        const std::size_t ints = static_cast<std::size_t>(cs->n_slots) * row_ints_per_slot(cs->row_cap);
        st = map_hip(hipMemset(cs->row_key, 0xff, ints * sizeof(int)));
        if (st != LCBG_OK) return st;
    }
    for (const PolarScratch* ps : {&v.scr.polar, &v.scr.polar_large}) {
        st = map_hip(hipMemset(ps->phi_head, 0xff, static_cast<std::size_t>(ps->n_slots) * kPhiHash * sizeof(int)));
        if (st != LCBG_OK) return st;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    st = map_hip(hipDeviceSynchronize());
    if (st != LCBG_OK) return st;
    d.cus = cus;
    d.ptr = static_cast<unsigned char*>(p);
    return LCBG_OK;
}

}  // namespace

}  // namespace lcbg

using namespace lcbg;

extern "C" lcbg_status lcbg_light_curve_device_noalloc(const lcbg_curve* curve, const lcbg_options* options,
                                                       const double* times, int64_t n, double* mag,
                                                       int32_t* method, void* stream)
{
    if (!curve || n < 0 || (n > 0 && (!times || !mag || !method))) return LCBG_INVALID_ARGUMENT;
    if (n > kNaMaxEpochs) return LCBG_INVALID_ARGUMENT;
    if (!std::isfinite(curve->tE) || curve->tE == 0.0 || !std::isfinite(curve->q) || !std::isfinite(curve->s))
        return LCBG_INVALID_ARGUMENT;  // LensParameters::is_valid (lens_parameters.cpp:7-11)
    const lcbg_options o = options ? *options : lcbg_default_options();
    const int bins = o.caustic_bins > 32 ? o.caustic_bins : 32;
    if (bins > kNaMaxBins) return LCBG_INVALID_ARGUMENT;
    // the retry and the refinement need workspace this array does not hold
    if (o.adaptive_source_bins != 0) return LCBG_UNSUPPORTED;
    if (n == 0) return LCBG_OK;
    const hipStream_t st_ = static_cast<hipStream_t>(stream);
    int dev = 0;
    lcbg_status st = map_hip(hipGetDevice(&dev));
    if (st != LCBG_OK) return st;
    if (dev < 0 || dev >= kNaDevices) return LCBG_UNSUPPORTED;
    NaDevice* d = &g_na[dev];
    {
        std::lock_guard<std::mutex> lock(g_na_mu);
        if (!d->ptr) {
            st = na_setup(*d, dev);
            if (st != LCBG_OK) return st;
        }
        if (d->table_bins != bins) {
            // synchronous upload of a few tens of KB: the first call on a
            // device and a call with another caustic_bins are not for capture
            std::vector<Pt2> host;
            angle_tables(bins, host);
            NaViews v{};
            (void)na_layout(reinterpret_cast<std::uintptr_t>(d->ptr), v);
            st = map_hip(hipMemcpy(v.tables, host.data(), host.size() * sizeof(Pt2), hipMemcpyHostToDevice));
            if (st != LCBG_OK) return st;
            d->table_bins = bins;
        }
    }
    NaViews v{};
    (void)na_layout(reinterpret_cast<std::uintptr_t>(d->ptr), v);
    v.cache.bins = bins;
    const FsTables tab = table_view(v.tables, bins);
    FsOptions opt{};
    opt.source_bins = o.source_bins;
    opt.caustic_bins = o.caustic_bins;
    opt.mode = o.mode;
    opt.polar_source_bins = o.polar_source_bins;
    opt.grid_ratio = o.grid_ratio;
    opt.polar_grid_ratio = o.polar_grid_ratio;
    opt.kinji = o.point_source_threshold;
    opt.hex_threshold = o.hexadecapole_threshold;
    opt.adaptive_hex = o.adaptive_hex_threshold;
    opt.ps_images = 0;
    opt.heavy_budget = -1;  // every seed search runs to its end in its own wave
    opt.lead_items = kNaLeadItems;
    opt.step_budget = kStepBudget;
    FsEpochs ep = v.ep;
    ep.n = n;
    ep.offsets = v.offsets;
    ep.times = times;
    ep.mag = mag;
    ep.method = method;
    ep.cost = nullptr;
    ep.sched = nullptr;
    const CurveRec rec = make_curve_rec(*curve);
    IrStreams ss{};
    ss.polar = ss.polar_small = ss.big = st_;
    ss.kev = nullptr;
    hipLaunchKernelGGL(k_na_setup, dim3(1), dim3(kWave), 0, st_, rec, n, v.curve, v.offsets);
    st = check_launch();
    if (st == LCBG_OK) st = launch_reset_counters(ep.counters, kCounterWords, st_);
    if (st == LCBG_OK) st = launch_curve_prep(v.curve, 1, v.cache, tab, st_);
    if (st == LCBG_OK) st = launch_branch_cache(v.curve, 1, v.cache, st_);
    if (st == LCBG_OK) st = launch_fs_stage1(v.curve, 1, v.cache, opt, ep, st_);
    if (st == LCBG_OK) st = launch_fs_stage2(v.curve, d->cus, v.cache, opt, ep, kNaStage2Waves, st_);
    if (st == LCBG_OK) st = launch_fs_ir(v.curve, v.cache, tab, opt, ep, v.lists, v.scr, st_, ss, false);
    if (st == LCBG_OK) st = launch_first_nonfinite(mag, n, ep.counters, st_);
    return st;
}
