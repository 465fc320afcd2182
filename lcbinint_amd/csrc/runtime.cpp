// Host-side runtime: error mapping, kernel timing, batch handles.
#include "runtime.hpp"

#include <cmath>
#include <new>
#include <vector>

namespace lcbg {

lcbg_status map_hip(hipError_t e)
{
    if (e == hipSuccess) return LCBG_OK;
    if (e == hipErrorInvalidValue) return LCBG_INVALID_ARGUMENT;
    return LCBG_DEVICE_ERROR;
}

lcbg_status check_launch() { return map_hip(hipGetLastError()); }

CurveRec make_curve_rec(const lcbg_curve& cv)
{
    CurveRec r;
    r.t0 = cv.t0;
    r.inverse_tE = 1.0 / cv.tE;
    r.u0 = cv.u0;
    r.cos_alpha = std::cos(cv.alpha);
    r.sin_alpha = std::sin(cv.alpha);
    r.s = cv.s;
    r.q_eff = cv.q != 0.0 ? 1.0 / cv.q : cv.q;
    r.rho = std::abs(cv.rho);
    r.ld_c = cv.ld_c;
    r.ld_d = cv.ld_d;
    r.source_bins = 0;
    r.pad = 0;
    return r;
}

namespace {

struct ThreadTiming {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
    bool pending = false;
};

thread_local ThreadTiming g_timing;

}  // namespace

KernelTimer::KernelTimer(hipStream_t stream) : stream_(stream)
{
    if (!g_timing.start) {
        (void)hipEventCreate(&g_timing.start);
        (void)hipEventCreate(&g_timing.stop);
    }
    (void)hipEventRecord(g_timing.start, stream_);
}

void KernelTimer::stop()
{
    (void)hipEventRecord(g_timing.stop, stream_);
    g_timing.pending = true;
}

}  // namespace lcbg

using namespace lcbg;

extern "C" double lcbg_last_kernel_ms(void)
{
    if (!g_timing.pending) return -1.0;
    if (hipEventSynchronize(g_timing.stop) != hipSuccess) return -1.0;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, g_timing.start, g_timing.stop) != hipSuccess) return -1.0;
    return static_cast<double>(ms);
}

extern "C" const char* lcbg_status_string(lcbg_status status)
{
    switch (status) {
    case LCBG_OK: return "ok";
    case LCBG_INVALID_ARGUMENT: return "invalid argument";
    case LCBG_NUMERICAL_ERROR: return "numerical error";
    case LCBG_UNSUPPORTED: return "unsupported";
    case LCBG_DEVICE_ERROR: return "device error";
    }
    return "unknown status";
}

extern "C" lcbg_options lcbg_default_options(void)
{
    lcbg_options o;
    o.source_bins = 50;
    o.caustic_bins = 1400;
    o.mode = 4;
    o.polar_source_bins = 0;
    o.grid_ratio = 4.0;
    o.polar_grid_ratio = 0.0;
    o.point_source_threshold = 20.0;
    o.hexadecapole_threshold = 3.0;
    o.adaptive_hex_threshold = 0.001;
    o.finite_source_tol = 0.0;
    o.finite_source_reltol = 0.0;
    o.adaptive_source_bins = 0;
    o.max_source_bins = 400;
    return o;
}

extern "C" lcbg_status lcbg_batch_create(const lcbg_curve* curves, int32_t n_curves,
                                         const lcbg_options* options, lcbg_batch** out)
{
    if (!out || n_curves < 0 || (n_curves > 0 && !curves)) return LCBG_INVALID_ARGUMENT;
    *out = nullptr;
    for (int32_t c = 0; c < n_curves; ++c) {
        if (!std::isfinite(curves[c].tE) || curves[c].tE == 0.0 || !std::isfinite(curves[c].q) ||
            !std::isfinite(curves[c].s))
            return LCBG_INVALID_ARGUMENT;  // LensParameters::is_valid (lens_parameters.cpp:7-11)
    }
    lcbg_batch* b = new (std::nothrow) lcbg_batch();
    if (!b) return LCBG_DEVICE_ERROR;
    b->n_curves = n_curves;
    b->options = options ? *options : lcbg_default_options();
    lcbg_status st = map_hip(hipGetDevice(&b->device));
    if (st == LCBG_OK && n_curves > 0) {
        std::vector<CurveRec>& recs = b->h_curves;
        recs.resize(static_cast<std::size_t>(n_curves));
        for (int32_t c = 0; c < n_curves; ++c) recs[static_cast<std::size_t>(c)] = make_curve_rec(curves[c]);
        b->h_params.assign(curves, curves + n_curves);
        st = map_hip(hipMalloc(reinterpret_cast<void**>(&b->d_curves), recs.size() * sizeof(CurveRec)));
        if (st == LCBG_OK)
            st = map_hip(hipMemcpy(b->d_curves, recs.data(), recs.size() * sizeof(CurveRec),
                                   hipMemcpyHostToDevice));
    }
    if (st != LCBG_OK) {
        lcbg_batch_destroy(b);
        return st;
    }
    *out = b;
    return LCBG_OK;
}

extern "C" lcbg_status lcbg_batch_destroy(lcbg_batch* batch)
{
    if (!batch) return LCBG_OK;
    if (batch->dyn_child) lcbg_batch_destroy(batch->dyn_child);
    if (batch->dyn) dyn_destroy(batch->dyn);
    if (batch->graph) (void)hipGraphExecDestroy(batch->graph);
    if (batch->capture_stream) (void)hipStreamDestroy(batch->capture_stream);
    if (batch->fs) fs_destroy(batch->fs);
    if (batch->d_curves) (void)hipFree(batch->d_curves);
    delete batch;
    return LCBG_OK;
}

extern "C" lcbg_status lcbg_batch_update_curves(lcbg_batch* batch, const lcbg_curve* curves,
                                                void* stream)
{
    if (!batch || (batch->n_curves > 0 && !curves)) return LCBG_INVALID_ARGUMENT;
    for (int32_t c = 0; c < batch->n_curves; ++c) {
        if (!std::isfinite(curves[c].tE) || curves[c].tE == 0.0 || !std::isfinite(curves[c].q) ||
            !std::isfinite(curves[c].s))
            return LCBG_INVALID_ARGUMENT;  // LensParameters::is_valid (lens_parameters.cpp:7-11)
    }
    for (int32_t c = 0; c < batch->n_curves; ++c) {
        CurveRec& r = batch->h_curves[static_cast<std::size_t>(c)];
        const int nb = r.source_bins;
        r = make_curve_rec(curves[c]);
        r.source_bins = nb;
    }
    batch->h_params.assign(curves, curves + batch->n_curves);
    batch->caches_stale = true;
    if (batch->n_curves == 0) return LCBG_OK;
    // pageable source: the copy is staged before the call returns, and the
    // stream orders it after the batch's earlier work on that stream
    return map_hip(hipMemcpyAsync(batch->d_curves, batch->h_curves.data(),
                                  batch->h_curves.size() * sizeof(CurveRec), hipMemcpyHostToDevice,
                                  static_cast<hipStream_t>(stream)));
}

extern "C" lcbg_status lcbg_batch_set_source_bins(lcbg_batch* batch, const int32_t* source_bins)
{
    if (!batch || (batch->n_curves > 0 && !source_bins)) return LCBG_INVALID_ARGUMENT;
    for (int32_t c = 0; c < batch->n_curves; ++c) {
        if (source_bins[c] < 0) return LCBG_INVALID_ARGUMENT;
    }
    for (int32_t c = 0; c < batch->n_curves; ++c)
        batch->h_curves[static_cast<std::size_t>(c)].source_bins = source_bins[c];
    if (batch->n_curves == 0) return LCBG_OK;
    return map_hip(hipMemcpy(batch->d_curves, batch->h_curves.data(),
                             batch->h_curves.size() * sizeof(CurveRec), hipMemcpyHostToDevice));
}
