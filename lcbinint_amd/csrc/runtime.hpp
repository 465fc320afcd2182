// Host-side runtime shared by the entry points: HIP error mapping, per-launch
// device timing, and the batch handle with its device-resident records.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "lcbinint_amd.h"

namespace lcbg {

// Per-curve constants as the kernels read them.  cos/sin(alpha) are rounded by
// the host libm, as the reference's std::cos/std::sin
// (python/lcbinint_pybind.cpp:1122-1123) are, because the device libm rounds a
// small fraction of arguments differently.  q_eff is the VBM flip 1/q
// (python/lcbinint_pybind.cpp:1125).  source_bins is the curve's own inverse-ray
// lattice resolution (0: the batch option).
struct CurveRec {
    double t0, inverse_tE, u0, cos_alpha, sin_alpha, s, q_eff, rho, ld_c, ld_d;
    int source_bins;
    int pad;
};

CurveRec make_curve_rec(const lcbg_curve& cv);

lcbg_status map_hip(hipError_t e);
lcbg_status check_launch();

// Records HIP events around one kernel on its launch stream; the elapsed time
// is kept per host thread and read back by lcbg_last_kernel_ms().
class KernelTimer {
public:
    explicit KernelTimer(hipStream_t stream);
    void stop();

private:
    hipStream_t stream_;
};

struct FsState;
void fs_destroy(FsState* fs);
struct DynState;
void dyn_destroy(DynState* d);

}  // namespace lcbg

struct lcbg_batch {
    int device = 0;
    int32_t n_curves = 0;
    lcbg_options options{};
    lcbg::CurveRec* d_curves = nullptr;  // n_curves records
    std::vector<lcbg::CurveRec> h_curves; // host copy of the same records
    lcbg::FsState* fs = nullptr;         // finite-source state, built on first use
    int64_t* d_cost = nullptr;           // optional per-epoch inverse-ray cost output (device)
    bool caches_stale = true;            // per-curve caustic caches need a rebuild
    std::vector<lcbg_curve> h_params;    // the curves as passed (for host trajectories)
    bool dynamic_images = false;         // point-source sum and seeds of the LensModel path
    bool one_epoch_curves = false;       // every curve holds at most one epoch (pseudo-curves)
    lcbg_batch* dyn_child = nullptr;     // pseudo-curves of the dynamic path
    lcbg::DynState* dyn = nullptr;       // device buffers of the dynamic path
    const double* src_y = nullptr;       // per-epoch impact parameters replacing u0 (device), or null
    lcbg_schedule schedule{};            // launch schedule (0 fields: defaults)
    int64_t work_budget = 0;             // wave steps per inverse-ray epoch (0: default)
    hipGraphExec_t graph = nullptr;      // captured finite-source run
    hipStream_t capture_stream = nullptr;
};
