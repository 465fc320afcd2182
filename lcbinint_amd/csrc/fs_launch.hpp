// Host-side launchers of the finite-source kernels.
#pragma once

#include <hip/hip_runtime_api.h>

#include "fs_types.hpp"
#include "runtime.hpp"

#include <vector>

namespace lcbg {

// The angle tables of a caustic_bins setting (bins caustic phases, the 1400
// seed-scan phases, the boundary, angular and interior probe angles), rounded
// by the host libm as the reference's cos/sin calls are; their count; the
// table pointers of a device copy.
inline int angle_table_count(int bins)
{
    return bins + kSeedPhases + kBoundaryProbes + kAngularProbes + kInteriorProbes;
}
void angle_tables(int bins, std::vector<Pt2>& host);
FsTables table_view(const Pt2* device_tables, int bins);

// Caustic caches: critical-curve solves, bounding box and point cells (the
// part stage 1 reads), then the tracked branches, their cells and their order
// by cell (the part stage 2 and the seed search read).
lcbg_status launch_curve_prep(const CurveRec* d_curves, int n_curves, const FsCache& cache,
                              const FsTables& tab, hipStream_t st);
// only_near: after stage 1, for the curves with queued epochs only
// (cache.near_n[c] > 0); the others keep no branches.
lcbg_status launch_branch_cache(const CurveRec* d_curves, int n_curves, const FsCache& cache, hipStream_t st,
                                bool only_near = false);

// Stage 1 (one workgroup per curve; with one_epoch, for curves of at most one
// epoch each, one lane per curve): trajectory, point-source solve, the
// bounding-box exit; queues the remaining epochs for stage 2.
lcbg_status launch_fs_stage1(const CurveRec* d_curves, int n_curves, const FsCache& cache,
                             const FsOptions& opt, const FsEpochs& ep, hipStream_t st, bool one_epoch = false);

// Stage 2 (three launches over the queue: caustic distances, one resident
// workgroup per CU (n_cus) with a curve's caustic in LDS, or one wave per epoch
// for caches larger than LDS; the point-source derivative test, one thread per epoch; the
// hexadecapole test, one lane per sample): queues the epochs that need
// inverse-ray shooting.  n_waves: waves of the per-epoch launches.
lcbg_status launch_fs_stage2(const CurveRec* d_curves, int n_cus, const FsCache& cache,
                             const FsOptions& opt, const FsEpochs& ep, int n_waves, hipStream_t st);

// Stage 3: the inverse-ray epochs ordered by estimated cost, their image seeds
// (persistent waves), then the Cartesian row scans on `st` beside the polar
// flood fills and the large-slot pass of the largest epochs on the two side
// streams (forked and joined through the events), then a large-slot retry
// pass for the epochs that outgrew a small slot.
struct IrStreams {
    hipStream_t polar;        // polar large-slot pass of the largest polar epochs
    hipStream_t polar_small;  // polar small-slot pass
    hipStream_t big;          // Cartesian large-slot pass of the largest epochs
    hipStream_t grad;         // gradient cell enumeration, beside the area passes
    hipEvent_t fork, fork_lead, join_polar, join_polar_small, join_big;
    hipEvent_t cache_fork, cache_join;  // branch caches on `big` beside stage 1
    hipEvent_t join_grad;               // end of the gradient cell enumeration
    // per-launch timing (profiled runs only, else null): events 2k and 2k+1
    // bracket launch k on the stream it runs on, in the order of IrLaunch
    hipEvent_t* kev;
};

// The kernel launches of the inverse-ray stage, in launch order.
enum IrLaunch {
    kIrSeeds = 0,      // k_ir_seeds, main stream
    kIrCartBig = 1,    // k_ir_cart, large slots, big-epoch stream
    kIrPolarBig = 2,   // k_ir_polar, large slots, polar stream
    kIrPolarSmall = 3, // k_ir_polar, small slots, polar small-slot stream
    kIrCartSmall = 4,  // k_ir_cart, small slots, main stream
    kIrCartRetry = 5,  // k_ir_cart, large slots, main stream
    kIrPolarRetry = 6, // k_ir_polar, large slots, main stream
    kIrLaunches = 7
};

// With every stream of `ss` equal to `st` the passes run in order on `st`
// and no event is used.  farm = false leaves out the probe-farm launches
// (heavy list, farm, resumed searches): valid when opt.heavy_budget < 0, so
// that no search defers.
lcbg_status launch_fs_ir(const CurveRec* d_curves, const FsCache& cache, const FsTables& tab,
                         const FsOptions& opt, const FsEpochs& ep, const IrLists& lists,
                         const IrScratch& scr, hipStream_t st, const IrStreams& ss, bool farm = true);

// One inverse-ray epoch of the last run on slot 0 (diagnostics): seed search
// (n_in < 0) or given seeds, then the magnification with `method`.
lcbg_status launch_ir_epoch_debug(const CurveRec* d_curves, const FsCache& cache, const FsTables& tab,
                                  const FsOptions& opt, const FsEpochs& ep, const IrScratch& scr,
                                  int64_t e, int method, int last_phase, Pt2* seeds_io, int n_in,
                                  int* n_out, double* mag_out, hipStream_t st);

// Gradient kernels of a value run.  launch_grad_enumerate needs the seeds of
// the inverse-ray items (after the seed search) and runs on its own slots
// (small, then large ones for the items that outgrow them);
// launch_grad_rest, after it and after the value run, computes the tangents of
// every epoch (point source, hexadecapole, antialiased inverse ray), then
// per-epoch derivatives dmag [n][8] and, if chi2 is not null, chi^2 and its
// gradient per curve.  mag_aa, dmag, flux/sigma/chi2/grad may be null.
lcbg_status launch_grad_enumerate(const CurveRec* d_curves, const FsOptions& opt, const FsEpochs& ep,
                                  const IrLists& lists, const PolarScratch& slots,
                                  const PolarScratch& large_slots, const GradWork& work, hipStream_t st);
lcbg_status launch_grad_rest(const CurveRec* d_curves, int n_curves, const FsOptions& opt,
                             const FsEpochs& ep, const IrLists& lists, const GradWork& work, int n_cus,
                             double* mag_aa, double* dmag, const double* flux, const double* sigma,
                             double fs, double fb, double* chi2, double* grad, hipStream_t st);

// Workgroups of one inverse-ray kernel resident per CU (kind 0 seeds,
// 1 Cartesian, 2 polar), from its register and LDS use.
int ir_resident_per_cu(int kind);
int adapt_resident_per_cu(unsigned lds);

// Records the first epoch of mag[0, n) that is not finite in
// counters[kFirstFailWord] (n - e; 0: none).
lcbg_status launch_first_nonfinite(const double* mag, int64_t n, unsigned* counters, hipStream_t st);

// Per-epoch result records (lcbg_result) of the run whose arrays `ep` views.
lcbg_status launch_epoch_results(const CurveRec* d_curves, const FsOptions& opt, const FsEpochs& ep,
                                 double tol, double reltol, lcbg_result* out, hipStream_t st);

// Zeroes the queue counters in-stream with a kernel (graph-capturable).
lcbg_status launch_reset_counters(unsigned* counters, int n, hipStream_t st);

// Critical curves (to_source false) or caustics of one lens at n phases
// (trig: cos, sin of 2 pi i / n): pts [n][4] the points, br [4][n] the tracked
// branches.
lcbg_status launch_geometry_branches(double s, double q_eff, const Pt2* trig, int n, bool to_source, Pt2* pts,
                                     Pt2* br, hipStream_t st);

// Device trajectories (traj.hip).  consts: n_curves TrajConst<double>, or
// TrajConst<TDual<9>> with with_jac (then jac [epoch][3][9] is written);
// per epoch the source position xs, ys in the binary frame and the
// separation ss; *bad |= 1 if an orbit is not finite.
struct EarthRowD;
lcbg_status launch_traj(const void* consts, bool with_jac, int n_curves, const int64_t* offsets,
                        const double* times, const EarthRowD* earth, int n_rows, double* xs, double* ys,
                        double* ss, double* jac, unsigned* bad, hipStream_t st);
// Records of the pseudo-curves of epochs [base, base + m): curve k has the
// one epoch base + k at "time" xs, with u0 = ys, s = ss, t0 = 0, tE = 1,
// alpha = 0 and its parent's q, rho, LD and lattice size; curves m..chunk-1
// are idle.  child_off gets chunk + 1 offsets, child_t the m times.
lcbg_status launch_pseudo_curves(const CurveRec* parent, int n_parent, const int64_t* offsets, const double* xs,
                                 const double* ys, const double* ss, int64_t base, int m, int chunk, CurveRec* child,
                                 int64_t* child_off, double* child_t, hipStream_t st);
// dmag [epoch][12] of epochs [base, base + m) from the pseudo-curves'
// derivatives gp [k][8] and the trajectory Jacobian.
lcbg_status launch_dyn_chain(const double* gp, const double* jac, int64_t base, int m, double* dmag, hipStream_t st);

}  // namespace lcbg
