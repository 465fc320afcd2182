// Device trajectories of the LensModel path (annual parallax, circular and
// Keplerian orbital motion) and the per-epoch pseudo-curves the dynamic light
// curves run on:
//
//   k_traj          one workgroup per curve, one thread per epoch: source
//                   position in the instantaneous binary frame and separation
//                   (traj_epoch), optionally with their derivatives; the Earth
//                   table is staged in LDS;
//   k_pseudo_curves one thread per epoch of a chunk: the record of a curve whose
//                   static source at "time" x, "u0" y sits where the epoch's
//                   source does, with the epoch's separation (t0 = 0, tE = 1,
//                   alpha = 0), so that the batched finite-source pipeline runs
//                   every epoch on the caustic of its own separation;
//   k_dyn_chain     the chain rule from the pseudo-curve derivatives and the
//                   trajectory Jacobian to the 12 curve parameters.
#include "fs_launch.hpp"
#include "traj.hpp"

namespace lcbg {

namespace {

constexpr int kTrajBlock = 256;
constexpr int kEarthLdsRows = 1024;  // rows staged in LDS (larger tables are read from HBM)

template <class T>
__global__ void __launch_bounds__(kTrajBlock) k_traj(const TrajConst<T>* __restrict__ consts,
                                                      const int64_t* __restrict__ offsets,
                                                      const double* __restrict__ times,
                                                      const EarthRowD* __restrict__ earth, int n_rows,
                                                      double* xs, double* ys, double* ss, double* jac,
                                                      unsigned* bad)
{
    __shared__ EarthRowD lds_rows[kEarthLdsRows];
    const int c = blockIdx.x;
    const TrajConst<T>& K = consts[c];
    const EarthRowD* rows = earth;
    if (K.parallax && n_rows <= kEarthLdsRows) {
        for (int i = threadIdx.x; i < n_rows; i += kTrajBlock) lds_rows[i] = earth[i];
        __syncthreads();
        rows = lds_rows;
    }
    const int64_t lo = offsets[c], hi = offsets[c + 1];
    bool ok_all = true;
    for (int64_t e = lo + threadIdx.x; e < hi; e += kTrajBlock) {
        T x, y, s;
        const bool ok = traj_epoch(K, rows, n_rows, times[e], &x, &y, &s);
        ok_all = ok_all && ok;
        xs[e] = tval(x);
        ys[e] = tval(y);
        ss[e] = tval(s);
        if constexpr (!std::is_same<T, double>::value) {
            const T out[3] = {x, y, s};
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < kTrajParams; ++k) jac[27 * e + 9 * j + k] = out[j].d[k];
        }
    }
    if (!ok_all) atomicOr(bad, 1u);
}

__global__ void __launch_bounds__(kTrajBlock) k_pseudo_curves(const CurveRec* __restrict__ parent, int n_parent,
                                                               const int64_t* __restrict__ offsets,
                                                               const double* __restrict__ xs,
                                                               const double* __restrict__ ys,
                                                               const double* __restrict__ ss, int64_t base, int m,
                                                               int chunk, CurveRec* child, int64_t* child_off,
                                                               double* child_t)
{
    const int k = static_cast<int>(blockIdx.x) * kTrajBlock + static_cast<int>(threadIdx.x);
    if (k > chunk) return;
    if (k == chunk) {
        child_off[chunk] = m;
        return;
    }
    child_off[k] = k < m ? k : m;
    CurveRec r;
    r.t0 = 0.0;
    r.inverse_tE = 1.0;
    r.cos_alpha = 1.0;
    r.sin_alpha = 0.0;
    r.pad = 0;
    if (k >= m) {
        // an idle curve with no epochs
        r.u0 = 0.0;
        r.s = 0.0;
        r.q_eff = 1.0;
        r.rho = 0.0;
        r.ld_c = 0.0;
        r.ld_d = 0.0;
        r.source_bins = 0;
        child[k] = r;
        return;
    }
    const int64_t e = base + k;
    // the parent curve: the last c with offsets[c] <= e
    int lo = 0, hi = n_parent - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    const CurveRec pc = parent[lo];
    r.u0 = ys[e];
    r.s = ss[e];
    r.q_eff = pc.q_eff;
    r.rho = pc.rho;
    r.ld_c = pc.ld_c;
    r.ld_d = pc.ld_d;
    r.source_bins = pc.source_bins;
    child[k] = r;
    child_t[k] = xs[e];
}

// dmag[12 e + k] for the 12 parameters (t0, tE, u0, alpha, s, q, rho, u_LD,
// piEN, piEE, g1, g2) from the pseudo-curve's derivatives gp (t0, tE, u0,
// alpha, s, q, rho, u_LD of a curve with alpha = 0, tE = 1: dA/dx = -d/dt0,
// dA/dy = d/du0, dA/ds) and the trajectory Jacobian J (x, y, s by t0, tE, u0,
// alpha, s, piEN, piEE, g1, g2).
__global__ void __launch_bounds__(kTrajBlock) k_dyn_chain(const double* __restrict__ gp_all,
                                                           const double* __restrict__ jac, int64_t base, int m,
                                                           double* dmag)
{
    const int k = static_cast<int>(blockIdx.x) * kTrajBlock + static_cast<int>(threadIdx.x);
    if (k >= m) return;
    const int64_t e = base + k;
    const double* gp = gp_all + static_cast<size_t>(k) * kGradParams;
    const double da[3] = {-gp[0], gp[2], gp[4]};
    const double* J = jac + 27 * e;
    double tr[kTrajParams];
    for (int j = 0; j < kTrajParams; ++j) tr[j] = da[0] * J[j] + da[1] * J[kTrajParams + j] + da[2] * J[2 * kTrajParams + j];
    double* out = dmag + 12 * e;
    out[0] = tr[0];
    out[1] = tr[1];
    out[2] = tr[2];
    out[3] = tr[3];
    out[4] = tr[4];
    out[5] = gp[5];
    out[6] = gp[6];
    out[7] = gp[7];
    out[8] = tr[5];
    out[9] = tr[6];
    out[10] = tr[7];
    out[11] = tr[8];
}

}  // namespace

lcbg_status launch_traj(const void* consts, bool with_jac, int n_curves, const int64_t* offsets,
                        const double* times, const EarthRowD* earth, int n_rows, double* xs, double* ys,
                        double* ss, double* jac, unsigned* bad, hipStream_t st)
{
    if (n_curves <= 0) return LCBG_OK;
    if (with_jac) {
        hipLaunchKernelGGL(k_traj<TDual<kTrajParams>>, dim3(static_cast<unsigned>(n_curves)), dim3(kTrajBlock), 0, st,
                           static_cast<const TrajConst<TDual<kTrajParams>>*>(consts), offsets, times, earth, n_rows,
                           xs, ys, ss, jac, bad);
    } else {
        hipLaunchKernelGGL(k_traj<double>, dim3(static_cast<unsigned>(n_curves)), dim3(kTrajBlock), 0, st,
                           static_cast<const TrajConst<double>*>(consts), offsets, times, earth, n_rows, xs, ys,
                           ss, nullptr, bad);
    }
    return check_launch();
}

lcbg_status launch_pseudo_curves(const CurveRec* parent, int n_parent, const int64_t* offsets, const double* xs,
                                 const double* ys, const double* ss, int64_t base, int m, int chunk, CurveRec* child,
                                 int64_t* child_off, double* child_t, hipStream_t st)
{
    const unsigned blocks = static_cast<unsigned>((chunk + 1 + kTrajBlock - 1) / kTrajBlock);
    hipLaunchKernelGGL(k_pseudo_curves, dim3(blocks), dim3(kTrajBlock), 0, st, parent, n_parent, offsets, xs, ys, ss,
                       base, m, chunk, child, child_off, child_t);
    return check_launch();
}

lcbg_status launch_dyn_chain(const double* gp, const double* jac, int64_t base, int m, double* dmag, hipStream_t st)
{
    if (m <= 0) return LCBG_OK;
    const unsigned blocks = static_cast<unsigned>((m + kTrajBlock - 1) / kTrajBlock);
    hipLaunchKernelGGL(k_dyn_chain, dim3(blocks), dim3(kTrajBlock), 0, st, gp, jac, base, m, dmag);
    return check_launch();
}

}  // namespace lcbg
