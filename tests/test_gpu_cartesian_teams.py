"""Cartesian teams on the longest epochs of the batched workload.

The Cartesian kernel runs four waves per workgroup; a long epoch offers its
later seeds to idle teammates, which scan those images in their own slots, and
the owner accounts every image in seed order.  These epochs (found by timing
the 2048-curve workload, seed 3) have 3-5 images that survive the overlap rule
among ~130 seeds, with 10^4-10^5 rows in all, so every part of that hand-over
runs: offers, images taken by teammates, images withdrawn by the owner, jobs
cancelled when the owner's accounting finds their seed covered.  Results must
match the reference's lattice sums (1e-9) and be bitwise identical between a
first run (queue order from the |A_ps| nbin^2 estimate) and a second one
(queue order from the first run's device times).
"""
import numpy as np
import pytest

from oracle.bindings import default_settings

pytestmark = pytest.mark.gpu

IR_RTOL = 1e-9
# epoch index in finite_source_batch() (2048 curves, seed 3): rows, images
HEAVY = [962656,   # 97,055 rows, 5 images
         1051623,  # 47,353 rows, 5 images
         990488,   # 38,025 rows, 4 images
         838307]   # 39,614 rows, 3 images


def _sub_batch():
    from lcbinint_amd.workloads import finite_source_batch
    curves, offsets, times, nbin = finite_source_batch()
    sub, ts, bins = [], [], []
    for e in HEAVY:
        c = int(np.searchsorted(offsets, e, side="right") - 1)
        lo, hi = max(offsets[c], e - 1), min(offsets[c + 1], e + 2)
        sub.append(curves[c])
        ts.append(times[lo:hi])
        bins.append(int(nbin[c]))
    off = np.concatenate([[0], np.cumsum([len(x) for x in ts])]).astype(np.int64)
    return sub, off, np.concatenate(ts), np.asarray(bins, dtype=np.int32)


def test_team_scans_match_reference_and_repeat_bitwise(gpu, ref):
    import torch
    from lcbinint_amd.capi import Batch, Curve, finite_source
    curves, off, tt, bins = _sub_batch()
    batch = Batch([Curve(**c) for c in curves])
    batch.set_source_bins(bins)
    d_off = torch.from_numpy(off).cuda()
    d_t = torch.from_numpy(tt).cuda()
    runs = []
    for _ in range(2):
        d_m = torch.full((len(tt),), float("nan"), dtype=torch.float64, device="cuda")
        d_k = torch.full((len(tt),), -9, dtype=torch.int32, device="cuda")
        prof = finite_source(batch, d_off.data_ptr(), d_t.data_ptr(), len(tt), d_m.data_ptr(),
                             d_k.data_ptr(), torch.cuda.current_stream().cuda_stream, profile=True)
        torch.cuda.synchronize()
        assert prof[2] == 0, prof
        runs.append((d_m.cpu().numpy(), d_k.cpu().numpy()))
    batch.close()
    (m0, k0), (m1, k1) = runs
    assert np.array_equal(k0, k1)
    assert np.array_equal(m0.view(np.int64), m1.view(np.int64)), "second run differs"
    for ci, cv in enumerate(curves):
        sl = slice(off[ci], off[ci + 1])
        st = default_settings(source_bins=int(bins[ci]), limb_darkening_c=cv["ld_c"])
        exp, emeth, _ = ref.light_curve(st, tt[sl], cv["t0"], cv["tE"], cv["u0"], cv["alpha"],
                                        cv["s"], cv["q"], cv["rho"])
        assert np.array_equal(k0[sl], emeth), (ci, k0[sl], emeth)
        ir = emeth >= 2
        assert ir.any(), ci
        rel = np.abs(m0[sl][ir] - exp[ir]) / np.abs(exp[ir])
        print(f"curve {ci} (epoch {HEAVY[ci]}): {ir.sum()} inverse-ray epochs, worst rel {rel.max():.3e}")
        assert rel.max() <= IR_RTOL, (ci, rel)
        ps = ~ir
        assert np.array_equal(m0[sl][ps].view(np.int64), exp[ps].view(np.int64))
