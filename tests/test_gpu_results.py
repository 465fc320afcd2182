"""Per-epoch result records, run status and the first failing epoch on the GPU.

The records (lcbg_result: the fields of the reference's lcbi_result plus the
method) are compared with the reference compiled from source, running its
evaluate_binary_light_curve_info loop (python/lcbinint_pybind.cpp:1176-1253)
on the same curves: methods, image counts, source positions and point-source
magnifications exactly; magnifications bitwise for point source and
hexadecapole, 1e-9 relative for inverse ray; error estimates bitwise for the
point-source derivative check and the hexadecapole, 1e-12 relative for the
Cartesian area indicator (its inputs, the row, gap, overlap and fold counts of
the scan, are compared exactly with the reference's LegacyAreaDiagnostics);
converged flags exactly, under zero and nonzero tolerances.
"""
import ctypes

import numpy as np
import pytest

from oracle.bindings import default_settings

pytestmark = pytest.mark.gpu

IR_RTOL = 1e-9
ERR_RTOL = 1e-12

README = dict(t0=0.0, tE=1.0, u0=-0.01, alpha=0.5, s=0.95, q=1e-2, rho=5e-3, ld_c=0.5, ld_d=0.0)


def _records(gpu, cv, times, mode=4, nbin=50, tol=0.0, reltol=0.0):
    o = gpu.lcbg_default_options()
    o.mode = mode
    o.source_bins = nbin
    o.finite_source_tol = tol
    o.finite_source_reltol = reltol
    rec, st, first = gpu.light_curve_results(gpu.Curve(**cv), times, o)
    return rec, st, first


def _reference(ref, cv, times, mode=4, nbin=50, tol=0.0, reltol=0.0):
    st = default_settings(source_bins=nbin, finite_mode=mode, limb_darkening_c=cv["ld_c"])
    out, status = ref.light_curve_info(st, times, cv["t0"], cv["tE"], cv["u0"], cv["alpha"], cv["s"],
                                       cv["q"], cv["rho"], tol, reltol)
    assert status == 0
    return out


def _compare_records(rec, exp, label):
    meth = rec["method"]
    assert np.array_equal(meth, exp["method"]), (label, np.nonzero(meth != exp["method"])[0][:10])
    for k in ("point_source_magnification", "source_x", "source_y"):
        assert np.array_equal(rec[k].view(np.int64), exp[k].view(np.int64)), (label, k)
    assert np.array_equal(rec["image_count"], exp["image_count"]), label
    assert np.array_equal(rec["refinement_level"], exp["refinement_level"]), label
    ps_hex = meth <= 1
    for k in ("magnification", "finite_source_magnification"):
        assert np.array_equal(rec[k][ps_hex].view(np.int64), exp[k][ps_hex].view(np.int64)), (label, k)
        ir = ~ps_hex
        rel = np.abs(rec[k][ir] - exp[k][ir]) / np.abs(exp[k][ir])
        assert rel.size == 0 or rel.max() <= IR_RTOL, (label, k, rel.max())
    e, x = rec["error_estimate"], exp["error_estimate"]
    assert np.array_equal(e[meth != 2].view(np.int64), x[meth != 2].view(np.int64)), (label, "err")
    cart = meth == 2
    rel = np.abs(e[cart] - x[cart]) / np.maximum(np.abs(x[cart]), 1e-300)
    assert rel.size == 0 or rel.max() <= ERR_RTOL, (label, "cartesian err", rel.max())
    assert np.array_equal(rec["converged"], exp["converged"]), (
        label, np.nonzero(rec["converged"] != exp["converged"])[0][:10])
    print(f"{label}: methods {np.bincount(meth, minlength=4).tolist()}, Cartesian err worst rel "
          f"{float(rel.max()) if rel.size else 0.0:.2e}, converged {int(rec['converged'].sum())}/{len(meth)}")


@pytest.mark.parametrize("mode", [4, 1])
@pytest.mark.parametrize("tolerance", ["zero", "median"])
def test_readme_curve_records_match_reference(gpu, ref, mode, tolerance):
    times = np.linspace(-0.5, 0.5, 200)
    reltol = 0.0
    if tolerance == "median":
        # the median relative error of the Cartesian epochs: about half of
        # them converge under it
        r0 = _reference(ref, README, times, mode=mode)
        cart = r0["method"] == 2
        reltol = float(np.median(r0["error_estimate"][cart] / np.maximum(np.abs(r0["magnification"][cart]), 1.0)))
    rec, st, first = _records(gpu, README, times, mode=mode, reltol=reltol)
    assert st == 0 and first == -1
    exp = _reference(ref, README, times, mode=mode, reltol=reltol)
    _compare_records(rec, exp, f"README curve mode {mode} reltol {reltol}")
    if mode == 1:
        assert (rec["method"] == 2).sum() > 30
    if reltol > 0.0:
        cart = rec["method"] == 2
        # the tolerance separates converged from unconverged Cartesian epochs
        assert 0 < rec["converged"][cart].sum() < cart.sum()


def test_batched_curve_records_match_reference(gpu, ref):
    from lcbinint_amd.workloads import finite_source_batch
    curves, offsets, times, nbin = finite_source_batch()
    rng = np.random.default_rng(11)
    picks = rng.choice(len(curves), 24, replace=False)
    n_cart = 0
    for c in picks:
        t = times[offsets[c]:offsets[c + 1]][::3]
        rec, st, first = _records(gpu, curves[c], t, nbin=int(nbin[c]), reltol=5e-3)
        assert st == 0 and first == -1, (c, st, first)
        exp = _reference(ref, curves[c], t, nbin=int(nbin[c]), reltol=5e-3)
        _compare_records(rec, exp, f"curve {c}")
        n_cart += int((rec["method"] == 2).sum())
    assert n_cart > 20


def test_cartesian_diagnostics_match_reference(gpu, ref):
    """The scan's diagnostic counts per Cartesian epoch, against the reference's
    LegacyAreaDiagnostics from inverse_ray_cartesian_binary_mag on the
    reference's own seeds."""
    import torch
    cv = README
    times = np.linspace(-0.5, 0.5, 200)
    o = gpu.lcbg_default_options()
    o.mode = 1
    batch = gpu.Batch([gpu.Curve(**cv)], o)
    n = len(times)
    cost = torch.zeros(15 * n, dtype=torch.int64, device="cuda")
    gpu.check(gpu.lcbg_batch_set_epoch_cost(batch.handle, cost.data_ptr()), "set_epoch_cost")
    d_off = torch.tensor([0, n], dtype=torch.int64, device="cuda")
    d_t = torch.from_numpy(times).cuda()
    d_m = torch.empty(n, dtype=torch.float64, device="cuda")
    d_k = torch.empty(n, dtype=torch.int32, device="cuda")
    gpu.finite_source(batch, d_off.data_ptr(), d_t.data_ptr(), n, d_m.data_ptr(), d_k.data_ptr(),
                      torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    batch.close()
    c = cost.cpu().numpy().reshape(15, n)
    meth = d_k.cpu().numpy()
    st = default_settings(source_bins=50, finite_mode=1, limb_darkening_c=cv["ld_c"])
    ca, sa = np.cos(cv["alpha"]), np.sin(cv["alpha"])
    checked = 0
    for e in np.nonzero(meth == 2)[0]:
        x = times[e] * ca - cv["u0"] * sa
        y = times[e] * sa + cv["u0"] * ca
        seeds = ref.image_seeds(st, cv["s"], 1.0 / cv["q"], x, y, cv["rho"])
        _, diag = ref.cartesian_area(st, cv["s"], 1.0 / cv["q"], x, y, cv["rho"], seeds)
        # diag: seed_count, processed, fold, rows, gaps, overlaps, max_jump, error
        got = [c[5, e], c[6, e], c[7, e], c[8, e], c[9, e]]
        assert got == [int(v) for v in diag[1:6]], (e, got, diag)
        jump = np.int64(c[10, e]).view(np.float64)
        assert jump == diag[6], (e, jump, diag[6])
        checked += 1
    assert checked > 30


def test_last_run_status_and_work_budget(gpu):
    """A small work budget makes the largest inverse-ray epochs give up
    normally: their magnification is NaN, the status is a numerical error with
    the budget flag, the first failing epoch is reported, and every other
    epoch equals the run without the budget."""
    import torch
    cv = README
    times = np.linspace(-0.5, 0.5, 200)
    n = len(times)
    o = gpu.lcbg_default_options()
    o.mode = 1
    d_off = torch.tensor([0, n], dtype=torch.int64, device="cuda")
    d_t = torch.from_numpy(times).cuda()
    sp = torch.cuda.current_stream().cuda_stream
    runs = []
    for budget in (0, 400):
        batch = gpu.Batch([gpu.Curve(**cv)], o)
        gpu.check(gpu.lcbg_batch_set_work_budget(batch.handle, budget), "set_work_budget")
        d_m = torch.empty(n, dtype=torch.float64, device="cuda")
        d_k = torch.empty(n, dtype=torch.int32, device="cuda")
        gpu.finite_source(batch, d_off.data_ptr(), d_t.data_ptr(), n, d_m.data_ptr(), d_k.data_ptr(), sp)
        torch.cuda.synchronize()
        runs.append((d_m.cpu().numpy(), gpu.last_run_status(batch)))
        batch.close()
    (full, (st0, fl0, first0)), (cut, (st1, fl1, first1)) = runs
    assert (st0, fl0, first0) == (0, 0, -1)
    bad = np.nonzero(~np.isfinite(cut))[0]
    assert len(bad) > 0, "the budget abandoned no epoch"
    assert st1 == 2 and (fl1 & 16), (st1, fl1)
    assert first1 == bad[0], (first1, bad[:5])
    ok = np.isfinite(cut)
    assert np.array_equal(cut[ok].view(np.int64), full[ok].view(np.int64))
    print(f"budget 400: {len(bad)} of {n} epochs abandoned, first {first1}")
    # the host path reports the same first failing epoch through its status
    o2 = gpu.lcbg_default_options()
    o2.mode = 1
    rec, st, first = gpu.light_curve_results(gpu.Curve(**cv), times, o2)
    assert st == 0 and first == -1 and np.isfinite(rec["magnification"]).all()
