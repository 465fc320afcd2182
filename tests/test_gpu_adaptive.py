"""Adaptive inverse ray on the GPU against the reference compiled from source.

With adaptive_source_bins = 1 and a tolerance, fixed_inverse_ray_binary
(finite_source_magnifier.cpp:3071-3296) records a refinement tape during the
Cartesian scan, may rescan an epoch at a finer lattice (the topology retry,
:1108-1121, 3166-3196) and refines the tape's cells largest error first; the
dispatcher widens its bounding-box margin and takes its point-source and
hexadecapole thresholds from the tolerances (:4679-4790).  Each test runs the
reference's info loop (python/lcbinint_pybind.cpp:1176-1253) on the same
curves and compares every epoch's record: methods, refinement levels and
converged flags exactly; point-source and hexadecapole values bitwise;
inverse-ray magnifications to 1e-9 relative (the refined value is the area
plus the refinement's correction; the area's row sums are added in another
order than the reference's); error estimates to 1e-9 relative.
"""
import numpy as np
import pytest

from oracle.bindings import default_settings

pytestmark = pytest.mark.gpu

IR_RTOL = 1e-9
ERR_RTOL = 1e-9

README = dict(t0=0.0, tE=1.0, u0=-0.01, alpha=0.5, s=0.95, q=1e-2, rho=5e-3, ld_c=0.5, ld_d=0.0)
# A curve whose epoch 19 of linspace(-0.05, 0.05, 21) at nbin 40 shows the
# global topology pattern (>= 16 seeds, >= 100 overlaps, >= 1000 gap repairs,
# a jump of >= 1000 cells) and is rescanned at 80 bins by the reference.
RETRY = dict(t0=0.0, tE=1.0, u0=0.005, alpha=1.2, s=0.5, q=0.3, rho=4e-3, ld_c=0.5, ld_d=0.0)


def _run(gpu, ref, cv, times, nbin=50, mode=4, tol=0.0, reltol=0.0, adaptive=1, max_bins=400):
    o = gpu.lcbg_default_options()
    o.mode = mode
    o.source_bins = nbin
    o.finite_source_tol = tol
    o.finite_source_reltol = reltol
    o.adaptive_source_bins = adaptive
    o.max_source_bins = max_bins
    rec, st, first = gpu.light_curve_results(gpu.Curve(**cv), times, o)
    assert st == 0 and first == -1, (st, first)
    s = default_settings(source_bins=nbin, finite_mode=mode, limb_darkening_c=cv["ld_c"],
                         adaptive_source_bins=adaptive, max_source_bins=max_bins)
    exp, status = ref.light_curve_info(s, times, cv["t0"], cv["tE"], cv["u0"], cv["alpha"], cv["s"],
                                       cv["q"], cv["rho"], tol, reltol)
    assert status == 0
    return rec, exp


def _compare(rec, exp, label):
    meth = rec["method"]
    assert np.array_equal(meth, exp["method"]), (label, np.nonzero(meth != exp["method"])[0][:10])
    lv = rec["refinement_level"]
    assert np.array_equal(lv, exp["refinement_level"]), (
        label, np.nonzero(lv != exp["refinement_level"])[0][:10])
    assert np.array_equal(rec["image_count"], exp["image_count"]), label
    ps_hex = meth <= 1
    m, x = rec["magnification"], exp["magnification"]
    assert np.array_equal(m[ps_hex].view(np.int64), x[ps_hex].view(np.int64)), (label, "ps/hex")
    rel = np.abs(m[~ps_hex] - x[~ps_hex]) / np.abs(x[~ps_hex])
    assert rel.size == 0 or rel.max() <= IR_RTOL, (label, rel.max())
    e, xe = rec["error_estimate"], exp["error_estimate"]
    assert np.array_equal(e[meth != 2].view(np.int64), xe[meth != 2].view(np.int64)), (label, "err")
    cart = meth == 2
    erel = np.abs(e[cart] - xe[cart]) / np.maximum(np.abs(xe[cart]), 1e-300)
    assert erel.size == 0 or erel.max() <= ERR_RTOL, (label, "cartesian err", erel.max())
    assert np.array_equal(rec["converged"], exp["converged"]), (
        label, np.nonzero(rec["converged"] != exp["converged"])[0][:10])
    print(f"{label}: methods {np.bincount(meth, minlength=4).tolist()}, levels "
          f"{np.bincount(lv).tolist()}, IR worst rel {float(rel.max()) if rel.size else 0.0:.1e}, "
          f"Cartesian err worst rel {float(erel.max()) if erel.size else 0.0:.1e}, "
          f"converged {int(rec['converged'].sum())}/{len(meth)}", flush=True)
    return int((lv > 0).sum())


@pytest.mark.parametrize("mode", [4, 1])
@pytest.mark.parametrize("tols", [(0.0, 1e-3), (0.0, 1e-4), (1e-3, 0.0)])
def test_readme_curve_adaptive_matches_reference(gpu, ref, mode, tols):
    times = np.linspace(-0.5, 0.5, 200)
    tol, reltol = tols
    rec, exp = _run(gpu, ref, README, times, mode=mode, tol=tol, reltol=reltol)
    refined = _compare(rec, exp, f"README curve mode {mode} tol {tol} reltol {reltol}")
    assert refined > 5


def test_uniform_source_adaptive_matches_reference(gpu, ref):
    # weights 0 or 1: the refinement queue sees many equal error weights, so
    # the order in which ties leave it decides which cells are split
    cv = dict(README, ld_c=0.0)
    times = np.linspace(-0.5, 0.5, 200)
    rec, exp = _run(gpu, ref, cv, times, mode=1, reltol=1e-3)
    assert _compare(rec, exp, "uniform source") > 5


@pytest.mark.parametrize("nbin", [20, 31, 64])
def test_small_and_large_lattices_match_reference(gpu, ref, nbin):
    # nbin < 32: safety 8 and never converged; nbin 64: fewer levels to max_source_bins
    times = np.linspace(-0.5, 0.5, 120)
    rec, exp = _run(gpu, ref, README, times, nbin=nbin, mode=1, reltol=1e-3, max_bins=200)
    _compare(rec, exp, f"README curve nbin {nbin}")
    if nbin < 32:
        assert not rec["converged"][rec["method"] == 2].any()


@pytest.mark.parametrize("reltol", [0.0, 1e-3])
def test_topology_retry_matches_reference(gpu, ref, reltol):
    times = np.linspace(-0.05, 0.05, 21)
    rec, exp = _run(gpu, ref, RETRY, times, nbin=40, reltol=reltol)
    _compare(rec, exp, f"retry curve reltol {reltol}")
    # the retried epoch; without a tolerance it is the only one at level 1
    assert rec["refinement_level"][19] == 1
    if reltol == 0.0:
        assert int((rec["refinement_level"] > 0).sum()) == 1
    # adaptive_source_bins = 0: no retry, the plain scan's value
    plain, exp0 = _run(gpu, ref, RETRY, times, nbin=40, adaptive=0)
    _compare(plain, exp0, "retry curve, plain")
    assert plain["magnification"][19] != rec["magnification"][19]


def test_batched_curves_adaptive_match_reference(gpu, ref):
    from lcbinint_amd.workloads import finite_source_batch
    curves, offsets, times, nbin = finite_source_batch()
    rng = np.random.default_rng(12)
    picks = rng.choice(len(curves), 24, replace=False)
    refined = 0
    for c in picks:
        t = times[offsets[c]:offsets[c + 1]][::4]
        rec, exp = _run(gpu, ref, curves[c], t, nbin=int(nbin[c]), reltol=1e-3)
        refined += _compare(rec, exp, f"curve {c} nbin {nbin[c]}")
    assert refined > 20


def test_gradient_rejects_adaptive(gpu):
    import ctypes

    import torch

    from lcbinint_amd import capi
    o = capi.lcbg_default_options()
    o.adaptive_source_bins = 1
    o.finite_source_reltol = 1e-3
    t = torch.linspace(-0.5, 0.5, 16, dtype=torch.float64, device="cuda")
    off = torch.tensor([0, 16], dtype=torch.int64, device="cuda")
    mag = torch.empty_like(t)
    meth = torch.empty(16, dtype=torch.int32, device="cuda")
    dmag = torch.empty(16 * 8, dtype=torch.float64, device="cuda")
    b = capi.Batch([capi.Curve(**README)], o)
    try:
        st = capi.lcbg_batch_finite_source_grad_device(
            b.handle, ctypes.c_void_p(off.data_ptr()), ctypes.c_void_p(t.data_ptr()), 16,
            ctypes.c_void_p(mag.data_ptr()), ctypes.c_void_p(meth.data_ptr()), None,
            ctypes.c_void_p(dmag.data_ptr()), None)
        assert capi.STATUS_NAMES[st] == "unsupported"
    finally:
        b.close()


def test_light_curve_api_adaptive_matches_reference(ref):
    import lcbinint

    times = np.linspace(-0.5, 0.5, 200)
    lc = lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=50, adaptive_source_bins=1,
                                                      reltol=1e-3),
                             limb_darkening=lcbinint.LimbDarkening.linear(0.5))
    params = dict(t0=0.0, tE=1.0, u0=-0.01, alpha=0.5, s=0.95, q=1e-2, rho=5e-3)
    got = np.asarray(lc(times, params))
    s = default_settings(source_bins=50, limb_darkening_c=0.5, adaptive_source_bins=1)
    exp, status = ref.light_curve_info(s, times, 0.0, 1.0, -0.01, 0.5, 0.95, 1e-2, 5e-3, 0.0, 1e-3)
    assert status == 0
    rel = np.abs(got / exp["magnification"] - 1.0)
    assert rel.max() <= IR_RTOL, rel.max()
    info = lc.info(times, params)
    assert np.array_equal(np.asarray(info.finite_source_refinement_levels), exp["refinement_level"])
    assert np.array_equal(np.asarray(info.finite_source_converged, dtype=np.int32), exp["converged"])
