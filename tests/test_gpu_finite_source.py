"""Finite-source parity on the GPU against the reference compiled from source.

Per epoch the method (point source, hexadecapole, Cartesian or polar inverse
ray) must equal the reference's; point-source and hexadecapole magnifications
must be bitwise equal, inverse-ray magnifications within 1e-9 relative (same
lattice cells, sums in a different order).
"""
import numpy as np
import pytest

from oracle.bindings import default_settings

pytestmark = pytest.mark.gpu

IR_RTOL = 1e-9


def _gpu_curves(gpu, curves, offsets, times, options=None):
    import torch
    from lcbinint_amd.capi import Batch, Curve, finite_source
    batch = Batch([Curve(**c) for c in curves], options)
    d_off = torch.from_numpy(np.ascontiguousarray(offsets)).cuda()
    d_t = torch.from_numpy(np.ascontiguousarray(times)).cuda()
    d_m = torch.full((len(times),), float("nan"), dtype=torch.float64, device="cuda")
    d_k = torch.full((len(times),), -9, dtype=torch.int32, device="cuda")
    prof = finite_source(batch, d_off.data_ptr(), d_t.data_ptr(), len(times), d_m.data_ptr(),
                         d_k.data_ptr(), torch.cuda.current_stream().cuda_stream, profile=True)
    torch.cuda.synchronize()
    batch.close()
    return d_m.cpu().numpy(), d_k.cpu().numpy(), prof


def _options(gpu, **kw):
    o = gpu.lcbg_default_options()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _settings_for(o, c=0.0):
    return default_settings(source_bins=o.source_bins, caustic_bins=o.caustic_bins,
                            grid_ratio=o.grid_ratio, polar_source_bins=o.polar_source_bins,
                            polar_grid_ratio=o.polar_grid_ratio, finite_mode=o.mode,
                            kinji_threshold=o.point_source_threshold,
                            hex_threshold=o.hexadecapole_threshold,
                            adaptive_hex_threshold=o.adaptive_hex_threshold,
                            limb_darkening_c=c)


def _compare(got, meth, exp, emeth, label):
    assert np.array_equal(meth, emeth), (label, np.nonzero(meth != emeth)[0][:10],
                                         meth[meth != emeth][:10], emeth[meth != emeth][:10])
    ps_hex = emeth <= 1
    bad = np.nonzero(ps_hex & (got.view(np.int64) != exp.view(np.int64)))[0]
    assert len(bad) == 0, (label, "point-source/hexadecapole not bitwise", bad[:10])
    ir = emeth >= 2
    rel = np.abs(got[ir] - exp[ir]) / np.abs(exp[ir])
    worst = float(rel.max()) if rel.size else 0.0
    print(f"{label}: {ir.sum()} inverse-ray epochs, worst rel err {worst:.3e}; methods "
          f"{np.bincount(emeth, minlength=4).tolist()}")
    assert worst <= IR_RTOL, (label, worst, np.nonzero(ir)[0][np.argmax(rel)])


@pytest.mark.parametrize("mode", [4, 1])
def test_plain_curve_matches_reference(gpu, ref, mode):
    from lcbinint_amd.workloads import plain_curve
    cv, t = plain_curve()
    o = _options(gpu, mode=mode)
    got, meth, prof = _gpu_curves(gpu, [cv], np.array([0, len(t)]), t, o)
    assert prof[2] == 0, prof
    exp, emeth, st = ref.light_curve(_settings_for(o, cv["ld_c"]), t, cv["t0"], cv["tE"], cv["u0"],
                                     cv["alpha"], cv["s"], cv["q"], cv["rho"])
    assert st == 0
    _compare(got, meth, exp, emeth, f"plain curve mode {mode}")


@pytest.mark.parametrize("bins", [1000, 1800])
def test_caustic_bins_matches_reference(gpu, ref, bins):
    # 1800 caustic samples per branch no longer fit in LDS: the caustic
    # distances then run one wave per epoch from global memory
    from lcbinint_amd.workloads import plain_curve
    cv, t = plain_curve()
    o = _options(gpu, caustic_bins=bins)
    got, meth, prof = _gpu_curves(gpu, [cv], np.array([0, len(t)]), t, o)
    assert prof[2] == 0, prof
    exp, emeth, st = ref.light_curve(_settings_for(o, cv["ld_c"]), t, cv["t0"], cv["tE"], cv["u0"],
                                     cv["alpha"], cv["s"], cv["q"], cv["rho"])
    assert st == 0
    assert (emeth == 1).any() and (emeth >= 2).any()
    _compare(got, meth, exp, emeth, f"plain curve, {bins} caustic bins")


@pytest.mark.parametrize("kinji,hex_thr", [(5.0, 3.0), (20.0, 1.0), (20.0, 0.0), (2.0, 6.0)])
def test_dispatch_thresholds_match_reference(gpu, ref, kinji, hex_thr):
    # The caustic-distance search stops refining past 3.2 hex_threshold rho and
    # skips the nearest-point scan only when the bounding-box margin
    # (point_source_threshold rho) lies beyond that: (5, 3) and (2, 6) never
    # skip it, (20, 1) skips it more often than the defaults, and (20, 0)
    # (no distance test at all) never caps.
    from lcbinint_amd.workloads import finite_source_batch, plain_curve
    curves, offsets, times, _ = finite_source_batch()
    sizes = np.diff(offsets)
    pick = [i for i in range(len(curves)) if sizes[i] <= 300 and curves[i]["rho"] <= 3e-3][:6]
    cv0, t0 = plain_curve()
    sel = [cv0] + [curves[i] for i in pick]
    ts = [t0] + [times[offsets[i]:offsets[i + 1]] for i in pick]
    off = np.concatenate([[0], np.cumsum([len(t) for t in ts])]).astype(np.int64)
    tt = np.concatenate(ts)
    o = _options(gpu, point_source_threshold=kinji, hexadecapole_threshold=hex_thr)
    got, meth, prof = _gpu_curves(gpu, sel, off, tt, o)
    assert prof[2] == 0, prof
    for ci, cv in enumerate(sel):
        sl = slice(off[ci], off[ci + 1])
        exp, emeth, st = ref.light_curve(_settings_for(o, cv["ld_c"]), tt[sl], cv["t0"], cv["tE"],
                                         cv["u0"], cv["alpha"], cv["s"], cv["q"], cv["rho"])
        assert st == 0
        _compare(got[sl], meth[sl], exp, emeth, f"kinji {kinji}, hex {hex_thr}, curve {ci}")


def test_forced_inverse_ray_positions(gpu, ref):
    # bbox exit, derivative and hexadecapole switched off as binary_inverse_ray
    # does (python/lcbinint_pybind.cpp:3826-3847)
    rng = np.random.default_rng(31)
    curves, times = [], []
    for k in range(24):
        s = float(np.exp(rng.uniform(np.log(0.5), np.log(2.0))))
        q = float(10 ** rng.uniform(-3, 0))
        rho = float(10 ** rng.uniform(-3, -1.5))
        c = float(rng.uniform(0.0, 0.8)) if k % 3 else 0.0
        curves.append(dict(t0=0.0, tE=1.0, u0=float(rng.uniform(-0.3, 0.3)),
                           alpha=float(rng.uniform(0, 6.28)), s=s, q=q, rho=rho, ld_c=c, ld_d=0.0))
        times.append(np.linspace(-0.6, 0.6, 9))
    offsets = np.arange(len(curves) + 1) * 9
    times = np.concatenate(times)
    for mode in (1, 2):
        o = _options(gpu, mode=mode, point_source_threshold=1e18, hexadecapole_threshold=0.0,
                     adaptive_hex_threshold=0.0)
        got, meth, prof = _gpu_curves(gpu, curves, offsets, times, o)
        assert prof[2] == 0, prof
        for ci, cv in enumerate(curves):
            sl = slice(offsets[ci], offsets[ci + 1])
            exp, emeth, st = ref.light_curve(_settings_for(o, cv["ld_c"]), times[sl], cv["t0"],
                                             cv["tE"], cv["u0"], cv["alpha"], cv["s"], cv["q"],
                                             cv["rho"])
            _compare(got[sl], meth[sl], exp, emeth, f"forced IR mode {mode} curve {ci}")


def test_batched_subset_matches_reference(gpu, ref):
    from lcbinint_amd.workloads import finite_source_batch
    curves, offsets, times, nbin = finite_source_batch(n_curves=48, seed=3)
    # one options set per call: group curves by nbin
    for nb in sorted(set(nbin.tolist()))[:6]:
        idx = [i for i in range(len(curves)) if nbin[i] == nb]
        sub = [curves[i] for i in idx]
        ts = [times[offsets[i]:offsets[i + 1]][::7] for i in idx]
        off = np.concatenate([[0], np.cumsum([len(x) for x in ts])])
        tt = np.concatenate(ts)
        o = _options(gpu, source_bins=int(nb))
        got, meth, prof = _gpu_curves(gpu, sub, off, tt, o)
        assert prof[2] == 0, prof
        for k, cv in enumerate(sub):
            sl = slice(off[k], off[k + 1])
            exp, emeth, st = ref.light_curve(_settings_for(o, cv["ld_c"]), tt[sl], cv["t0"],
                                             cv["tE"], cv["u0"], cv["alpha"], cv["s"], cv["q"],
                                             cv["rho"])
            _compare(got[sl], meth[sl], exp, emeth, f"batch nbin {nb} curve {idx[k]}")
