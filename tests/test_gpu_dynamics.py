"""Light curves with annual parallax and orbital motion on the device against
the reference core (oracle/_ref) on the same Earth table: per epoch the same
method, point-source and hexadecapole values bitwise, inverse-ray values to
1e-9 relative (the same lattices; only the order of the FP64 sums differs)."""
import os
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
TABLE = ROOT / "tests" / "golden" / "earth_ephemeris_substitute.txt"
REF = ROOT / "oracle" / "_ref" / "liblcbinint_ref.so"
os.environ.setdefault("LCBININT_EARTH_EPHEMERIS", str(TABLE))

pytestmark = pytest.mark.gpu


def _curves(n_curves, n_epochs, seed, mode, width=0.6):
    from lcbinint_amd.workloads import dynamic_batch
    curves, dyn, _, _ = dynamic_batch(n_curves=n_curves, seed=seed, min_epochs=2, max_epochs=2)
    rng = np.random.default_rng(seed + 100)
    times = []
    for c, d in zip(curves, dyn):
        if mode == 2:
            d["g3"] = d["g2"] * 0.5
            d["lom_szs"] = 0.3
            d["lom_ar"] = 1.5
        # epochs close to the peak, where the finite source matters
        times.append(np.sort(c["t0"] + c["tE"] * rng.uniform(-width, width, n_epochs)))
    offsets = np.arange(n_curves + 1, dtype=np.int64) * n_epochs
    return curves, dyn, offsets, np.concatenate(times)


@pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_dynamic_light_curves_match_reference(gpu, mode):
    from lcbinint_amd import capi
    from oracle.bindings import default_settings, load_ref
    capi.load_earth_ephemeris(TABLE)
    ref = load_ref(str(REF))
    curves, dyn, offsets, times = _curves(12, 80, 7 + mode, mode, width=0.15)
    for c in curves:
        # large sources close to the lens axis: many finite-source epochs
        c["rho"] = max(c["rho"], 5e-3)
        c["u0"] *= 0.1
    b = capi.Batch([capi.Curve(**c) for c in curves])
    dobj = [capi.Dynamics(**d) for d in dyn]
    print(f"dynamic light curves, orbit mode {mode}: {len(times)} epochs", flush=True)
    mag, meth = capi.light_curves_dynamic(b, dobj, mode, offsets, times)
    b.close()
    n_fs = 0
    for i, (c, d) in enumerate(zip(curves, dyn)):
        sl = slice(offsets[i], offsets[i + 1])
        p = dict(c, **d)
        st = default_settings(source_bins=50, limb_darkening_c=c["ld_c"])
        a, m, _, status = ref.light_curve_dynamic(st, times[sl], p, mode)
        assert status == 0
        assert np.array_equal(meth[sl], m), (i, meth[sl], m)
        ps = m <= 1
        assert np.array_equal(mag[sl][ps].view(np.int64), a[ps].view(np.int64)), i
        ir = m >= 2
        n_fs += int(ir.sum())
        if ir.any():
            rel = np.abs(mag[sl][ir] - a[ir]) / np.abs(a[ir])
            assert rel.max() <= 1e-9, (i, rel.max())
    print(f"inverse-ray epochs compared: {n_fs}", flush=True)


@pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")
def test_lightcurve_api_parallax_orbit(gpu):
    import lcbinint_amd as lcbinint
    from lcbinint_amd import capi
    from oracle.bindings import default_settings, load_ref
    capi.load_earth_ephemeris(TABLE)
    ref = load_ref(str(REF))
    t0 = 10100.0
    params = dict(t0=t0, tE=30.0, u0=0.05, alpha=0.7, s=1.1, q=0.02, rho=3e-3, piEN=0.2, piEE=-0.1,
                  g1=0.3 / 1.1 / 365.25, g2=-0.5 / 365.25, g3=0.0)
    times = np.linspace(t0 - 15, t0 + 15, 301)
    lc = lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=50),
                             limb_darkening=lcbinint.LimbDarkening.linear(0.4),
                             event=lcbinint.EventCoordinates(ra=270.0, dec=-30.0, tfix=t0),
                             orbital_motion_mode=lcbinint.OrbitalMotionMode.CIRCULAR, parallax=True)
    got = lc(times, params)
    a, m, _, status = ref.light_curve_dynamic(default_settings(source_bins=50, limb_darkening_c=0.4), times,
                                              dict(params, ra=270.0, dec=-30.0, tfix=t0), 1)
    assert status == 0
    rel = np.abs(got - a) / np.abs(a)
    assert rel.max() <= 1e-9, rel.max()
    # parallax is read only with parallax=True (python/lcbinint_pybind.cpp:2734-2735)
    lc0 = lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=50),
                              limb_darkening=lcbinint.LimbDarkening.linear(0.4),
                              event=lcbinint.EventCoordinates(ra=270.0, dec=-30.0, tfix=t0),
                              orbital_motion_mode=lcbinint.OrbitalMotionMode.CIRCULAR)
    a0, _, _, _ = ref.light_curve_dynamic(default_settings(source_bins=50, limb_darkening_c=0.4), times,
                                          dict(params, piEN=0.0, piEE=0.0, ra=270.0, dec=-30.0, tfix=t0), 1)
    got0 = lc0(times, params)
    assert np.max(np.abs(got0 - a0) / np.abs(a0)) <= 1e-9


@pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")
def test_dynamic_gradients_match_finite_differences(gpu):
    # derivatives of the magnification with respect to all twelve parameters,
    # trajectory included, against central differences (epochs whose method
    # changes across the step are left out; inverse-ray epochs through the
    # antialiased quadrature, as in the static gradient tests)
    from lcbinint_amd import capi
    capi.load_earth_ephemeris(TABLE)
    curves, dyn, offsets, times = _curves(8, 40, 11, 1, width=0.15)
    for c in curves:
        c["rho"] = max(c["rho"], 5e-3)
        c["u0"] *= 0.1
    b = capi.Batch([capi.Curve(**c) for c in curves])

    def run(cs, ds):
        b.update_curves([capi.Curve(**c) for c in cs])
        return capi.light_curves_dynamic_grad(b, [capi.Dynamics(**d) for d in ds], 1, offsets, times)

    mag, meth, aa, dm = run(curves, dyn)
    assert np.all(np.isfinite(dm))
    ir = meth >= 2
    print(f"dynamic gradient: {len(times)} epochs, {int(ir.sum())} inverse-ray", flush=True)
    for k, name in enumerate(capi.DYN_GRAD_PARAMS):
        v = np.array([c[name] if name in c else d[name] for c, d in zip(curves, dyn)])
        # the nearly edge-on circular orbits of g3 = 0 lose digits to
        # cancellation in g1 and g2, so those take a larger relative step
        rel = 1e-4 if name in ("g1", "g2") else 1e-6
        step = {"t0": 1e-5 * np.array([c["tE"] for c in curves])}.get(name, rel * np.abs(v))
        step = np.where(step == 0.0, 1e-8, step)
        out = []
        for sgn in (1, -1):
            cs = [dict(c) for c in curves]
            ds = [dict(d) for d in dyn]
            for i in range(len(cs)):
                (cs[i] if name in cs[i] else ds[i])[name] = v[i] + sgn * step[i]
            out.append(run(cs, ds))
        h = np.repeat(step, np.diff(offsets))
        fd = (out[0][2] - out[1][2]) / (2 * h)
        same = (out[0][1] == meth) & (out[1][1] == meth)
        scale = np.maximum(np.abs(fd), 1e-3 * np.abs(fd[same]).max() + 1e-12)
        err = np.abs(dm[:, k] - fd) / scale
        for sel, med_bar, label in ((same & ~ir, 1e-4, "point source / hexadecapole"),
                                    (same & ir, 1e-2, "inverse ray")):
            if sel.sum() >= 5:
                assert np.median(err[sel]) <= med_bar, (name, label, np.median(err[sel]), err[sel].max())
    b.close()


def test_dynamic_torch_bridge(gpu):
    import torch
    from lcbinint_amd import capi
    from lcbinint_amd.torch_ops import DYNAMIC_PARAM_NAMES, DynamicLightCurves
    capi.load_earth_ephemeris(TABLE)
    curves, dyn, offsets, times = _curves(6, 50, 13, 1, width=0.3)
    rows = [[c[k] if k in c else d[k] for k in DYNAMIC_PARAM_NAMES] for c, d in zip(curves, dyn)]
    params = torch.tensor(rows, dtype=torch.float64, device="cuda", requires_grad=True)
    lc = DynamicLightCurves(offsets, times, dyn, orbit_mode=capi.ORBIT_CIRCULAR)
    mag = lc(params)
    w = torch.linspace(0.5, 1.5, len(times), dtype=torch.float64, device="cuda")
    (w * mag).sum().backward()
    # the same contraction from the C-ABI derivatives
    b = capi.Batch([capi.Curve(**c) for c in curves])
    m, _, _, dm = capi.light_curves_dynamic_grad(b, [capi.Dynamics(**d) for d in dyn], capi.ORBIT_CIRCULAR,
                                                 offsets, times)
    b.close()
    assert np.array_equal(mag.detach().cpu().numpy(), m)
    exp = np.stack([np.add.reduceat(w.cpu().numpy() * dm[:, k], offsets[:-1]) for k in range(12)], axis=1)
    got = params.grad.cpu().numpy()
    assert np.allclose(got, exp, rtol=1e-12, atol=1e-14 * np.abs(exp).max())
    lc.close()
