"""Light curves with annual parallax and orbital motion on the device against
the reference core (oracle/_ref) on the same Earth table: per epoch the same
method, inverse-ray values to 1e-9 relative (the same lattices; only the order
of the FP64 sums differs), point-source and hexadecapole values bitwise without
orbital motion and to 1e-11 relative with it (the device evaluates the orbit's
cos, sin and atan2 per epoch with its own libm, a few ulp from the host's, so
the source position and separation differ in the last bits).

The device trajectory itself is checked against the host restatement
(lcbg_dynamic_sources, bitwise against the reference in the CPU tests):
bitwise without orbital motion, to a few ulp with it; its Jacobian likewise."""
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
        if mode == 0:
            assert np.array_equal(mag[sl][ps].view(np.int64), a[ps].view(np.int64)), i
        else:
            rel = np.abs(mag[sl][ps] - a[ps]) / np.abs(a[ps])
            assert rel.size == 0 or rel.max() <= 1e-11, (i, rel.max())
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


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_device_trajectory_matches_host(gpu, mode):
    """lcbg_batch_dynamic_sources_device against lcbg_dynamic_sources (and the
    Jacobians), curves with and without parallax."""
    import torch
    from lcbinint_amd import capi
    capi.load_earth_ephemeris(TABLE)
    curves, dyn, offsets, times = _curves(16, 300, 21 + mode, mode, width=3.0)
    for d in dyn[::2]:
        d["piEN"] = d["piEE"] = 0.0
    b = capi.Batch([capi.Curve(**c) for c in curves])
    dobj = (capi.Dynamics * len(dyn))(*[capi.Dynamics(**d) for d in dyn])
    n = len(times)
    d_off = torch.from_numpy(offsets).cuda()
    d_t = torch.from_numpy(times).cuda()
    out = torch.empty((3, n), dtype=torch.float64, device="cuda")
    jac = torch.empty((n, 3, 9), dtype=torch.float64, device="cuda")
    sp = torch.cuda.current_stream().cuda_stream
    print(f"device trajectory, orbit mode {mode}: {n} epochs", flush=True)
    for with_jac in (False, True):
        capi.check(capi.lcbg_batch_dynamic_sources_device(
            b.handle, dobj, mode, d_off.data_ptr(), d_t.data_ptr(), n, out[0].data_ptr(), out[1].data_ptr(),
            out[2].data_ptr(), jac.data_ptr() if with_jac else None, sp), "lcbg_batch_dynamic_sources_device")
        torch.cuda.synchronize()
        got = out.cpu().numpy().T
        got_j = jac.cpu().numpy()
        worst = 0.0
        for i, (c, d) in enumerate(zip(curves, dyn)):
            sl = slice(offsets[i], offsets[i + 1])
            host, host_j = capi.dynamic_sources_jacobian(capi.Curve(**c), capi.Dynamics(**d), mode, times[sl])
            # epochs outside the Earth table are NaN on both sides (the sign
            # bit of a NaN is not meaningful)
            nan = np.isnan(host)
            assert np.array_equal(np.isnan(got[sl]), nan), (i, with_jac)
            g, h = got[sl][~nan], host[~nan]
            if mode == 0:
                assert np.array_equal(g.view(np.int64), h.view(np.int64)), (i, with_jac)
            else:
                err = np.abs(g - h) / np.maximum(np.abs(h), 1.0)
                worst = max(worst, float(err.max(initial=0.0)))
                assert err.max(initial=0.0) <= 1e-13, (i, with_jac, err.max())
            if with_jac:
                jnan = np.isnan(host_j)
                assert np.array_equal(np.isnan(got_j[sl]), jnan), i
                gj, hj = got_j[sl][~jnan], host_j[~jnan]
                if mode == 0:
                    # the same operations; only the signs of zero derivatives may differ
                    assert np.all(gj == hj), i
                else:
                    # an ulp in the orbit phase can add or drop one Newton step
                    # of Kepler's equation (stopping at |dE| <= 1e-13), which
                    # moves the forward-mode derivatives in their ninth digit
                    scale = np.maximum(np.abs(hj), np.abs(hj).max(initial=0.0) * 1e-6 + 1e-300)
                    jerr = np.abs(gj - hj) / scale
                    assert jerr.max(initial=0.0) <= 1e-7, (i, jerr.max())
        print(f"  jacobian {with_jac}: worst relative difference {worst:.2e}", flush=True)
    b.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_device_entry_matches_host_entry(gpu, mode):
    """lcbg_batch_light_curves_dynamic_device (device arrays, values and
    value+grad) gives the host entry points' results bitwise; a second static
    run after the separations change rebuilds the shared caustic caches."""
    import torch
    from lcbinint_amd import capi
    capi.load_earth_ephemeris(TABLE)
    curves, dyn, offsets, times = _curves(10, 120, 31 + mode, mode, width=0.2)
    for c in curves:
        c["rho"] = max(c["rho"], 5e-3)
        c["u0"] *= 0.1
    dlist = [capi.Dynamics(**d) for d in dyn]
    dobj = (capi.Dynamics * len(dyn))(*dlist)
    n = len(times)
    d_off = torch.from_numpy(offsets).cuda()
    d_t = torch.from_numpy(times).cuda()
    sp = torch.cuda.current_stream().cuda_stream
    b = capi.Batch([capi.Curve(**c) for c in curves])
    for rnd in range(2):
        if rnd == 1:
            for c in curves:
                c["s"] *= 1.05
            b.update_curves([capi.Curve(**c) for c in curves])
        print(f"device entry, orbit mode {mode}, pass {rnd}: {n} epochs", flush=True)
        mag, meth = capi.light_curves_dynamic(b, dlist, mode, offsets, times)
        fresh = capi.Batch([capi.Curve(**c) for c in curves])
        mag_f, meth_f = capi.light_curves_dynamic(fresh, dlist, mode, offsets, times)
        fresh.close()
        assert np.array_equal(meth, meth_f) and np.array_equal(mag.view(np.int64), mag_f.view(np.int64))
        m_d = torch.empty(n, dtype=torch.float64, device="cuda")
        k_d = torch.empty(n, dtype=torch.int32, device="cuda")
        capi.check(capi.lcbg_batch_light_curves_dynamic_device(
            b.handle, dobj, mode, d_off.data_ptr(), d_t.data_ptr(), n, m_d.data_ptr(), k_d.data_ptr(), None,
            None, sp), "lcbg_batch_light_curves_dynamic_device")
        torch.cuda.synchronize()
        assert np.array_equal(k_d.cpu().numpy(), meth)
        assert np.array_equal(m_d.cpu().numpy().view(np.int64), mag.view(np.int64))
    g_mag, g_meth, g_aa, g_dm = capi.light_curves_dynamic_grad(b, dlist, mode, offsets, times)
    assert np.array_equal(g_mag.view(np.int64), mag.view(np.int64))
    aa = torch.empty(n, dtype=torch.float64, device="cuda")
    dm = torch.empty((n, 12), dtype=torch.float64, device="cuda")
    capi.check(capi.lcbg_batch_light_curves_dynamic_device(
        b.handle, dobj, mode, d_off.data_ptr(), d_t.data_ptr(), n, m_d.data_ptr(), k_d.data_ptr(), aa.data_ptr(),
        dm.data_ptr(), sp), "lcbg_batch_light_curves_dynamic_device")
    torch.cuda.synchronize()
    assert np.array_equal(m_d.cpu().numpy().view(np.int64), g_mag.view(np.int64))
    assert np.array_equal(aa.cpu().numpy().view(np.int64), g_aa.view(np.int64))
    assert np.array_equal(dm.cpu().numpy().view(np.int64), g_dm.view(np.int64))
    b.close()


def test_lightcurve_keeps_its_handle_across_calls(gpu):
    """LightCurve reuses one device handle for the LensModel path across calls
    (curves updated in place); every call equals a fresh LightCurve's."""
    import lcbinint
    from lcbinint_amd import capi
    capi.load_earth_ephemeris(TABLE)
    t0 = 10100.0
    times = np.linspace(t0 - 10, t0 + 10, 121)

    def make():
        return lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=40),
                                   limb_darkening=lcbinint.LimbDarkening.linear(0.4),
                                   event=lcbinint.EventCoordinates(ra=270.0, dec=-30.0, tfix=t0),
                                   orbital_motion_mode=lcbinint.OrbitalMotionMode.CIRCULAR, parallax=True)

    kept = make()
    for k, s in enumerate((1.1, 0.9, 1.25)):
        params = dict(t0=t0, tE=30.0, u0=0.03 + 0.01 * k, alpha=0.7, s=s, q=0.02, rho=3e-3, piEN=0.2,
                      piEE=-0.1, g1=0.3 / s / 365.25, g2=-0.5 / 365.25)
        a = kept(times, params)
        b = make()(times, params)
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), k
        info = kept.info(times, params)
        assert np.array_equal(np.asarray(info.magnifications).view(np.int64), a.view(np.int64)), k
