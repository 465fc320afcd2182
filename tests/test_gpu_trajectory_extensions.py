"""Terrestrial parallax, xallarap and binary sources on the GPU against the
reference's own C API (lcbi_magnification_array, src/lcbinint/lcbinint.cpp:88-119,
which runs LensModel::magnification over the epochs) compiled from source, on the
substitute Earth table.

The device evaluates the new trajectory terms per epoch
(src/lcbinint/model/trajectory.cpp:287-401): the terrestrial hour angle's cos
and sin, the xallarap phase's cos and sin and its Kepler solve come from the
device libm, a few ulp from the host's, so source positions agree to ~1e-13
and magnifications to the contract's 1e-11 (point source, hexadecapole) and
1e-9 (inverse ray) relative.  A binary source is the flux-weighted blend of two
such curves at (u0, t0) and (u0_2, t0_2), computed as the reference's
light_curve_from_dict computes it (python/lcbinint_pybind.cpp:2325-2356)."""
import os
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
TABLE = ROOT / "tests" / "golden" / "earth_ephemeris_substitute.txt"
REF = ROOT / "oracle" / "_ref" / "liblcbinint_ref.so"
os.environ.setdefault("LCBININT_EARTH_EPHEMERIS", str(TABLE))

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")]

PS_RTOL = 1e-11
IR_RTOL = 1e-9
# With orbital motion the separation s(t) comes from the device's cos, sin and
# atan2 of the orbit phase, an ulp or so from the host's.  An inverse-ray epoch
# with a lattice cell that close to the source edge (the contract's exception
# for rays within 1e-12 of the edge) can then differ from the reference by
# that cell and what the flood fill reaches through it.  Such an epoch (past
# IR_RTOL) is evaluated again by the reference at the device's own source
# position and separation (ref_lens_model_at), which must give the device's
# method and value to IR_RTOL.


def _capi_ref():
    from oracle.bindings import RefCApi
    return RefCApi(str(REF))


def _curves(n, seed, kind):
    """n curves near their peaks with large sources (many finite-source epochs)
    and the trajectory extension `kind`."""
    rng = np.random.default_rng(seed)
    curves, dyns = [], []
    for _ in range(n):
        t0 = rng.uniform(10080.0, 10120.0)
        tE = np.exp(rng.uniform(np.log(10.0), np.log(60.0)))
        c = dict(t0=t0, tE=tE, u0=rng.uniform(-0.05, 0.05), alpha=rng.uniform(0, 2 * np.pi),
                 s=np.exp(rng.uniform(np.log(0.6), np.log(1.6))), q=np.exp(rng.uniform(np.log(1e-3), 0.0)),
                 rho=np.exp(rng.uniform(np.log(2e-3), np.log(1e-2))), ld_c=rng.uniform(0.1, 0.9), ld_d=0.0)
        d = dict(ra=rng.uniform(250, 280), dec=rng.uniform(-35, -20), tfix=t0, lom_ar=1.0)
        if kind == "terrestrial":
            d.update(piEN=rng.uniform(-0.5, 0.5), piEE=rng.uniform(-0.5, 0.5), obs_lat=rng.uniform(-40, 40),
                     obs_lon=rng.uniform(-180, 180))
        elif kind == "angular":
            d.update(xallarap_type=1, xi_1=rng.uniform(-0.1, 0.1), xi_2=rng.uniform(-0.1, 0.1),
                     omega_xa=rng.uniform(0.05, 0.5), inc_xa=rng.uniform(0, np.pi), phi_xa=rng.uniform(0, 2 * np.pi))
        elif kind == "kepler":
            d.update(xallarap_type=2, piEN_xa=rng.uniform(-0.1, 0.1), piEE_xa=rng.uniform(-0.1, 0.1),
                     period_xa=rng.uniform(10, 100), ecc_xa=rng.uniform(0, 0.7), peri_xa=rng.uniform(0, 2 * np.pi),
                     inc_xa=rng.uniform(0, np.pi))
        elif kind == "all":
            d.update(piEN=rng.uniform(-0.5, 0.5), piEE=rng.uniform(-0.5, 0.5), obs_lat=rng.uniform(-40, 40),
                     obs_lon=rng.uniform(-180, 180), xallarap_type=1, xi_1=rng.uniform(-0.1, 0.1),
                     xi_2=rng.uniform(-0.1, 0.1), omega_xa=rng.uniform(0.05, 0.5), inc_xa=rng.uniform(0, np.pi),
                     phi_xa=rng.uniform(0, 2 * np.pi), g1=rng.uniform(-1, 1) / 365.25, g2=rng.uniform(-1, 1) / 365.25)
        curves.append(c)
        dyns.append(d)
    return curves, dyns


def _reference(api, c, d, times, orbit_mode):
    p = api.params(t0=c["t0"], tE=c["tE"], umin=c["u0"], theta=c["alpha"], sep=c["s"], q=c["q"], rho=c["rho"],
                   limb_darkening_c=c["ld_c"], piEN=d.get("piEN", 0.0), piEE=d.get("piEE", 0.0), ra=d["ra"],
                   dec=d["dec"], tfix=d["tfix"], obs_lat=d.get("obs_lat", 0.0), obs_lon=d.get("obs_lon", 0.0),
                   orbital_motion_mode=orbit_mode, g1=d.get("g1", 0.0), g2=d.get("g2", 0.0),
                   xi_1=d.get("xi_1", 0.0), xi_2=d.get("xi_2", 0.0), omega_xa=d.get("omega_xa", 0.0),
                   inc_xa=d.get("inc_xa", 0.0), phi_xa=d.get("phi_xa", 0.0), piEN_xa=d.get("piEN_xa", 0.0),
                   piEE_xa=d.get("piEE_xa", 0.0), period_xa=d.get("period_xa", 0.0), ecc_xa=d.get("ecc_xa", 0.0),
                   peri_xa=d.get("peri_xa", 0.0))
    o = api.options(xallarap_param_type=d.get("xallarap_type", 0))
    rec, st = api.curve(times, p, o)
    assert st == 0, st
    return rec


@pytest.mark.parametrize("kind,orbit_mode", [("terrestrial", 0), ("angular", 0), ("kepler", 0), ("all", 1)])
def test_extended_trajectories_match_reference(gpu, kind, orbit_mode):
    from lcbinint_amd import capi
    capi.load_earth_ephemeris(TABLE)
    api = _capi_ref()
    curves, dyns = _curves(10, {"terrestrial": 1, "angular": 2, "kepler": 3, "all": 4}[kind], kind)
    rng = np.random.default_rng(9)
    n_ep = 60
    times = np.concatenate([np.sort(c["t0"] + c["tE"] * rng.uniform(-0.12, 0.12, n_ep)) for c in curves])
    offsets = np.arange(len(curves) + 1, dtype=np.int64) * n_ep
    import ctypes

    import torch

    from oracle.bindings import default_settings, load_ref
    b = capi.Batch([capi.Curve(**c) for c in curves])
    dobj = [capi.Dynamics(**d) for d in dyns]
    print(f"{kind}, orbit mode {orbit_mode}: {len(times)} epochs", flush=True)
    rec, st = capi.light_curves_dynamic_results(b, dobj, orbit_mode, offsets, times)
    assert st == 0, st
    # the device trajectory's separation per epoch
    d_off = torch.from_numpy(offsets).cuda()
    d_t = torch.from_numpy(times).cuda()
    xs, ys, ss = (torch.empty_like(d_t) for _ in range(3))
    capi.check(capi.lcbg_batch_dynamic_sources_device(
        b.handle, (capi.Dynamics * len(dobj))(*dobj), orbit_mode, ctypes.c_void_p(d_off.data_ptr()),
        ctypes.c_void_p(d_t.data_ptr()), len(times), ctypes.c_void_p(xs.data_ptr()),
        ctypes.c_void_p(ys.data_ptr()), ctypes.c_void_p(ss.data_ptr()), None, None), "dynamic sources")
    torch.cuda.synchronize()
    sep = ss.cpu().numpy()
    b.close()
    ref = load_ref(str(REF))
    meth = rec["method"]
    worst_ps = worst_ir = worst_src = 0.0
    for i, (c, d) in enumerate(zip(curves, dyns)):
        sl = slice(offsets[i], offsets[i + 1])
        exp = _reference(api, c, d, times[sl], orbit_mode)
        m, x = rec["magnification"][sl], exp["magnification"]
        rel = np.abs(m / x - 1.0)
        ps = meth[sl] <= 1
        worst_ps = max(worst_ps, float(rel[ps].max()) if ps.any() else 0.0)
        ir_rel = np.where(ps, 0.0, rel)
        if orbit_mode != 0:
            for k in np.nonzero(ir_rel > IR_RTOL)[0]:
                e = offsets[i] + k
                a, mth = ref.lens_model_at(default_settings(limb_darkening_c=c["ld_c"]), float(sep[e]), c["q"],
                                           float(rec["source_x"][e]), float(rec["source_y"][e]), c["rho"])
                print(f"  curve {i} epoch {k}: {ir_rel[k]:.1e} from the reference's trajectory; at the "
                      f"device's geometry {abs(m[k] / a - 1.0):.1e} (method {mth} / {meth[e]})", flush=True)
                assert mth == meth[e] and abs(m[k] / a - 1.0) <= IR_RTOL
                ir_rel[k] = 0.0
        worst_ir = max(worst_ir, float(ir_rel.max()))
        for k in ("source_x", "source_y"):
            worst_src = max(worst_src, float(np.abs(rec[k][sl] - exp[k]).max()))
        assert np.array_equal(rec["image_count"][sl], exp["image_count"]), (i, kind)
    n_fs = int((meth >= 2).sum())
    print(f"  methods {np.bincount(meth, minlength=4).tolist()}; worst rel PS/hex {worst_ps:.1e}, IR "
          f"{worst_ir:.1e}; worst |source| difference {worst_src:.1e}", flush=True)
    assert worst_ps <= PS_RTOL and worst_ir <= IR_RTOL and worst_src <= 1e-12
    assert n_fs > 20


def test_extension_moves_the_curve(gpu):
    # the terms are not no-ops: each changes the light curve by more than the
    # tolerance against the same curve without it
    from lcbinint_amd import capi
    capi.load_earth_ephemeris(TABLE)
    for kind in ("terrestrial", "angular", "kepler"):
        curves, dyns = _curves(3, 20, kind)
        base = []
        for d in dyns:
            e = {k: v for k, v in d.items() if k in ("ra", "dec", "tfix", "lom_ar", "piEN", "piEE")}
            base.append(e)
        times = np.concatenate([np.linspace(c["t0"] - 2, c["t0"] + 2, 40) for c in curves])
        offsets = np.arange(4, dtype=np.int64) * 40
        b = capi.Batch([capi.Curve(**c) for c in curves])
        m1, _ = capi.light_curves_dynamic(b, [capi.Dynamics(**d) for d in dyns], 0, offsets, times)
        m0, _ = capi.light_curves_dynamic(b, [capi.Dynamics(**d) for d in base], 0, offsets, times)
        b.close()
        assert np.abs(m1 / m0 - 1.0).max() > 1e-6, kind


def test_mismatched_xallarap_form_is_unsupported(gpu):
    from lcbinint_amd import capi
    curves, dyns = _curves(1, 30, "angular")
    dyns[0]["xallarap_type"] = 2
    b = capi.Batch([capi.Curve(**curves[0])])
    t = np.linspace(curves[0]["t0"] - 1, curves[0]["t0"] + 1, 8)
    rec, st = capi.light_curves_dynamic_results(b, [capi.Dynamics(**dyns[0])], 0, np.array([0, 8]), t)
    b.close()
    assert capi.STATUS_NAMES[st] == "unsupported"


def test_light_curve_api_binary_source_and_extensions(gpu):
    import lcbinint
    api = _capi_ref()
    t = np.linspace(-0.5, 0.5, 200)
    base = dict(t0=0.0, tE=1.0, u0=-0.01, alpha=0.5, s=0.95, q=1e-2, rho=5e-3)
    two = dict(base, u0_2=0.02, t0_2=0.1, flux_ratio=0.35)
    lc = lcbinint.LightCurve(source="binary", options=lcbinint.Options(coordinates="vbm", nbin=50),
                             limb_darkening=lcbinint.LimbDarkening.linear(0.5))
    got = np.asarray(lc(t, two))

    def ref_curve(u0, t0):
        p = api.params(t0=t0, tE=1.0, umin=u0, theta=0.5, sep=0.95, q=1e-2, rho=5e-3, limb_darkening_c=0.5)
        r, st = api.curve(t, p, api.options())
        assert st == 0
        return r["magnification"]
    v1, v2 = ref_curve(-0.01, 0.0), ref_curve(0.02, 0.1)
    fr = 0.35
    exp = (v1 + fr * v2) * (1.0 / (1.0 + fr))
    rel = np.abs(got / exp - 1.0)
    print(f"binary source: worst rel {rel.max():.1e}", flush=True)
    assert rel.max() <= IR_RTOL
    info = lc.info(t, two)
    assert np.abs(np.asarray(info.magnifications) / exp - 1.0).max() <= IR_RTOL
    # one epoch: magnification_from_dict divides by (1 + f)
    one = lc.magnification(0.05, two)
    tt1 = np.array([0.05])
    a1 = api.curve(tt1, api.params(t0=0.0, tE=1.0, umin=-0.01, theta=0.5, sep=0.95, q=1e-2, rho=5e-3,
                                   limb_darkening_c=0.5), api.options())[0]["magnification"][0]
    a2 = api.curve(tt1, api.params(t0=0.1, tE=1.0, umin=0.02, theta=0.5, sep=0.95, q=1e-2, rho=5e-3,
                                   limb_darkening_c=0.5), api.options())[0]["magnification"][0]
    assert abs(one / ((a1 + fr * a2) / (1.0 + fr)) - 1.0) <= IR_RTOL
    # xallarap and terrestrial parallax through the drop-in API
    ev = lcbinint.EventCoordinates(ra=270.0, dec=-30.0, tfix=10100.0)
    ev.obs_lat, ev.obs_lon = -30.2, -70.7
    lc2 = lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=40,
                                                       xallarap_param_type="angular_velocity"),
                              limb_darkening=lcbinint.LimbDarkening.linear(0.4), parallax=True,
                              terrestrial_parallax=True, event=ev)
    params = dict(t0=10100.0, tE=20.0, u0=0.01, alpha=0.7, s=1.1, q=0.1, rho=4e-3, piEN=0.3, piEE=-0.2,
                  xi_1=0.04, xi_2=-0.02, omega_xa=0.3, inc_xa=0.8, phi_xa=1.1)
    tt = np.linspace(10098.0, 10102.0, 120)
    got2 = np.asarray(lc2(tt, params))
    p = api.params(t0=10100.0, tE=20.0, umin=0.01, theta=0.7, sep=1.1, q=0.1, rho=4e-3, limb_darkening_c=0.4,
                   piEN=0.3, piEE=-0.2, ra=270.0, dec=-30.0, tfix=10100.0, obs_lat=-30.2, obs_lon=-70.7,
                   xi_1=0.04, xi_2=-0.02, omega_xa=0.3, inc_xa=0.8, phi_xa=1.1)
    o = api.options(source_bins=40, xallarap_param_type=1)
    r2, st = api.curve(tt, p, o)
    assert st == 0
    rel2 = np.abs(got2 / r2["magnification"] - 1.0)
    print(f"terrestrial + xallarap through LightCurve: worst rel {rel2.max():.1e}", flush=True)
    assert rel2.max() <= IR_RTOL


@pytest.mark.parametrize("kind,orbit_mode", [("all", 1), ("kepler", 0)])
def test_extended_gradients_match_finite_differences(gpu, kind, orbit_mode):
    # the new terms move with t0 (the xallarap phase and elapsed time) and with
    # piEN, piEE (terrestrial parallax): those derivatives of the magnification
    # against central differences (epochs whose method changes across the
    # step are left out; inverse-ray epochs through the antialiased quadrature)
    from lcbinint_amd import capi
    capi.load_earth_ephemeris(TABLE)
    curves, dyns = _curves(6, 40 + orbit_mode, kind)
    rng = np.random.default_rng(5)
    n_ep = 30
    times = np.concatenate([np.sort(c["t0"] + c["tE"] * rng.uniform(-0.12, 0.12, n_ep)) for c in curves])
    offsets = np.arange(len(curves) + 1, dtype=np.int64) * n_ep
    b = capi.Batch([capi.Curve(**c) for c in curves])

    def run(cs, ds):
        b.update_curves([capi.Curve(**c) for c in cs])
        return capi.light_curves_dynamic_grad(b, [capi.Dynamics(**d) for d in ds], orbit_mode, offsets, times)

    mag, meth, aa, dm = run(curves, dyns)
    assert np.all(np.isfinite(dm))
    names = ["t0", "u0", "tE"] + (["piEN", "piEE"] if kind == "all" else [])
    for name in names:
        k = capi.DYN_GRAD_PARAMS.index(name)
        v = np.array([c[name] if name in c else d[name] for c, d in zip(curves, dyns)])
        step = {"t0": 1e-5 * np.array([c["tE"] for c in curves])}.get(name, 1e-6 * np.abs(v))
        step = np.where(step == 0.0, 1e-8, step)
        out = []
        for sgn in (1, -1):
            cs = [dict(c) for c in curves]
            ds = [dict(d) for d in dyns]
            for i in range(len(cs)):
                (cs[i] if name in cs[i] else ds[i])[name] = v[i] + sgn * step[i]
            out.append(run(cs, ds))
        h = np.repeat(step, np.diff(offsets))
        fd = (out[0][2] - out[1][2]) / (2 * h)
        same = (out[0][1] == meth) & (out[1][1] == meth)
        scale = np.maximum(np.abs(fd), 1e-3 * np.abs(fd[same]).max() + 1e-12)
        err = np.abs(dm[:, k] - fd) / scale
        ir = meth >= 2
        for sel, bar, label in ((same & ~ir, 1e-4, "point source / hexadecapole"), (same & ir, 1e-2, "inverse ray")):
            if sel.sum() >= 5:
                print(f"{kind} d/d{name} {label}: {int(sel.sum())} epochs, median {np.median(err[sel]):.1e}",
                      flush=True)
                assert np.median(err[sel]) <= bar, (name, label, np.median(err[sel]), err[sel].max())
    b.close()
