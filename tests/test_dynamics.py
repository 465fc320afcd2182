"""Annual parallax and orbital motion: host trajectories against the reference
core compiled from its sources (oracle/_ref), and invariants that hold for any
Earth table (the reference's own table is not distributed; both sides read the
substitute table of tests/golden/earth_ephemeris_substitute.txt)."""
import os
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
TABLE = ROOT / "tests" / "golden" / "earth_ephemeris_substitute.txt"
REF = ROOT / "oracle" / "_ref" / "liblcbinint_ref.so"
os.environ.setdefault("LCBININT_EARTH_EPHEMERIS", str(TABLE))

DYN_KEYS = ("piEN", "piEE", "ra", "dec", "tfix", "g1", "g2", "g3", "lom_szs", "lom_ar")


@pytest.fixture(scope="module")
def capi():
    from lcbinint_amd import capi as c
    c.load_earth_ephemeris(TABLE)
    return c


def _case(rng, mode, parallax, tfix_shift):
    t0 = 10100.0 + rng.uniform(-20, 20)
    return dict(t0=t0, tE=rng.uniform(10, 80), u0=rng.uniform(-0.5, 0.5), alpha=rng.uniform(0, 6.28),
                s=float(np.exp(rng.uniform(np.log(0.3), np.log(3)))), q=float(np.exp(rng.uniform(np.log(1e-5), 0))),
                rho=0.0, piEN=rng.uniform(-0.5, 0.5) if parallax else 0.0,
                piEE=rng.uniform(-0.5, 0.5) if parallax else 0.0, ra=270.0, dec=-30.0,
                tfix=0.0 if tfix_shift is None else t0 + tfix_shift,
                g1=rng.uniform(-1, 1) / 365.25, g2=rng.uniform(-1, 1) / 365.25,
                g3=rng.uniform(-1, 1) / 365.25 if mode == 2 else 0.0, lom_szs=rng.uniform(-1, 1),
                lom_ar=rng.uniform(1.0, 3.0))


def _mine(capi, p, mode, times):
    cv = capi.Curve(t0=p["t0"], tE=p["tE"], u0=p["u0"], alpha=p["alpha"], s=p["s"], q=p["q"], rho=p["rho"],
                    ld_c=0.0, ld_d=0.0)
    return capi.dynamic_sources(cv, capi.Dynamics(**{k: p[k] for k in DYN_KEYS}), mode, times)


@pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("parallax", [False, True])
def test_trajectory_bitwise_vs_reference(capi, mode, parallax):
    if mode == 0 and not parallax:
        pytest.skip("static trajectory: covered by the static path")
    from oracle.bindings import default_settings, load_ref
    ref = load_ref(str(REF))
    rng = np.random.default_rng(10 * mode + parallax)
    for k in range(12):
        p = _case(rng, mode, parallax, None if k % 2 else 3.0)
        times = np.sort(rng.uniform(10000, 10200, 200))
        _, _, src, status = ref.light_curve_dynamic(default_settings(), times, p, mode)
        assert status == 0
        mine = _mine(capi, p, mode, times)
        assert np.array_equal(mine.view(np.int64), src.view(np.int64)), (mode, parallax, k)


def test_parallax_vanishes_at_reference_time(capi):
    # the offsets are relative to the Earth's position and velocity at tfix
    # (trajectory.cpp:245-257): exactly zero there.  Near tfix they grow only
    # with the difference between the table's velocity and the slope of its
    # linear interpolation (~1e-4 AU/day), linearly within one table interval
    rng = np.random.default_rng(3)
    p = _case(rng, 0, True, 5.0)
    tf = p["tfix"]
    base = dict(p, piEN=0.0, piEE=0.0)
    times = np.array([tf, tf + 0.01, tf + 0.02])
    off = _mine(capi, p, 0, times)[:, :2] - _mine(capi, base, 0, times)[:, :2]
    assert np.all(off[0] == 0.0)
    assert np.all(np.abs(off[1:]) < 1e-5)
    r = np.linalg.norm(off[2]) / np.linalg.norm(off[1])
    assert 1.9 < r < 2.1
    # a year-scale offset is of order pi_E: one AU in the plane of the sky
    far = _mine(capi, p, 0, np.array([tf + 90.0]))[0, :2] - _mine(capi, base, 0, np.array([tf + 90.0]))[0, :2]
    assert 0.05 * np.hypot(p["piEN"], p["piEE"]) < np.hypot(*far) < 3.0 * np.hypot(p["piEN"], p["piEE"])


def test_orbit_reference_state(capi):
    # at the reference time the orbit returns the input separation and angle
    rng = np.random.default_rng(4)
    for mode in (1, 2):
        for _ in range(10):
            p = _case(rng, mode, False, None)
            out = _mine(capi, p, mode, np.array([p["t0"]]))
            assert abs(out[0, 2] - p["s"]) <= 1e-12 * p["s"]
            # the source rotation by alpha(t0) - alpha is then the identity
            tau = 0.0
            x = tau * np.cos(p["alpha"]) - p["u0"] * np.sin(p["alpha"])
            y = tau * np.sin(p["alpha"]) + p["u0"] * np.cos(p["alpha"])
            assert abs(out[0, 0] - x) < 1e-9 and abs(out[0, 1] - y) < 1e-9


def test_circular_orbit_period(capi):
    # a circular orbit repeats after 2 pi / w_orb; with g1 = g3 = 0 the
    # separation is constant and the angle advances at g2
    rng = np.random.default_rng(5)
    p = _case(rng, 1, False, None)
    p.update(g1=0.0, g3=0.0, g2=0.01)
    t = p["t0"] + np.array([0.0, 10.0])
    out = _mine(capi, p, 1, t)
    assert abs(out[1, 2] - p["s"]) < 1e-12
    # source radius in the rotating frame is unchanged
    assert abs(np.hypot(*out[1, :2]) - np.hypot(((t[1] - p["t0"]) / p["tE"]), p["u0"])) < 1e-12


def test_ephemeris_table_errors(capi):
    assert capi.lcbg_load_earth_ephemeris(b"/nonexistent/earth.txt") != 0
    capi.load_earth_ephemeris(TABLE)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_trajectory_jacobian_vs_finite_differences(capi, mode):
    # forward-mode derivatives of (x, y, s) against central differences of the
    # same host trajectory (step 1e-4 relative: the nearly edge-on circular
    # orbits of g3 = 0 lose digits to cancellation, which smaller steps expose);
    # errors are relative to each component's largest derivative over the epochs
    rng = np.random.default_rng(20 + mode)
    for k in range(4):
        p = _case(rng, mode, True, 3.0 if k % 2 else None)
        times = np.sort(rng.uniform(10000, 10200, 40))
        cv = capi.Curve(t0=p["t0"], tE=p["tE"], u0=p["u0"], alpha=p["alpha"], s=p["s"], q=p["q"],
                        rho=0.0, ld_c=0.0, ld_d=0.0)
        dy = capi.Dynamics(**{kk: p[kk] for kk in DYN_KEYS})
        base, jac = capi.dynamic_sources_jacobian(cv, dy, mode, times)
        assert np.array_equal(base.view(np.int64), capi.dynamic_sources(cv, dy, mode, times).view(np.int64))
        for j, name in enumerate(capi.TRAJ_PARAMS):
            v = p[name]
            h = 1e-5 * p["tE"] if name == "t0" else (1e-4 * abs(v) if v != 0.0 else 1e-8)
            if mode == 0 and name in ("g1", "g2"):
                assert np.all(jac[:, :, j] == 0.0)
                continue
            outs = []
            for sgn in (1, -1):
                q = dict(p)
                q[name] = v + sgn * h
                cq = capi.Curve(t0=q["t0"], tE=q["tE"], u0=q["u0"], alpha=q["alpha"], s=q["s"], q=q["q"],
                                rho=0.0, ld_c=0.0, ld_d=0.0)
                dq = capi.Dynamics(**{kk: q[kk] for kk in DYN_KEYS})
                outs.append(capi.dynamic_sources(cq, dq, mode, times))
            fd = (outs[0] - outs[1]) / (2 * h)
            # (a floor of 1e-9 for derivatives that vanish identically, such as
            # those of the Keplerian frame angle with respect to s)
            scale = np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max(axis=0, keepdims=True) + 1e-9)
            err = np.abs(jac[:, :, j] - fd) / scale
            assert np.median(err) < 1e-5 and err.max() < 1e-2, (mode, name, np.median(err), err.max())
