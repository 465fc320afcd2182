"""Geometry helpers of LightCurve on the GPU against the reference compiled
from source (oracle/_ref):

* caustics / critical_curves (geometry_branches,
  python/lcbinint_pybind.cpp:1882-1926): the points of every phase come from
  the reference's own critical-curve solve (ref_critical_points_at_phase) or
  caustic map (ref_caustic_points_at_phase); the branch tracking is restated
  here (py_append_tracked_points, :621-658: first phase ordered by atan2, then
  the permutation of least total squared step, the first one in
  std::next_permutation order).  Bitwise, including the separation of an
  orbiting lens at a given time;
* source_trajectory (source_for_geometry, :1840-1880): against the
  reference's Trajectory and orbital_state composed as the binding composes
  them (ref_source_for_geometry): bitwise without orbital motion, 1e-13 with
  it (the device evaluates the orbit's cos, sin and atan2 per epoch).
"""
import itertools
import math
import os
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
TABLE = ROOT / "tests" / "golden" / "earth_ephemeris_substitute.txt"
REF = ROOT / "oracle" / "_ref" / "liblcbinint_ref.so"
os.environ.setdefault("LCBININT_EARTH_EPHEMERIS", str(TABLE))

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not REF.exists(), reason="oracle/_ref not built")]

PERMS = list(itertools.permutations(range(4)))  # std::next_permutation order from the identity


def _track(per_phase):
    """Four branches from the points of each phase (phases whose solve did not
    give four points are skipped, as the reference skips them)."""
    branches = [[] for _ in range(4)]
    for pts in per_phase:
        if len(pts) != 4:
            continue
        if not branches[0]:
            order = sorted(range(4), key=lambda i: math.atan2(pts[i][1], pts[i][0]))
            for b in range(4):
                branches[b].append((float(pts[order[b]][0]), float(pts[order[b]][1])))
            continue
        best, best_cost = PERMS[0], math.inf
        for p in PERMS:
            cost = 0.0
            for b in range(4):
                bx, by = branches[b][-1]
                dx = bx - float(pts[p[b]][0])
                dy = by - float(pts[p[b]][1])
                cost += dx * dx + dy * dy
            if cost < best_cost:
                best, best_cost = p, cost
        for b in range(4):
            branches[b].append((float(pts[best[b]][0]), float(pts[best[b]][1])))
    return np.array([[p[0] for p in br] for br in branches]), np.array([[p[1] for p in br] for br in branches])


def _expected(ref, s, q_eff, n, caustic):
    pts = []
    for i in range(n):
        phase = 2.0 * math.pi * i / n
        pts.append(ref.caustic_points(s, q_eff, phase) if caustic else ref.critical_points(s, q_eff, phase))
    return _track(pts)


def _same(got, exp, label):
    got = np.asarray(got, dtype=np.float64)
    assert got.shape == exp.shape, (label, got.shape, exp.shape)
    assert np.array_equal(got.view(np.int64), exp.view(np.int64)), (label, np.abs(got - exp).max())


@pytest.mark.parametrize("s,q", [(0.95, 1e-2), (1.3, 0.2), (0.6, 1e-3), (2.5, 0.5)])
def test_caustics_and_critical_curves_match_reference(gpu, s, q):
    import lcbinint
    from oracle.bindings import load_ref
    ref = load_ref(str(REF))
    lc = lcbinint.LightCurve()
    n = 160
    print(f"geometry s={s} q={q}, {n} phases", flush=True)
    for caustic in (True, False):
        ex, ey = _expected(ref, s, 1.0 / q, n, caustic)
        got = lc.caustics(s=s, q=q, n_points=n) if caustic else lc.critical_curves(s=s, q=q, n_points=n)
        _same(got.x, ex, ("x", caustic))
        _same(got.y, ey, ("y", caustic))
    # the dict form: the lens at its reference separation
    got = lc.caustics({"s": s, "q": q}, n)
    ex, ey = _expected(ref, s, 1.0 / q, n, True)
    _same(got.x, ex, "dict x")
    _same(got.y, ey, "dict y")
    with pytest.raises(ValueError):
        lc.caustics(s=s, q=q, n_points=7)


def test_caustics_of_an_orbiting_lens(gpu):
    """With orbital motion and a finite time, the caustic is the one of the
    separation at that time (the reference's orbit state)."""
    import lcbinint
    from oracle.bindings import default_settings, load_ref
    ref = load_ref(str(REF))
    t0 = 10100.0
    p = dict(t0=t0, tE=30.0, u0=0.05, alpha=0.7, s=1.1, q=0.02, g1=0.3 / 1.1 / 365.25, g2=-0.5 / 365.25,
             ra=270.0, dec=-30.0, tfix=t0)
    lc = lcbinint.LightCurve(event=lcbinint.EventCoordinates(ra=270.0, dec=-30.0, tfix=t0),
                             orbital_motion_mode=lcbinint.OrbitalMotionMode.CIRCULAR)
    for t in (t0 - 40.0, t0 + 25.0):
        _, _, src, st = ref.light_curve_dynamic(default_settings(), [t], dict(p, rho=0.0), 1)
        assert st == 0
        s_t = float(src[0, 2])
        assert s_t != p["s"]
        ex, ey = _expected(ref, s_t, 1.0 / p["q"], 120, True)
        got = lc.caustics(s=p["s"], q=p["q"], n_points=120, time=t, t0=t0, tE=p["tE"], u0=p["u0"],
                          alpha=p["alpha"], g1=p["g1"], g2=p["g2"])
        _same(got.x, ex, ("orbit x", t))
        _same(got.y, ey, ("orbit y", t))
    # without a time the reference separation is used
    got = lc.caustics(s=p["s"], q=p["q"], n_points=120)
    ex, _ = _expected(ref, p["s"], 1.0 / p["q"], 120, True)
    _same(got.x, ex, "no time")


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("parallax", [False, True])
def test_source_trajectory_matches_reference(gpu, mode, parallax):
    import lcbinint
    from lcbinint_amd import capi
    from oracle.bindings import load_ref
    capi.load_earth_ephemeris(TABLE)
    ref = load_ref(str(REF))
    t0 = 10100.0
    p = dict(t0=t0, tE=30.0, u0=0.05, alpha=0.7, s=1.1, q=0.02, g1=0.3 / 1.1 / 365.25, g2=-0.5 / 365.25,
             g3=0.2 / 365.25 if mode == 2 else 0.0, lom_szs=0.3 if mode == 2 else 0.0,
             lom_ar=1.5 if mode == 2 else 1.0)
    if parallax:
        p.update(piEN=0.2, piEE=-0.1)
    times = np.linspace(t0 - 60.0, t0 + 60.0, 301)
    modes = [lcbinint.OrbitalMotionMode.STATIC, lcbinint.OrbitalMotionMode.CIRCULAR,
             lcbinint.OrbitalMotionMode.KEPLER]
    lc = lcbinint.LightCurve(event=lcbinint.EventCoordinates(ra=270.0, dec=-30.0, tfix=t0),
                             orbital_motion_mode=modes[mode], parallax=parallax)
    print(f"source trajectory, orbit mode {mode}, parallax {parallax}", flush=True)
    got = lc.source_trajectory(times, p)
    assert np.array_equal(np.asarray(got.times), times)
    xy = np.stack([np.asarray(got.x), np.asarray(got.y)], axis=1)
    exp = ref.source_for_geometry(times, dict(p, ra=270.0, dec=-30.0, tfix=t0), mode)
    assert np.isfinite(exp).all()
    if mode == 0:
        assert np.array_equal(xy.view(np.int64), exp.view(np.int64))
    else:
        err = np.abs(xy - exp) / np.maximum(np.abs(exp), 1.0)
        assert err.max() <= 1e-13, err.max()
    # the keyword forms (piEN / piEE only with parallax=True)
    kw = {k: v for k, v in p.items() if k not in ("piEN", "piEE")}
    if parallax:
        got2 = lc.source_trajectory(times, piEN=p["piEN"], piEE=p["piEE"], **kw)
        assert np.array_equal(np.asarray(got2.x), np.asarray(got.x))
    else:
        got2 = lc.source_trajectory(times, **kw)
        assert np.array_equal(np.asarray(got2.x), np.asarray(got.x))
        with pytest.raises(ValueError):
            lc.source_trajectory(times, piEN=0.1, **kw)
