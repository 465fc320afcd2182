"""The oracle's critical-curve and hexadecapole restatements pinned bitwise
against the reference compiled from source, and the limb-darkening table
against its closed form."""
import math

import numpy as np
import pytest

from oracle.bindings import default_settings


def test_critical_curve_bitwise_vs_reference(oracle, ref):
    rng = np.random.default_rng(41)
    for _ in range(300):
        s = float(np.exp(rng.uniform(np.log(0.3), np.log(3))))
        q = float(np.exp(rng.uniform(np.log(1e-5), np.log(1e5))))
        phase = 2.0 * math.pi * int(rng.integers(0, 1400)) / 1400.0
        ca, cr = oracle.critical_curve(s, q, math.cos(phase), math.sin(phase))
        assert np.array_equal(ca, ref.caustic_points(s, q, phase))
        assert np.array_equal(cr, ref.critical_sources(s, q, phase))


def test_caustic_points_satisfy_critical_condition(oracle):
    # mapped critical points lie on the caustic: the image of a critical point
    # has |J| = 0, so the Jacobian of the lens map at the root vanishes
    s, q = 1.1, 0.01
    for k in range(0, 1400, 97):
        phase = 2.0 * math.pi * k / 1400.0
        ca, cr = oracle.critical_curve(s, q, math.cos(phase), math.sin(phase))
        assert np.allclose(ca, cr, atol=1e-12)  # both maps agree on the source plane


def test_limb_darkening_table_closed_form(oracle):
    for c in (0.0, 0.3, 0.9):
        t = oracle.ld_table(c)
        r2 = np.arange(20001) / 20000.0
        exact = 1.0 - c * (1.0 - np.sqrt(1.0 - r2))
        assert np.max(np.abs(t - exact)) < 1e-15
        # nearest-node lookup (finite_source_magnifier.cpp:4001-4012)
        assert oracle.ld_lookup(t, 0.123456) == t[int(0.123456 * 20000.0 + 0.5)]
        assert oracle.ld_lookup(t, -1.0) == t[0] and oracle.ld_lookup(t, 2.0) == t[20000]


def test_hexadecapole_bitwise_and_limit(oracle, ref):
    rng = np.random.default_rng(43)
    for _ in range(200):
        s = float(np.exp(rng.uniform(np.log(0.3), np.log(3))))
        q = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e4))))
        x, y = rng.uniform(-1, 1, 2)
        rho = float(10 ** rng.uniform(-4, -1.5))
        c = float(rng.uniform(0, 0.9))
        a0, _ = oracle.binary_mag0(s, q, x, y)
        m, e = oracle.hexadecapole(s, q, x, y, rho, c, 0.0, a0)
        mr, er = ref.hexadecapole(default_settings(limb_darkening_c=c), s, q, x, y, rho, a0)
        assert (m, e) == (mr, er)
    # rho -> 0: the hexadecapole magnification tends to the point-source one
    a0, _ = oracle.binary_mag0(1.2, 0.2, 0.7, 0.4)
    m, _ = oracle.hexadecapole(1.2, 0.2, 0.7, 0.4, 1e-6, 0.5, 0.0, a0)
    assert abs(m - a0) < 1e-9 * a0
