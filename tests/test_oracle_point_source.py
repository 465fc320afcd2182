"""The point-source oracle pinned before it is trusted.

Pins: literal golden values of the reference's own tests, the single-lens
closed form in the q -> 0 limit, polynomial residuals, and bitwise equality
with the reference compiled from source (oracle/_ref) on randomized inputs.
"""
import math

import numpy as np
import pytest

# tests/regression/test_vbm_consistency.py:8-13 (VBMicrolensing BinaryMag0 values, 1e-10)
BINARY_POINT_CASES = [
    (1.0, 0.1, 0.2, 0.1, 5.871444912771214),
    (0.7, 0.3, -0.4, 0.2, 2.116643550532278),
    (1.5, 1.0, 0.05, -0.2, 1.5493462433112466),
    (1.0, 1.0e-3, 0.3, 0.4, 2.1789388609029046),
]


@pytest.mark.parametrize("s,q,x,y,expected", BINARY_POINT_CASES)
def test_point_source_golden_values(oracle, s, q, x, y, expected):
    a, _ = oracle.binary_mag0(s, q, x, y)
    assert abs(a - expected) <= 1e-10 * abs(expected)


def test_unit_core_legacy_cases(oracle):
    # tests/unit/test_core.cpp:73-94: lcbi_magnification at t=0.2 with umin=0.1,
    # theta=0, legacy coordinates (x = tau, y = umin), effective q = q.
    a, _ = oracle.binary_mag0(1.0, 0.1, 0.2, 0.1)
    assert abs(a - 5.871444912771214) <= 1e-10
    # sep=1.5, q=1: the legacy wide-binary offset m2*s - m2/s (lens_model.cpp:30-41)
    m2 = 1.0 / (1.0 + 1.0)
    offset = m2 * 1.5 - m2 / 1.5
    a, _ = oracle.binary_mag0(1.5, 1.0, 0.2 - offset, 0.1)
    assert abs(a - 3.5659775904852786) <= 1e-10


@pytest.mark.parametrize("u", [0.05, 0.1, 0.3, 0.7, 1.2])
def test_single_lens_limit(oracle, u):
    # q -> 0 with the planet far away: A -> (u^2+2)/(u sqrt(u^2+4)); the input
    # coordinates are centred on the centre of mass, i.e. on the primary.
    a, n = oracle.binary_mag0(5.0, 1e-12, u * math.cos(0.3), u * math.sin(0.3))
    exact = (u * u + 2.0) / (u * math.sqrt(u * u + 4.0))
    assert abs(a - exact) <= 1e-9 * exact
    assert n == 3


def test_zero_separation_and_bad_mass_ratio(oracle):
    assert oracle.binary_mag0(0.0, 0.1, 0.1, 0.1)[0] == 0.0
    assert oracle.binary_mag0(1.0, 0.0, 0.1, 0.1)[0] == 0.0
    assert oracle.binary_mag0(1.0, -1.0, 0.1, 0.1)[0] == 0.0


# tests/regression/test_solver_vbm_consistency.py:4-33
ROOT_CASES = [
    [complex(-1.0, 0.0), 0j, 0j, complex(1.0, 0.0)],
    [complex(-1.0, 0.25), complex(0.5, -0.75), complex(-1.0, 0.0), complex(0.25, 0.5),
     complex(-0.25, 0.0), complex(1.0, 0.0)],
    [complex(0.2, -0.1), complex(-0.4, 0.3), complex(0.1, 0.2), complex(-0.7, 0.0),
     complex(0.4, -0.6), complex(-0.2, 0.1), complex(0.3, 0.0), complex(1.0, 0.0)],
]


@pytest.mark.parametrize("coeffs", ROOT_CASES)
def test_roots_residual_and_numpy(oracle, coeffs):
    roots = oracle.roots_gen(coeffs)
    poly = np.polynomial.Polynomial(coeffs)
    assert np.all(np.abs(poly(roots)) < 1e-9)
    expected = np.roots(list(reversed(coeffs)))
    for r in expected:
        assert np.min(np.abs(roots - r)) < 1e-9


def test_cubic_roots_of_unity_exact(oracle):
    roots = oracle.roots_gen(ROOT_CASES[0])
    expected = [np.exp(2j * np.pi * k / 3) for k in range(3)]
    for e in expected:
        assert np.min(np.abs(roots - e)) < 1e-14


def test_quintic_images_satisfy_lens_equation(oracle):
    rng = np.random.default_rng(7)
    for _ in range(200):
        s = float(np.exp(rng.uniform(np.log(0.3), np.log(3))))
        q = float(np.exp(rng.uniform(np.log(1e-5), 0)))
        x, y = rng.uniform(-1, 1, 2)
        cands = oracle.image_candidates(s, q, x, y)
        phys = cands[cands[:, 4] == 1.0]
        assert len(phys) in (3, 5)
        assert np.all(phys[:, 3] < 1e-8)
        # magnification identity: sum of 1/|J| over physical images
        a, n = oracle.binary_mag0(s, q, x, y)
        assert n == len(phys)
        assert abs(a - np.sum(1.0 / np.abs(phys[:, 2]))) <= 1e-12 * a


def test_bitwise_vs_reference_point_source(oracle, ref):
    rng = np.random.default_rng(11)
    n = 50000
    s = np.exp(rng.uniform(np.log(0.3), np.log(3), n))
    q = np.exp(rng.uniform(np.log(1e-5), 0, n))
    q = np.where(rng.uniform(size=n) < 0.5, q, 1 / q)
    x, y = rng.uniform(-2, 2, (2, n))
    a, ia = oracle.binary_mag0_many(s, q, x, y)
    b, ib = ref.binary_mag0_many(s, q, x, y)
    assert np.array_equal(a.view(np.int64), b.view(np.int64))
    assert np.array_equal(ia, ib)


def test_bitwise_vs_reference_derivative_indicator(oracle, ref):
    rng = np.random.default_rng(12)
    for _ in range(2000):
        s = float(np.exp(rng.uniform(np.log(0.3), np.log(3))))
        q = float(np.exp(rng.uniform(np.log(1e-5), 0)))
        x, y = rng.uniform(-1.5, 1.5, 2)
        assert oracle.binary_mag0_with_derivatives(s, q, x, y) == \
            ref.binary_mag0_with_derivatives(s, q, x, y)


def test_bitwise_vs_reference_candidates_and_roots(oracle, ref):
    rng = np.random.default_rng(13)
    for _ in range(2000):
        s = float(np.exp(rng.uniform(np.log(0.3), np.log(3))))
        q = float(np.exp(rng.uniform(np.log(1e-5), 0)))
        x, y = rng.uniform(-1.5, 1.5, 2)
        assert np.array_equal(oracle.image_candidates(s, q, x, y), ref.image_candidates(s, q, x, y))
    for _ in range(500):
        deg = int(rng.integers(3, 8))
        c = rng.normal(size=deg + 1) + 1j * rng.normal(size=deg + 1)
        assert np.array_equal(oracle.roots_gen(c), ref.roots_gen(c))


def test_light_curve_point_source_matches_reference_loop(oracle, ref):
    from oracle.bindings import default_settings
    times = np.linspace(-1.5, 1.5, 301)
    a = oracle.light_curve_ps(times, 0.1, 0.8, 0.05, 2.1, 1.1, 0.03)
    b, meth, st = ref.light_curve(default_settings(), times, 0.1, 0.8, 0.05, 2.1, 1.1, 0.03, 0.0)
    assert st == 0
    assert np.array_equal(a, b)
