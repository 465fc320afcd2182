"""Synthetic light-curve workloads with fixed seeds.

point_source_batch: 4096 parameter sets, s log-uniform [0.3, 3], q log-uniform
[1e-5, 1], u0 uniform [-0.5, 0.5], alpha uniform [0, 2 pi), tE = 1, t0 = 0;
256 sorted epochs per set uniform in [-2, 2]; one quintic solve per epoch
(1,048,576 solves).

finite_source_batch: 2048 sets with the geometry above plus rho log-uniform
[1e-4, 3e-2] and linear limb darkening u uniform [0.1, 0.9]; per-curve epoch
count floor(log-uniform [50, 2001)), epochs uniform in [-1, 1]; nbin uniform
integer in [25, 100]; automatic Cartesian/polar choice.

differentiation_batch: 256 sets with the point-source geometry plus rho
log-uniform [1e-4, 3e-2] and linear limb darkening u uniform [0.1, 0.9]; 1000
sorted epochs per set uniform in [-2, 2]; observed fluxes are made from the
model itself (F = fs A + fb, fs = 1, fb = 0.1, Gaussian noise of 1% of F).

dynamic_batch: 1024 curves with annual parallax and circular orbital motion;
epochs floor(U{500..3000}) per curve, uniform over 200 days (reduced JD
10000-10200, the Earth table of tests/golden/earth_ephemeris_substitute.txt
covers it); t0 uniform [10050, 10150], tE log-uniform [10, 100] days, the
point-source geometry above, rho log-uniform [1e-4, 1e-2], linear LD u uniform
[0.1, 0.9]; piEN, piEE uniform [-0.5, 0.5]; ds/dt, dalpha/dt uniform [-1, 1]
per year, as VBM rates g1 = (ds/dt) / s / 365.25, g2 = (dalpha/dt) / 365.25,
g3 = 0; event at ra 270, dec -30 degrees, tfix = t0; nbin 50.

plain_curve: the single light curve t0=0, tE=1, u0=-0.01, alpha=0.5, s=0.95,
q=1e-2, rho=5e-3, linear LD u=0.5, 200 epochs linspace(-0.5, 0.5), nbin 50.
"""
from __future__ import annotations

import numpy as np


def point_source_batch(n_curves: int = 4096, n_epochs: int = 256, seed: int = 2):
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(0.3), np.log(3.0), n_curves))
    q = np.exp(rng.uniform(np.log(1e-5), 0.0, n_curves))
    u0 = rng.uniform(-0.5, 0.5, n_curves)
    alpha = rng.uniform(0.0, 2.0 * np.pi, n_curves)
    times = np.sort(rng.uniform(-2.0, 2.0, (n_curves, n_epochs)), axis=1)
    curves = [dict(t0=0.0, tE=1.0, u0=float(u0[i]), alpha=float(alpha[i]), s=float(s[i]),
                   q=float(q[i]), rho=0.0, ld_c=0.0, ld_d=0.0) for i in range(n_curves)]
    offsets = np.arange(n_curves + 1, dtype=np.int64) * n_epochs
    return curves, offsets, times.ravel()


def finite_source_batch(n_curves: int = 2048, seed: int = 3):
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(0.3), np.log(3.0), n_curves))
    q = np.exp(rng.uniform(np.log(1e-5), 0.0, n_curves))
    u0 = rng.uniform(-0.5, 0.5, n_curves)
    alpha = rng.uniform(0.0, 2.0 * np.pi, n_curves)
    rho = np.exp(rng.uniform(np.log(1e-4), np.log(3e-2), n_curves))
    ld = rng.uniform(0.1, 0.9, n_curves)
    counts = np.floor(np.exp(rng.uniform(np.log(50), np.log(2001), n_curves))).astype(np.int64)
    nbin = rng.integers(25, 101, n_curves)
    times = [np.sort(rng.uniform(-1.0, 1.0, c)) for c in counts]
    curves = [dict(t0=0.0, tE=1.0, u0=float(u0[i]), alpha=float(alpha[i]), s=float(s[i]),
                   q=float(q[i]), rho=float(rho[i]), ld_c=float(ld[i]), ld_d=0.0)
              for i in range(n_curves)]
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return curves, offsets, np.concatenate(times), nbin


def plain_curve():
    curve = dict(t0=0.0, tE=1.0, u0=-0.01, alpha=0.5, s=0.95, q=1e-2, rho=5e-3, ld_c=0.5,
                 ld_d=0.0)
    return curve, np.linspace(-0.5, 0.5, 200)


def differentiation_batch(n_curves: int = 256, n_epochs: int = 1000, seed: int = 4):
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(0.3), np.log(3.0), n_curves))
    q = np.exp(rng.uniform(np.log(1e-5), 0.0, n_curves))
    u0 = rng.uniform(-0.5, 0.5, n_curves)
    alpha = rng.uniform(0.0, 2.0 * np.pi, n_curves)
    rho = np.exp(rng.uniform(np.log(1e-4), np.log(3e-2), n_curves))
    ld = rng.uniform(0.1, 0.9, n_curves)
    times = np.sort(rng.uniform(-2.0, 2.0, (n_curves, n_epochs)), axis=1)
    curves = [dict(t0=0.0, tE=1.0, u0=float(u0[i]), alpha=float(alpha[i]), s=float(s[i]),
                   q=float(q[i]), rho=float(rho[i]), ld_c=float(ld[i]), ld_d=0.0)
              for i in range(n_curves)]
    offsets = np.arange(n_curves + 1, dtype=np.int64) * n_epochs
    return curves, offsets, times.ravel()


def synthetic_fluxes(mag, fs: float = 1.0, fb: float = 0.1, noise: float = 0.01, seed: int = 5):
    """F = fs A + fb with Gaussian noise of `noise` times F; returns (flux, sigma)."""
    rng = np.random.default_rng(seed)
    f = fs * np.asarray(mag) + fb
    sigma = noise * f
    return f + sigma * rng.standard_normal(len(f)), sigma


def dynamic_batch(n_curves: int = 1024, seed: int = 5, min_epochs: int = 500, max_epochs: int = 3000):
    """Returns curves (dicts of the lcbg_curve fields), dynamics (dicts of the
    lcbg_dynamics fields), offsets, times; orbital motion is circular."""
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(0.3), np.log(3.0), n_curves))
    q = np.exp(rng.uniform(np.log(1e-5), 0.0, n_curves))
    u0 = rng.uniform(-0.5, 0.5, n_curves)
    alpha = rng.uniform(0.0, 2.0 * np.pi, n_curves)
    rho = np.exp(rng.uniform(np.log(1e-4), np.log(1e-2), n_curves))
    ld = rng.uniform(0.1, 0.9, n_curves)
    t0 = rng.uniform(10050.0, 10150.0, n_curves)
    tE = np.exp(rng.uniform(np.log(10.0), np.log(100.0), n_curves))
    pin = rng.uniform(-0.5, 0.5, n_curves)
    pie = rng.uniform(-0.5, 0.5, n_curves)
    dsdt = rng.uniform(-1.0, 1.0, n_curves)
    dadt = rng.uniform(-1.0, 1.0, n_curves)
    counts = rng.integers(min_epochs, max_epochs + 1, n_curves)
    times = [np.sort(rng.uniform(10000.0, 10200.0, c)) for c in counts]
    curves = [dict(t0=float(t0[i]), tE=float(tE[i]), u0=float(u0[i]), alpha=float(alpha[i]), s=float(s[i]),
                   q=float(q[i]), rho=float(rho[i]), ld_c=float(ld[i]), ld_d=0.0) for i in range(n_curves)]
    dyn = [dict(piEN=float(pin[i]), piEE=float(pie[i]), ra=270.0, dec=-30.0, tfix=float(t0[i]),
                g1=float(dsdt[i] / s[i] / 365.25), g2=float(dadt[i] / 365.25), g3=0.0, lom_szs=0.0,
                lom_ar=1.0) for i in range(n_curves)]
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return curves, dyn, offsets, np.concatenate(times)
