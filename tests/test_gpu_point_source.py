"""Point-source parity on the GPU: HIP kernels vs the oracle, bitwise.

Sizes: 200k random rows; the full 4096 x 256 point-source batch (1,048,576
solves); ragged and empty curves.  The bar is bit equality of every
magnification (the point-source layer has no reordered reductions).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_binary_mag0_device_bitwise(gpu, oracle):
    import torch
    rng = np.random.default_rng(21)
    n = 200_000
    s = np.exp(rng.uniform(np.log(0.3), np.log(3), n))
    q = np.exp(rng.uniform(np.log(1e-5), 0, n))
    q = np.where(rng.uniform(size=n) < 0.5, q, 1.0 / q)
    x, y = rng.uniform(-2, 2, (2, n))
    s[:3] = [0.0, 1.0, 1.0]
    q[:3] = [0.5, 0.0, -2.0]
    ds, dq, dx, dy = (_dev(torch, v) for v in (s, q, x, y))
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    nimg = torch.empty(n, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    gpu.check(gpu.lcbg_binary_mag0_device(ds.data_ptr(), dq.data_ptr(), dx.data_ptr(),
                                          dy.data_ptr(), n, out.data_ptr(), nimg.data_ptr(),
                                          stream), "binary_mag0")
    torch.cuda.synchronize()
    a = out.cpu().numpy()
    ia = nimg.cpu().numpy()
    b, ib = oracle.binary_mag0_many(s, q, x, y)
    mism = np.nonzero(a.view(np.int64) != b.view(np.int64))[0]
    assert len(mism) == 0, (len(mism), mism[:5], a[mism[:5]], b[mism[:5]])
    assert np.array_equal(ia, ib)
    assert a[0] == 0.0 and a[1] == 0.0 and a[2] == 0.0


def test_golden_values_on_gpu(gpu):
    import torch
    cases = np.array([(1.0, 0.1, 0.2, 0.1, 5.871444912771214),
                      (0.7, 0.3, -0.4, 0.2, 2.116643550532278),
                      (1.5, 1.0, 0.05, -0.2, 1.5493462433112466),
                      (1.0, 1.0e-3, 0.3, 0.4, 2.1789388609029046)])
    cols = [_dev(torch, cases[:, k].copy()) for k in range(4)]
    out = torch.empty(4, dtype=torch.float64, device="cuda")
    gpu.check(gpu.lcbg_binary_mag0_device(*(c.data_ptr() for c in cols), 4, out.data_ptr(), None,
                                          torch.cuda.current_stream().cuda_stream), "golden")
    got = out.cpu().numpy()
    assert np.all(np.abs(got - cases[:, 4]) <= 1e-10 * cases[:, 4])


def _run_batch(gpu, curves, offsets, times):
    import torch
    from lcbinint_amd.capi import Batch, Curve
    batch = Batch([Curve(**c) for c in curves])
    d_off = _dev(torch, offsets)
    d_t = _dev(torch, times)
    d_m = torch.full((len(times),), float("nan"), dtype=torch.float64, device="cuda")
    batch.point_source(d_off.data_ptr(), d_t.data_ptr(), d_m.data_ptr(),
                       torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    batch.close()
    return d_m.cpu().numpy()


def test_point_source_batch_full_size_bitwise(gpu, oracle):
    from lcbinint_amd.workloads import point_source_batch
    curves, offsets, times = point_source_batch()
    got = _run_batch(gpu, curves, offsets, times)
    # the oracle on the same epochs: trajectory then binary_mag0(s, 1/q, y)
    exp = np.empty_like(got)
    for c, cv in enumerate(curves):
        sl = slice(offsets[c], offsets[c + 1])
        exp[sl] = oracle.light_curve_ps(times[sl], cv["t0"], cv["tE"], cv["u0"], cv["alpha"],
                                        cv["s"], cv["q"])
    mism = np.nonzero(got.view(np.int64) != exp.view(np.int64))[0]
    assert len(mism) == 0, (len(mism), mism[:5])
    assert np.all(np.isfinite(got)) and np.all(got >= 1.0 - 1e-12)


def test_ragged_and_empty_curves(gpu, oracle):
    rng = np.random.default_rng(5)
    counts = np.array([0, 1, 7, 0, 300, 65, 0, 256, 257, 1])
    curves = []
    times = []
    for k, c in enumerate(counts):
        curves.append(dict(t0=rng.uniform(-0.2, 0.2), tE=rng.uniform(0.5, 2.0),
                           u0=rng.uniform(-0.5, 0.5), alpha=rng.uniform(0, 6.28),
                           s=rng.uniform(0.5, 2.0), q=10 ** rng.uniform(-4, 0.5), rho=0.0,
                           ld_c=0.0, ld_d=0.0))
        times.append(np.sort(rng.uniform(-2, 2, c)))
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    times = np.concatenate(times)
    got = _run_batch(gpu, curves, offsets, times)
    for c, cv in enumerate(curves):
        sl = slice(offsets[c], offsets[c + 1])
        exp = oracle.light_curve_ps(times[sl], cv["t0"], cv["tE"], cv["u0"], cv["alpha"],
                                    cv["s"], cv["q"])
        assert np.array_equal(got[sl], exp)


def test_host_entry_point(gpu, oracle):
    from lcbinint_amd import capi
    cv = capi.Curve(0.0, 1.0, -0.01, 0.5, 0.95, 1e-2, 0.0, 0.0, 0.0)
    t = np.linspace(-0.5, 0.5, 200)
    out = np.empty_like(t)
    capi.check(capi.lcbg_light_curve_point_source(ctypes.byref(cv), t.ctypes.data_as(capi.c_double_p),
                                                  len(t), out.ctypes.data_as(capi.c_double_p)),
               "host")
    exp = oracle.light_curve_ps(t, 0.0, 1.0, -0.01, 0.5, 0.95, 1e-2)
    assert np.array_equal(out, exp)
