"""The allocation-free entry point against the batch entry point it mirrors.

lcbg_light_curve_device_noalloc keeps all of its workspace in one static
device array; its results must equal those of lcbg_batch_finite_source_device
on a one-curve batch of the same curve: methods equal, point-source and
hexadecapole magnifications bitwise, inverse-ray magnifications within 1e-9
relative (measured: bitwise; the schedule changes no result).  Two calls in a
row, the second with fewer epochs, check that the array's contents left by a
call do not leak into the next.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

IR_RTOL = 1e-9


def _batch_run(gpu, cv, times, opt):
    import torch
    n = len(times)
    batch = gpu.Batch([gpu.Curve(**cv)], opt)
    d_off = torch.tensor([0, n], dtype=torch.int64, device="cuda")
    d_t = torch.from_numpy(np.ascontiguousarray(times)).cuda()
    d_m = torch.empty(n, dtype=torch.float64, device="cuda")
    d_k = torch.empty(n, dtype=torch.int32, device="cuda")
    gpu.finite_source(batch, d_off.data_ptr(), d_t.data_ptr(), n, d_m.data_ptr(), d_k.data_ptr(),
                      torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    batch.close()
    return d_m.cpu().numpy(), d_k.cpu().numpy()


def _noalloc_run(gpu, cv, times, opt):
    import torch
    n = len(times)
    d_t = torch.from_numpy(np.ascontiguousarray(times)).cuda()
    d_m = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
    d_k = torch.full((n,), -9, dtype=torch.int32, device="cuda")
    rec = gpu.Curve(**cv)
    st = gpu.lcbg_light_curve_device_noalloc(ctypes.byref(rec), ctypes.byref(opt), d_t.data_ptr(), n,
                                             d_m.data_ptr(), d_k.data_ptr(),
                                             torch.cuda.current_stream().cuda_stream)
    gpu.check(st, "lcbg_light_curve_device_noalloc")
    torch.cuda.synchronize()
    return d_m.cpu().numpy(), d_k.cpu().numpy()


def _same(got, gk, exp, ek, label):
    assert np.array_equal(gk, ek), (label, np.nonzero(gk != ek)[0][:10])
    ps_hex = ek <= 1
    assert np.array_equal(got[ps_hex].view(np.int64), exp[ps_hex].view(np.int64)), label
    ir = ~ps_hex
    rel = np.abs(got[ir] - exp[ir]) / np.abs(exp[ir])
    assert rel.size == 0 or rel.max() <= IR_RTOL, (label, rel.max())
    bitwise = bool(np.array_equal(got.view(np.int64), exp.view(np.int64)))
    print(f"{label}: {len(got)} epochs, methods {np.bincount(ek, minlength=4).tolist()}, "
          f"inverse-ray worst rel {float(rel.max()) if rel.size else 0.0:.2e}, bitwise {bitwise}")


def _heavy_curve():
    """A curve of the batched workload with many inverse-ray epochs at nbin >= 80."""
    from lcbinint_amd.workloads import finite_source_batch
    curves, offsets, times, nbin = finite_source_batch()
    counts = np.diff(offsets)
    # close approaches (|u0| small) at high resolution and many epochs
    order = np.argsort([abs(c["u0"]) for c in curves])
    for c in order:
        if counts[c] >= 1000 and nbin[c] >= 80 and curves[c]["rho"] >= 1e-3:
            return curves[c], times[offsets[c]:offsets[c + 1]], int(nbin[c])
    raise AssertionError("no heavy curve in the workload")


@pytest.mark.parametrize("mode", [4, 1])
def test_noalloc_matches_batch_entry_two_calls(gpu, mode):
    heavy, t_heavy, nb = _heavy_curve()
    readme = dict(t0=0.0, tE=1.0, u0=-0.01, alpha=0.5, s=0.95, q=1e-2, rho=5e-3, ld_c=0.5, ld_d=0.0)
    t_readme = np.linspace(-0.5, 0.5, 200)
    o1 = gpu.lcbg_default_options()
    o1.mode = mode
    o1.source_bins = nb
    o2 = gpu.lcbg_default_options()
    o2.mode = mode
    # first call: the larger input; second: the smaller one, on the array the first left
    got1, gk1 = _noalloc_run(gpu, heavy, t_heavy, o1)
    got2, gk2 = _noalloc_run(gpu, readme, t_readme, o2)
    exp1, ek1 = _batch_run(gpu, heavy, t_heavy, o1)
    exp2, ek2 = _batch_run(gpu, readme, t_readme, o2)
    _same(got1, gk1, exp1, ek1, f"heavy curve mode {mode} nbin {nb}")
    _same(got2, gk2, exp2, ek2, f"README curve mode {mode}")
    assert (ek1 >= 2).sum() >= 10


def test_noalloc_rejects_oversized_input(gpu):
    import torch
    n = 8193
    d = torch.zeros(n, dtype=torch.float64, device="cuda")
    k = torch.zeros(n, dtype=torch.int32, device="cuda")
    rec = gpu.Curve(0.0, 1.0, 0.1, 0.0, 1.0, 0.1, 1e-3, 0.0, 0.0)
    st = gpu.lcbg_light_curve_device_noalloc(ctypes.byref(rec), None, d.data_ptr(), n, d.data_ptr(),
                                             k.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert st == 1  # LCBG_INVALID_ARGUMENT
    o = gpu.lcbg_default_options()
    o.caustic_bins = 4096
    st = gpu.lcbg_light_curve_device_noalloc(ctypes.byref(rec), ctypes.byref(o), d.data_ptr(), 10,
                                             d.data_ptr(), k.data_ptr(),
                                             torch.cuda.current_stream().cuda_stream)
    assert st == 1
