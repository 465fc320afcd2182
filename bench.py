"""Throughput benchmark (one JSON line on stdout).

Headline workload: the batched finite-source light curves - 2048 binary-lens
parameter sets with per-curve epoch counts floor(log-uniform[50, 2001)) (1,068,351
epochs), rho log-uniform [1e-4, 3e-2], linear limb darkening u uniform
[0.1, 0.9], nbin uniform in [25, 100] per curve, automatic Cartesian/polar
choice, FP64, synthetic inputs with a fixed seed
(lcbinint_amd.workloads.finite_source_batch).  One step re-evaluates every
light curve from its parameters: the curve records are uploaded again (which
marks the per-curve caustic caches stale), then one finite-source run rebuilds
the caches and runs point source, dispatcher, hexadecapole and inverse-ray
shooting over all epochs.  Epoch times are resident in HBM.

Also reported: the plain GPU method on the single README curve (200 epochs,
nbin 50): one epoch per launch on a reused one-curve batch, value and value +
gradient, and the host-pointer drop-in entry point called once per curve and
once per epoch; the 4096 x 256 point-source batch; the value + gradient
workload (256 x 1000); the parallax / orbital-motion workload; the reference
core compiled from source (oracle/_ref) timed on the host cores, on a bounded
sample of the headline workload and on the README curve (one process).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# A finite-source run keeps five streams busy at once (the caller's, the
# large-slot Cartesian and polar passes, the polar small-slot pass, the
# gradient enumeration).  With HIP's default of four hardware queues per
# process two of them share a queue and run one after the other: on the value
# + gradient workload the main seed search then waits for the large-slot
# Cartesian pass (239 ms per step instead of 208 ms).  A replay of the captured
# graph needs more: with eight queues it takes 180 ms per headline step
# against 145 ms eager, with sixteen 145 ms.  Raised to sixteen (a larger value
# the caller exported is kept) before the HIP runtime starts.
try:
    _hw_queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))
except ValueError:
    _hw_queues = 0
if _hw_queues < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "LD finite-source mags/s and value+grad/s, batched vs plain GPU, 1e-4 rel err"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak (vendor datasheet)

# Algorithmic FP64 operations (+ - * / sqrt = 1 each) per unit of work.
# Point-source solve: mean over the 4096 x 256 point-source batch counted by the
# oracle's operation counter (oracle/count_flops.py).  Critical-curve solve
# (quartic + lens equation of its 4 roots): mean over 1400 phases of 128 curves
# of the batched workload, same counter.  Lattice cells: the lens map, distance
# and limb-darkening weight as written in finite_source_magnifier.cpp:65-80,
# 121-153, 1241-1311 (Cartesian 30, Jacobian-guarded Cartesian 53, polar 37).
FLOPS_PS_SOLVE = 3158.2
FLOPS_CRIT_SOLVE = 1694.6
FLOPS_CART_CELL = 30.0
FLOPS_CART_GUARDED_CELL = 53.0
FLOPS_POLAR_CELL = 37.0

# HBM bytes per k_ir_cart launch on the headline workload (mean of its three
# launches of a step), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (profiles/k_ir_batched_fs.md): KiB per dispatch x 1024, FETCH_SIZE scaled by
# 2.0 and WRITE_SIZE by 1.0, the factors that bring each counter's reading for
# a 256 MiB device copy to 256 MiB (bench.py --copy-calibration).
K_IR_CART_TRAFFIC_BYTES = (2.0 * 1051500.850 + 1.0 * 1858051.950) * 1024 / 3
# FP64 lane operations the hardware issues per k_ir_cart launch on the same
# workload (rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64, 64 lanes per
# wave instruction, FMA as 2; profiles/k_ir_batched_fs.md): an upper bound on
# the executed FP64 operations, reported beside the algorithmic rate.
K_IR_CART_HW_FP64_FLOP = 1.491016e+11
COPY_BYTES = 1 << 28


def _cpu_worker(args):
    lib_path, items = args
    from oracle.bindings import default_settings, load_ref
    ref = load_ref(lib_path)
    t0 = time.perf_counter()
    n = 0
    for cv, nb, tt in items:
        st = default_settings(source_bins=int(nb), limb_darkening_c=cv["ld_c"])
        _, _, status = ref.light_curve(st, tt, cv["t0"], cv["tE"], cv["u0"], cv["alpha"], cv["s"],
                                       cv["q"], cv["rho"])
        if status != 0:
            raise RuntimeError(f"reference light curve failed with status {status}")
        n += len(tt)
    return time.perf_counter() - t0, n


def cpu_baseline(curves, offsets, times, nbin, stride=2):
    """The reference core (oracle/_ref, compiled from its own sources) on every
    `stride`-th curve of the same workload, one process per core (the
    reference solver keeps static state, so it is not run threaded)."""
    from multiprocessing import get_context

    ref = ROOT / "oracle" / "_ref" / "liblcbinint_ref.so"
    if not ref.exists():
        return None
    idx = np.arange(0, len(curves), stride)
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    # longest-first round robin so the processes finish together
    counts = np.diff(offsets)[idx]
    order = idx[np.argsort(-counts, kind="stable")]
    parts = [[] for _ in range(cores)]
    for k, c in enumerate(order):
        parts[k % cores].append((curves[c], int(nbin[c]), times[offsets[c]:offsets[c + 1]]))
    with get_context("spawn").Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(str(ref), p) for p in parts])
    busy = max(r[0] for r in res)
    n = sum(r[1] for r in res)
    return {
        "value": n / busy,
        "unit": "mags/s",
        "cores": cores,
        "kind": "reference",
        "sample": f"every {stride}th curve of the batched workload ({len(idx)} curves, {n} epochs, "
                  f"{sum(r[0] for r in res):.1f} s of CPU work) split over {cores} processes; "
                  f"slowest process {busy:.2f} s",
    }


def bench_batched(capi, torch, steps, warmup, autotune=True):
    from lcbinint_amd.workloads import finite_source_batch
    curves, offsets, times, nbin = finite_source_batch()
    n_epochs = len(times)
    recs = (capi.Curve * len(curves))(*[capi.Curve(**c) for c in curves])
    batch = capi.Batch(list(recs))
    batch.set_source_bins(nbin)
    d_off = torch.from_numpy(offsets).cuda()
    d_t = torch.from_numpy(times).cuda()
    d_m = torch.empty_like(d_t)
    d_k = torch.empty(n_epochs, dtype=torch.int32, device="cuda")
    sptr = torch.cuda.current_stream().cuda_stream

    def step():
        capi.check(capi.lcbg_batch_update_curves(batch.handle, recs, sptr), "update_curves")
        return capi.finite_source(batch, d_off.data_ptr(), d_t.data_ptr(), n_epochs,
                                  d_m.data_ptr(), d_k.data_ptr(), sptr, profile=True)

    # launch schedule chosen by the autotuner on this workload (untimed)
    schedule, tuned_ms = capi.autotune(batch, d_off.data_ptr(), d_t.data_ptr(), n_epochs, d_m.data_ptr(),
                                       d_k.data_ptr(), sptr, max_runs=18) if autotune else (None, None)
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    stages, launches = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ms, qc, flags = step()
        if flags:
            raise RuntimeError(f"finite-source run reported capacity flags {flags}")
        stages.append(ms)
        launches.append(capi.ir_kernel_ms(batch))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    work = capi.work_counts(batch)
    meth = np.bincount(d_k.cpu().numpy(), minlength=4)
    finite = bool(torch.isfinite(d_m).all().item())
    if not finite:
        raise RuntimeError("non-finite magnification in the batched run")
    stage_ms = np.mean(np.array(stages), axis=0)
    launch_ms = {k: float(np.mean([l[k] for l in launches])) for k in launches[0]}
    # the same steps replayed from a captured HIP graph (same buffers)
    gm = torch.empty_like(d_t)
    gk = torch.empty(n_epochs, dtype=torch.int32, device="cuda")
    capi.capture_finite_source(batch, d_off.data_ptr(), d_t.data_ptr(), n_epochs, gm.data_ptr(),
                               gk.data_ptr(), sptr)
    capi.check(capi.lcbg_batch_update_curves(batch.handle, recs, sptr), "update_curves")
    capi.launch_graph(batch, sptr)
    torch.cuda.synchronize()
    tg = time.perf_counter()
    for _ in range(steps):
        capi.check(capi.lcbg_batch_update_curves(batch.handle, recs, sptr), "update_curves")
        capi.launch_graph(batch, sptr)
    torch.cuda.synchronize()
    graph_ms = 1e3 * (time.perf_counter() - tg) / steps
    graph_same = bool(torch.equal(gm, d_m)) and bool(torch.equal(gk, d_k))
    batch.close()
    return dict(n_curves=len(curves), n_epochs=n_epochs, wall=wall, steps=steps,
                stage_ms=stage_ms.tolist(), launch_ms=launch_ms, queue=qc, work=work,
                methods=meth.tolist(), graph_ms=graph_ms, graph_same=graph_same,
                schedule=schedule.as_dict() if schedule is not None else None, tuned_ms=tuned_ms)


def bench_value_grad(capi, torch, steps, warmup):
    """chi^2 and its gradient w.r.t. (t0, tE, u0, alpha, s, q, rho, u_LD) of
    256 curves x 1000 epochs (lcbinint_amd.workloads.differentiation_batch),
    against observed fluxes made from the model (fs = 1, fb = 0.1, 1% noise).
    A step uploads the curve parameters and evaluates magnifications, their
    derivatives and the per-curve chi^2 gradient.  The plain per-epoch path
    is plain_per_epoch(grad=True)."""
    from lcbinint_amd.workloads import differentiation_batch, synthetic_fluxes
    curves, offsets, times = differentiation_batch()
    n, nc = len(times), len(curves)
    recs = (capi.Curve * nc)(*[capi.Curve(**c) for c in curves])
    batch = capi.Batch(list(recs))
    d_off = torch.from_numpy(offsets).cuda()
    d_t = torch.from_numpy(times).cuda()
    mag = torch.empty(n, dtype=torch.float64, device="cuda")
    meth = torch.empty(n, dtype=torch.int32, device="cuda")
    sptr = torch.cuda.current_stream().cuda_stream
    capi.finite_source(batch, d_off.data_ptr(), d_t.data_ptr(), n, mag.data_ptr(), meth.data_ptr(),
                       sptr)
    flux, sigma = synthetic_fluxes(mag.cpu().numpy())
    d_flux = torch.from_numpy(flux).cuda()
    d_sig = torch.from_numpy(sigma).cuda()
    chi2 = torch.empty(nc, dtype=torch.float64, device="cuda")
    grad = torch.empty(nc * 8, dtype=torch.float64, device="cuda")

    def step():
        capi.check(capi.lcbg_batch_update_curves(batch.handle, recs, sptr), "update_curves")
        capi.chi2_grad(batch, d_off.data_ptr(), d_t.data_ptr(), n, d_flux.data_ptr(),
                       d_sig.data_ptr(), 1.0, 0.1, mag.data_ptr(), meth.data_ptr(),
                       chi2.data_ptr(), grad.data_ptr(), None, None, sptr)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    methods = np.bincount(meth.cpu().numpy(), minlength=4).tolist()
    finite = bool(torch.isfinite(grad).all().item()) and bool(torch.isfinite(chi2).all().item())
    batch.close()
    if not finite:
        raise RuntimeError("non-finite chi^2 gradient")
    # plain per-epoch path: one epoch per launch on a reused one-curve handle
    # (no allocation, caustic cache rebuilt only when the curve changes), the
    # epochs of a few curves in order, as the reference's epoch loop runs them
    per_epoch, per_epoch_meth = plain_per_epoch(capi, torch, grad=True)
    value_grad = n * steps / wall
    return {
        "workload": "chi^2 and gradient w.r.t. (t0,tE,u0,alpha,s,q,rho,u_LD): 256 curves x 1000 "
                    "epochs in [-2, 2], s log-uniform [0.3, 3], q log-uniform [1e-5, 1], rho "
                    "log-uniform [1e-4, 3e-2], linear LD u in [0.1, 0.9], nbin 50, seed 4; "
                    "fluxes F = A + 0.1 with 1% noise",
        "value_grad_per_s": value_grad, "unit": "epochs/s", "ms_per_step": 1e3 * wall / steps,
        "methods": methods, "plain_per_epoch_value_grad_per_s": per_epoch,
        "plain_per_epoch_methods": per_epoch_meth,
        "plain_per_epoch": "README curve (200 epochs, nbin 50, auto grid), one epoch per launch of "
                           "lcbg_batch_finite_source_grad_device on a reused one-curve batch (no "
                           "allocation, caustic cache built once), result read back per epoch",
        "batched_over_plain_per_epoch": value_grad / per_epoch,
    }


def plain_per_epoch(capi, torch, grad):
    """The plain GPU method on the README curve (200 epochs, nbin 50, auto
    grid): one epoch per launch, in time order, on one reused one-curve batch
    (no allocation; the caustic cache is built once), the result copied back
    after every launch, as the reference's epoch loop evaluates a curve.
    Returns (epochs/s, method counts)."""
    from lcbinint_amd.workloads import plain_curve
    cv, times = plain_curve()
    one_t = torch.empty(1, dtype=torch.float64, device="cuda")
    one_off = torch.tensor([0, 1], dtype=torch.int64, device="cuda")
    one_m = torch.empty(1, dtype=torch.float64, device="cuda")
    one_k = torch.empty(1, dtype=torch.int32, device="cuda")
    one_d = torch.empty(8, dtype=torch.float64, device="cuda")
    sptr = torch.cuda.current_stream().cuda_stream
    batch = capi.Batch([capi.Curve(**cv)])
    meth = np.zeros(4, dtype=np.int64)

    def one(t):
        one_t.fill_(float(t))
        if grad:
            capi.finite_source_grad(batch, one_off.data_ptr(), one_t.data_ptr(), 1, one_m.data_ptr(),
                                    one_k.data_ptr(), 0, one_d.data_ptr(), sptr)
            _ = one_d.cpu()
        else:
            capi.finite_source(batch, one_off.data_ptr(), one_t.data_ptr(), 1, one_m.data_ptr(),
                               one_k.data_ptr(), sptr)
        return int(one_k.cpu()[0])

    one(times[0])  # warm-up: workspace, scratch and the caustic cache
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in times:
        meth[one(t)] += 1
    elapsed = time.perf_counter() - t0
    batch.close()
    return len(times) / elapsed, meth.tolist()


def bench_dynamic(capi, n_curves=1024, steps=1):
    """Light curves with annual parallax and circular orbital motion: the
    1024-curve workload of lcbinint_amd.workloads.dynamic_batch (500-3000
    epochs per curve over 200 days), substitute Earth table of tests/golden.
    Value, and value plus the derivatives with respect to 12 parameters, through
    lcbg_batch_light_curves_dynamic_device (offsets and times resident on the
    device; the per-curve trajectory constants computed on the host inside the
    timed region).  The same curves with parallax alone (no orbital motion:
    one caustic cache per curve, reused across calls) are timed beside."""
    import torch
    from lcbinint_amd.workloads import dynamic_batch
    table = ROOT / "tests" / "golden" / "earth_ephemeris_substitute.txt"
    capi.load_earth_ephemeris(table)
    curves, dyn, off, t = dynamic_batch(n_curves=n_curves)
    n = len(t)
    b = capi.Batch([capi.Curve(**c) for c in curves])
    d = (capi.Dynamics * n_curves)(*[capi.Dynamics(**x) for x in dyn])
    d_off = torch.from_numpy(off).cuda()
    d_t = torch.from_numpy(t).cuda()
    mag = torch.empty(n, dtype=torch.float64, device="cuda")
    meth = torch.empty(n, dtype=torch.int32, device="cuda")
    dm = torch.empty((n, 12), dtype=torch.float64, device="cuda")
    sp = torch.cuda.current_stream().cuda_stream

    def run(mode, grad):
        capi.check(capi.lcbg_batch_light_curves_dynamic_device(
            b.handle, d, mode, d_off.data_ptr(), d_t.data_ptr(), n, mag.data_ptr(), meth.data_ptr(), None,
            dm.data_ptr() if grad else None, sp), "lcbg_batch_light_curves_dynamic_device")

    out = {}
    for mode, label in ((capi.ORBIT_CIRCULAR, "circular"), (capi.ORBIT_STATIC, "parallax_only")):
        rates = {}
        for grad in (False, True):
            run(mode, grad)  # warm-up (workspace and, without orbital motion, caustic caches)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                run(mode, grad)
            torch.cuda.synchronize()
            rates["value_grad" if grad else "value"] = n * steps / (time.perf_counter() - t0)
            finite = bool(torch.isfinite(mag).all()) and (not grad or bool(torch.isfinite(dm).all()))
            if not finite:
                raise RuntimeError(f"non-finite value or derivative in the dynamic workload ({label})")
        out[label] = {"value_epochs_per_s": rates["value"], "value_grad_epochs_per_s": rates["value_grad"],
                      "methods": np.bincount(meth.cpu().numpy(), minlength=4).tolist()}
    b.close()
    return {
        "workload": f"annual parallax + circular orbital motion: {n_curves} curves, {n} epochs over 200 days "
                    "(500-3000 per curve), rho log-uniform [1e-4, 1e-2], linear LD U(0.1, 0.9), nbin 50, "
                    "piEN/piEE U(-0.5, 0.5), ds/dt, dalpha/dt U(-1, 1)/yr, seed 5; device arrays",
        "grad_params": 12,
        **out,
    }


def bench_plain(capi, calls_per_epoch=50):
    """The unbatched drop-in path on the README curve: lcbg_light_curve with
    host arrays, once for the whole curve and once per epoch."""
    from lcbinint_amd.workloads import plain_curve
    cv, t = plain_curve()
    rec = capi.Curve(**cv)
    opt = capi.lcbg_default_options()
    mag = np.empty_like(t)
    meth = np.empty(len(t), dtype=np.int32)
    P = capi.c_double_p

    def run(tt, out, mt):
        capi.check(capi.lcbg_light_curve(ctypes.byref(rec), ctypes.byref(opt),
                                         tt.ctypes.data_as(P), len(tt), out.ctypes.data_as(P),
                                         mt.ctypes.data_as(capi.c_int32_p)), "lcbg_light_curve")

    run(t, mag, meth)  # warm-up
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        run(t, mag, meth)
    per_curve = reps * len(t) / (time.perf_counter() - t0)
    idx = np.linspace(0, len(t) - 1, calls_per_epoch).astype(int)
    one = np.empty(1)
    one_m = np.empty(1, dtype=np.int32)
    t0 = time.perf_counter()
    for i in idx:
        run(t[i:i + 1], one, one_m)
    per_epoch = len(idx) / (time.perf_counter() - t0)
    return per_curve, per_epoch, np.bincount(meth, minlength=4).tolist()


def cpu_readme_curve():
    """The reference core (oracle/_ref) on the README curve, one process, as
    LightCurve(...)(times, params) runs it: ms per epoch."""
    from lcbinint_amd.workloads import plain_curve
    from oracle.bindings import default_settings, load_ref
    path = ROOT / "oracle" / "_ref" / "liblcbinint_ref.so"
    if not path.exists():
        return None
    ref = load_ref(path)
    cv, t = plain_curve()
    st = default_settings(source_bins=50, limb_darkening_c=cv["ld_c"])
    args = (cv["t0"], cv["tE"], cv["u0"], cv["alpha"], cv["s"], cv["q"], cv["rho"])
    ref.light_curve(st, t, *args)  # warm-up
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        _, meth, status = ref.light_curve(st, t, *args)
    dt = (time.perf_counter() - t0) / reps
    return {"ms_per_epoch": 1e3 * dt / len(t), "mags_per_s": len(t) / dt, "cores": 1, "kind": "reference",
            "sample": f"README curve, {len(t)} epochs, nbin 50, auto grid, {reps} repeats, one process",
            "methods": np.bincount(meth, minlength=4).tolist()}


def bench_point_source(capi, torch, steps=10):
    from lcbinint_amd.workloads import point_source_batch
    curves, offsets, times = point_source_batch()
    batch = capi.Batch([capi.Curve(**c) for c in curves])
    d_off = torch.from_numpy(offsets).cuda()
    d_t = torch.from_numpy(times).cuda()
    d_m = torch.empty_like(d_t)
    sptr = torch.cuda.current_stream().cuda_stream
    batch.point_source(d_off.data_ptr(), d_t.data_ptr(), d_m.data_ptr(), sptr)
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.point_source(d_off.data_ptr(), d_t.data_ptr(), d_m.data_ptr(), sptr)
        kms.append(capi.lcbg_last_kernel_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    batch.close()
    kernel_ms = float(np.mean(kms))
    tflops = FLOPS_PS_SOLVE * len(times) / (kernel_ms * 1e-3) / 1e12
    return {"value": len(times) * steps / wall, "unit": "mags/s", "epochs": len(times),
            "kernel": "k_light_curves_ps", "kernel_ms": kernel_ms, "achieved_tflops": tflops,
            "frac": tflops / FP64_PEAK_TFLOPS}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the plain-path and point-source measurements")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch schedule")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three device copies of 256 MiB (the reference reading for "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for {args.gpus}): light curves are independent, "
              "so N GPUs would each run their own share of the curves with no exchange.",
              file=sys.stderr)
        return 2

    import torch

    from lcbinint_amd import capi

    torch.cuda.set_device(0)
    if args.copy_calibration:
        src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": COPY_BYTES, "copies": 3}))
        return 0
    print("[bench] batched finite-source workload", file=sys.stderr, flush=True)
    b = bench_batched(capi, torch, args.steps, args.warmup, autotune=not args.no_autotune)
    value = b["n_epochs"] * b["steps"] / b["wall"]
    w = b["work"]
    ir_flops = (w["quintic"] * FLOPS_PS_SOLVE + w["quartic"] * FLOPS_CRIT_SOLVE +
                w["cart"] * FLOPS_CART_CELL + w["cart_guarded"] * FLOPS_CART_GUARDED_CELL +
                w["polar"] * FLOPS_POLAR_CELL)
    ir_ms = b["stage_ms"][3]
    # dominant kernel: k_ir_cart (three launches per step: large slots, small
    # slots, retries); per-launch work and duration are the step's totals / 3
    lm = b["launch_ms"]
    cart_ms = (lm["cart_big"] + lm["cart_small"] + lm["cart_retry"]) / 3
    cart_flops = (w["cart"] * FLOPS_CART_CELL + w["cart_guarded"] * FLOPS_CART_GUARDED_CELL) / 3
    achieved = cart_flops / (cart_ms * 1e-3) / 1e12
    extras = {}
    if not args.no_extras:
        print("[bench] plain path and point-source batch", file=sys.stderr, flush=True)
        per_curve, per_epoch, plain_meth = bench_plain(capi)
        launch_rate, launch_meth = plain_per_epoch(capi, torch, grad=False)
        extras["plain_gpu"] = {
            "workload": "README curve: t0=0 tE=1 u0=-0.01 alpha=0.5 s=0.95 q=0.01 rho=5e-3, "
                        "linear LD 0.5, 200 epochs, nbin 50, auto grid",
            "per_curve_call_mags_per_s": per_curve, "per_epoch_call_mags_per_s": per_epoch,
            "methods": plain_meth,
            "batched_over_per_epoch_call": value / per_epoch,
            "batched_over_per_curve_call": value / per_curve,
            "per_epoch_launch_mags_per_s": launch_rate,
            "per_epoch_launch_methods": launch_meth,
            "per_epoch_launch": "the same curve, one epoch per launch of lcbg_batch_finite_source_device "
                                "on a reused one-curve batch (no allocation, caustic cache built once), "
                                "result read back per epoch",
            "batched_over_per_epoch_launch": value / launch_rate,
        }
        extras["point_source_batch"] = bench_point_source(capi, torch)
        print("[bench] value and gradient workload", file=sys.stderr, flush=True)
        extras["value_grad"] = bench_value_grad(capi, torch, max(2, args.steps // 2), 1)
        print("[bench] parallax and orbital motion workload", file=sys.stderr, flush=True)
        extras["dynamic"] = bench_dynamic(capi)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] reference CPU baseline", file=sys.stderr, flush=True)
        from lcbinint_amd.workloads import finite_source_batch
        curves, offsets, times, nbin = finite_source_batch()
        cpu = cpu_baseline(curves, offsets, times, nbin)
        extras["cpu_readme_curve"] = cpu_readme_curve()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "mags/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * b["wall"] / b["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched finite-source light curves: 2048 curves, 1,068,351 epochs "
                        "(per-curve counts log-uniform in [50, 2000]), rho log-uniform "
                        "[1e-4, 3e-2], linear LD u in [0.1, 0.9], nbin 25-100 per curve, auto "
                        "Cartesian/polar, seed 3; caustic caches rebuilt every step",
            "curves": b["n_curves"],
            "epochs": b["n_epochs"],
            "methods": b["methods"],
            "stage_ms": {"caustic_cache": b["stage_ms"][0], "point_source": b["stage_ms"][1],
                         "dispatch_hexadecapole": b["stage_ms"][2],
                         "inverse_ray": b["stage_ms"][3]},
            "parallelism": "one process, one GPU",
            "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
            "schedule": {"autotuned": b["schedule"], "tuning_run_ms": b["tuned_ms"]},
            "graph_replay": {"ms_per_step": b["graph_ms"], "mags_per_s": b["n_epochs"] / (b["graph_ms"] * 1e-3),
                             "bitwise_equal_to_eager": b["graph_same"]},
        },
        "roofline": {
            # the lattice scan issues FP64 VALU work (no matrix instructions);
            # the roof is MI355X's FP64 vector peak
            "bound": "fp64_valu",
            "kernel": "k_ir_cart",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": K_IR_CART_TRAFFIC_BYTES,
            "kernel_ms": cart_ms,
            "launches_per_step": 3,
            "algorithmic_flops_per_launch": cart_flops,
            "achieved_hbm_gbps": K_IR_CART_TRAFFIC_BYTES / (cart_ms * 1e-3) / 1e9,
            "hw_fp64_tflops": K_IR_CART_HW_FP64_FLOP / (cart_ms * 1e-3) / 1e12,
            "hw_fp64_frac": K_IR_CART_HW_FP64_FLOP / (cart_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "inverse_ray_stage_ms": ir_ms,
            "inverse_ray_stage_flops": ir_flops,
            # the three k_ir_cart launches overlap each other and the polar
            # and seed launches: the whole stage's algorithmic rate over its
            # wall time, beside the per-launch figure
            "inverse_ray_stage_tflops": ir_flops / (ir_ms * 1e-3) / 1e12,
            "inverse_ray_stage_frac": ir_flops / (ir_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "launch_ms": lm,
            "work": w,
        },
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
