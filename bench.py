"""Throughput benchmark (one JSON line on stdout).

Workload (default): the point-source batch - 4096 binary-lens parameter sets x
256 epochs, one quintic solve per epoch (1,048,576 solves), FP64, synthetic
inputs with a fixed seed (lcbinint_amd.workloads.point_source_batch).  A step is
one launch of lcbg_batch_point_source_device over every epoch, with the epoch
times already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "LD finite-source mags/s and value+grad/s, batched vs plain GPU, 1e-4 rel err"
FP64_VALU_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak (vendor datasheet)
# Algorithmic FP64 operations per point-source solve (+,-,*,/,sqrt = 1 each),
# counted by the oracle's operation counter over the full point-source batch
# (oracle/count_flops.py).
FLOPS_PER_PS_SOLVE = 3158.2


def _cpu_baseline_worker(args):
    lib_path, kind, s, q, x, y = args
    from oracle.bindings import load_oracle, load_ref
    lib = load_ref(lib_path) if kind == "reference" else load_oracle(lib_path)
    t0 = time.perf_counter()
    lib.binary_mag0_many(s, q, x, y)
    return time.perf_counter() - t0, len(s)


def cpu_baseline(curves, offsets, times, max_solves=1 << 20):
    """Times the reference core (oracle/_ref) on the host cores over the same
    epochs, one process per core (the reference solver keeps static state)."""
    from multiprocessing import get_context

    ref = ROOT / "oracle" / "_ref" / "liblcbinint_ref.so"
    port = ROOT / "oracle" / "liblcb_oracle.so"
    if ref.exists():
        lib_path, kind = ref, "reference"
    elif port.exists():
        lib_path, kind = port, "port"
    else:
        return None
    n = min(len(times), max_solves)
    s = np.empty(n)
    q = np.empty(n)
    x = np.empty(n)
    y = np.empty(n)
    for c, cv in enumerate(curves):
        lo, hi = int(offsets[c]), min(int(offsets[c + 1]), n)
        if lo >= n:
            break
        tau = (times[lo:hi] - cv["t0"]) * (1.0 / cv["tE"])
        ca, sa = np.cos(cv["alpha"]), np.sin(cv["alpha"])
        x[lo:hi] = tau * ca - cv["u0"] * sa
        y[lo:hi] = tau * sa + cv["u0"] * ca
        s[lo:hi] = cv["s"]
        q[lo:hi] = 1.0 / cv["q"]
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    chunks = np.array_split(np.arange(n), cores)
    work = [(str(lib_path), kind, s[c], q[c], x[c], y[c]) for c in chunks]
    t0 = time.perf_counter()
    with get_context("spawn").Pool(cores) as pool:
        res = pool.map(_cpu_baseline_worker, work)
    wall = time.perf_counter() - t0
    busy = max(r[0] for r in res)
    return {
        "value": n / busy,
        "unit": "mags/s",
        "cores": cores,
        "kind": kind,
        "sample": f"{n} point-source solves of the same batch split over {cores} processes "
                  f"(slowest process {busy:.2f} s, wall {wall:.2f} s incl. process start)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the batch shards over "
              "curves with no exchange, so N GPUs run N independent shares.", file=sys.stderr)
        return 2

    import torch

    from lcbinint_amd import capi
    from lcbinint_amd.workloads import point_source_batch

    torch.cuda.set_device(0)
    curves, offsets, times = point_source_batch()
    n_epochs = len(times)
    batch = capi.Batch([capi.Curve(**c) for c in curves])
    d_off = torch.from_numpy(offsets).cuda()
    d_t = torch.from_numpy(times).cuda()
    d_m = torch.empty_like(d_t)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    for _ in range(args.warmup):
        batch.point_source(d_off.data_ptr(), d_t.data_ptr(), d_m.data_ptr(), sptr)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.point_source(d_off.data_ptr(), d_t.data_ptr(), d_m.data_ptr(), sptr)
        kernel_ms.append(capi.lcbg_last_kernel_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    batch.close()
    ms_per_step = 1e3 * wall / args.steps
    value = n_epochs / (wall / args.steps)
    avg_kernel_ms = float(np.mean(kernel_ms))
    achieved = FLOPS_PER_PS_SOLVE * n_epochs / (avg_kernel_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "mags/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "point-source batch: 4096 binary-lens parameter sets x 256 epochs, "
                        "one quintic solve per epoch (1,048,576 solves), seed 2",
            "curves": len(curves),
            "epochs": n_epochs,
            "kernel": "k_light_curves_ps",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VALU_PEAK_TFLOPS,
            "traffic": None,
            "kernel_ms": avg_kernel_ms,
            "flops_per_unit": FLOPS_PER_PS_SOLVE,
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(curves, offsets, times),
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
