"""Counts the oracle's FP64 operations per point-source solve on the
point-source batch workload (4096 sets x 256 epochs, seed 2).

Convention: +, -, *, /, sqrt = 1 operation each; comparisons 0.  The count is
the algorithmic work figure behind the point-source kernel's roofline.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

from lcbinint_amd.workloads import point_source_batch  # noqa: E402
from oracle.bindings import load_oracle  # noqa: E402


def main():
    o = load_oracle(ROOT / "oracle" / "liblcb_oracle.so")
    curves, offsets, times = point_source_batch()
    o.flop_reset()
    n = 0
    for c, cv in enumerate(curves):
        sl = slice(offsets[c], offsets[c + 1])
        o.light_curve_ps(times[sl], cv["t0"], cv["tE"], cv["u0"], cv["alpha"], cv["s"], cv["q"])
        n += offsets[c + 1] - offsets[c]
    print(f"solves {n}  flops/solve {o.flop_count() / n:.1f}")


if __name__ == "__main__":
    main()
