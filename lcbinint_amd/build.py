"""Builds the gfx950 HIP library and the Python extension in-tree.

    python -m lcbinint_amd.build          # or lcbinint_amd.build.build_all()

Outputs (next to this file):
    liblcbinint_amd.so                    C-ABI library (include/lcbinint_amd.h)
    _lcbinint.cpython-*.so                pybind11 module behind lcbinint_amd.LightCurve
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = os.environ.get("LCBG_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-add anywhere, so every FP64 operation
# rounds exactly as the reference's x86-64 SSE2 build does.
# -packed-fp32-ops: keep the compiler from forming packed FP32 instructions.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{INCLUDE}", f"-I{CSRC}",
]

LIB = PKG / "liblcbinint_amd.so"


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... (exit {proc.returncode})")
    err = "\n".join(
        line for line in proc.stderr.splitlines()
        if "packed-fp32-ops" not in line and "not a recognized feature" not in line
        and line.strip()
    )
    if err:
        sys.stderr.write(err + "\n")


# Host-only sources are compiled by g++ so that their libm calls (cos/sin of
# the trajectory angle, the fixed probe and phase tables) are glibc's, as in the
# reference build; hipcc's host pass resolves std::cos/std::sin to a different
# implementation that rounds some arguments one ulp away.
HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
    f"-I{ROCM / 'include'}", f"-I{INCLUDE}", f"-I{CSRC}",
]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)])
    else:
        _run(["g++", *HOST_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_library() -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    sources = [s for s in sources if not s.name.startswith("py_")]
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_python_module() -> Path:
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = PKG / f"_lcbinint{ext}"
    tmp = out.with_suffix(".tmp")
    py_inc = sysconfig.get_paths()["include"]
    _run([
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
        f"-I{pybind11.get_include()}", f"-I{py_inc}", f"-I{INCLUDE}",
        str(CSRC / "py_lcbinint.cpp"), "-o", str(tmp),
        f"-L{PKG}", "-llcbinint_amd", "-Wl,-rpath,$ORIGIN",
    ])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_library()
    if (CSRC / "py_lcbinint.cpp").exists():
        build_python_module()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}")
