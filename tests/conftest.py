import ctypes
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

ORACLE_LIB = ROOT / "oracle" / "liblcb_oracle.so"
REF_LIB = ROOT / "oracle" / "_ref" / "liblcbinint_ref.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    """The C restatement (test infrastructure; see oracle/lcb_oracle.h)."""
    from oracle import bindings
    if not ORACLE_LIB.exists():
        pytest.fail(f"{ORACLE_LIB} missing: run `make -C oracle oracle`")
    return bindings.load_oracle(ORACLE_LIB)


@pytest.fixture(scope="session")
def ref():
    """The reference core compiled from source (oracle/_ref), if built."""
    from oracle import bindings
    if not REF_LIB.exists():
        pytest.skip("oracle/_ref not built (needs the reference sources at build time)")
    return bindings.load_ref(REF_LIB)


@pytest.fixture(scope="session")
def gpu():
    """The HIP library plus a torch CUDA device; fails loudly if either is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test scheduled without a visible GPU")
    from lcbinint_amd import capi
    return capi
