"""The C-ABI library loads and exports every function its header declares."""
import ctypes
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_header_functions():
    from lcbinint_amd import capi
    names = capi.header_functions()
    assert "lcbg_batch_point_source_device" in names
    lib = ctypes.CDLL(str(capi.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_and_defaults():
    from lcbinint_amd import capi
    assert capi.lcbg_status_string(0) == b"ok"
    assert capi.lcbg_status_string(2) == b"numerical error"
    o = capi.lcbg_default_options()
    # lcbi_default_options (src/lcbinint/lcbinint.cpp:36-54)
    assert (o.source_bins, o.caustic_bins, o.mode) == (50, 1400, 4)
    assert (o.grid_ratio, o.point_source_threshold, o.hexadecapole_threshold,
            o.adaptive_hex_threshold) == (4.0, 20.0, 3.0, 0.001)


def test_invalid_arguments_rejected_without_gpu():
    from lcbinint_amd import capi
    # argument validation happens before any device call
    assert capi.lcbg_binary_mag0_device(None, None, None, None, -1, None, None, None) == 1
    assert capi.lcbg_batch_point_source_device(None, None, None, None, None) == 1
    h = ctypes.c_void_p()
    bad = (capi.Curve * 1)(capi.Curve(0.0, 0.0, 0.1, 0.0, 1.0, 0.1, 0.0, 0.0, 0.0))
    assert capi.lcbg_batch_create(bad, 1, None, ctypes.byref(h)) == 1  # tE == 0
