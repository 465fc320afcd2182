"""ctypes binding of the C ABI declared in include/lcbinint_amd.h.

The library must have been built (``python -m lcbinint_amd.build``); importing
this module raises if it is missing, so no code path can silently fall back to
a CPU implementation.
"""
from __future__ import annotations

import ctypes

import numpy as np
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "liblcbinint_amd.so"
HEADER = PKG.parent / "include" / "lcbinint_amd.h"

# PyTorch-ROCm carries its own HIP runtime.  Loading torch before this
# library makes the library bind to that same runtime; loading the library
# first would start the system runtime, and torch would then find no GPU.
try:
    import torch  # noqa: F401
except ImportError:  # the library itself does not need torch
    pass

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -m lcbinint_amd.build` (needs hipcc for gfx950)"
    )

lib = ctypes.CDLL(str(LIB_PATH))

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_int32_p = ctypes.POINTER(ctypes.c_int32)

STATUS_NAMES = {0: "ok", 1: "invalid argument", 2: "numerical error", 3: "unsupported",
                4: "device error"}


class Curve(ctypes.Structure):
    """lcbg_curve: one light curve's model in VBM coordinates."""
    _fields_ = [(name, ctypes.c_double) for name in
                ("t0", "tE", "u0", "alpha", "s", "q", "rho", "ld_c", "ld_d")]


class Triple(ctypes.Structure):
    """lcbg_triple: a triple-lens light curve (VBM geometry)."""
    _fields_ = [(name, ctypes.c_double) for name in
                ("t0", "tE", "u0", "alpha", "s", "q", "q2", "sep2", "ang")]


class Options(ctypes.Structure):
    """lcbg_options: the finite-source options (plain method, retry, adaptive mode)."""
    _fields_ = [
        ("source_bins", ctypes.c_int), ("caustic_bins", ctypes.c_int),
        ("mode", ctypes.c_int), ("polar_source_bins", ctypes.c_int),
        ("grid_ratio", ctypes.c_double), ("polar_grid_ratio", ctypes.c_double),
        ("point_source_threshold", ctypes.c_double),
        ("hexadecapole_threshold", ctypes.c_double),
        ("adaptive_hex_threshold", ctypes.c_double),
        ("finite_source_tol", ctypes.c_double),
        ("finite_source_reltol", ctypes.c_double),
        ("adaptive_source_bins", ctypes.c_int),
        ("max_source_bins", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    """lcbg_result: one epoch's record (the fields of the reference's lcbi_result)."""
    _fields_ = [
        ("magnification", ctypes.c_double), ("point_source_magnification", ctypes.c_double),
        ("finite_source_magnification", ctypes.c_double), ("source_x", ctypes.c_double),
        ("source_y", ctypes.c_double), ("image_count", ctypes.c_int32), ("method", ctypes.c_int32),
        ("error_estimate", ctypes.c_double), ("refinement_level", ctypes.c_int32),
        ("converged", ctypes.c_int32),
    ]


# numpy view of an array of lcbg_result records
RESULT_DTYPE = np.dtype([(name, np.float64 if typ is ctypes.c_double else np.int32)
                         for name, typ in Result._fields_], align=True)
assert RESULT_DTYPE.itemsize == ctypes.sizeof(Result)


def _sig(name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


lcbg_default_options = _sig("lcbg_default_options", Options)
lcbg_status_string = _sig("lcbg_status_string", ctypes.c_char_p, ctypes.c_int)
lcbg_binary_mag0_device = _sig(
    "lcbg_binary_mag0_device", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
lcbg_batch_create = _sig(
    "lcbg_batch_create", ctypes.c_int, ctypes.POINTER(Curve), ctypes.c_int32,
    ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_void_p))
lcbg_batch_destroy = _sig("lcbg_batch_destroy", ctypes.c_int, ctypes.c_void_p)
lcbg_batch_update_curves = _sig("lcbg_batch_update_curves", ctypes.c_int, ctypes.c_void_p,
                                ctypes.POINTER(Curve), ctypes.c_void_p)
lcbg_batch_set_source_bins = _sig("lcbg_batch_set_source_bins", ctypes.c_int, ctypes.c_void_p,
                                  c_int32_p)
lcbg_batch_point_source_device = _sig(
    "lcbg_batch_point_source_device", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
lcbg_light_curve_point_source = _sig(
    "lcbg_light_curve_point_source", ctypes.c_int, ctypes.POINTER(Curve), c_double_p,
    ctypes.c_int64, c_double_p)
lcbg_last_kernel_ms = _sig("lcbg_last_kernel_ms", ctypes.c_double)
lcbg_batch_set_epoch_cost = _sig("lcbg_batch_set_epoch_cost", ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_void_p)
lcbg_wall_clock_rate_khz = _sig("lcbg_wall_clock_rate_khz", ctypes.c_int64)


class LcbgError(RuntimeError):
    def __init__(self, status: int, where: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}")
        self.status = status


def check(status: int, where: str) -> None:
    if status != 0:
        raise LcbgError(status, where)


def header_functions() -> list[str]:
    """Names of every function the public header declares."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lcbg_[a-z0-9_]+)\s*\(", text)))


class Batch:
    """Owns an lcbg_batch handle (per-curve records resident on the device)."""

    def __init__(self, curves, options: Options | None = None):
        arr = (Curve * len(curves))(*curves)
        handle = ctypes.c_void_p()
        opts = ctypes.byref(options) if options is not None else None
        check(lcbg_batch_create(arr, len(curves), opts, ctypes.byref(handle)), "lcbg_batch_create")
        self.handle = handle
        self.n_curves = len(curves)

    def update_curves(self, curves, stream: int = 0):
        arr = (Curve * len(curves))(*curves)
        if len(curves) != self.n_curves:
            raise ValueError(f"update_curves needs {self.n_curves} curves, got {len(curves)}")
        check(lcbg_batch_update_curves(self.handle, arr, stream), "lcbg_batch_update_curves")

    def set_source_bins(self, source_bins):
        import numpy as np
        nb = np.ascontiguousarray(source_bins, dtype=np.int32)
        if nb.shape != (self.n_curves,):
            raise ValueError(f"source_bins needs {self.n_curves} entries, got {nb.shape}")
        check(lcbg_batch_set_source_bins(self.handle, nb.ctypes.data_as(c_int32_p)),
              "lcbg_batch_set_source_bins")

    def point_source(self, offsets_ptr: int, times_ptr: int, mag_ptr: int, stream: int = 0):
        check(lcbg_batch_point_source_device(self.handle, offsets_ptr, times_ptr, mag_ptr, stream),
              "lcbg_batch_point_source_device")

    def close(self):
        if self.handle:
            lcbg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


lcbg_batch_finite_source_device = _sig(
    "lcbg_batch_finite_source_device", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
lcbg_batch_finite_source_profile = _sig(
    "lcbg_batch_finite_source_profile", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32),
    ctypes.POINTER(ctypes.c_uint32))
lcbg_light_curve = _sig(
    "lcbg_light_curve", ctypes.c_int, ctypes.POINTER(Curve), ctypes.POINTER(Options), c_double_p,
    ctypes.c_int64, c_double_p, c_int32_p)


def finite_source(batch: Batch, offsets_ptr: int, times_ptr: int, n_epochs: int, mag_ptr: int,
                  method_ptr: int, stream: int = 0, profile: bool = False):
    """Runs the finite-source pipeline; with profile=True returns
    (stage_ms[4], queue_counts[2], error_flags)."""
    if not profile:
        check(lcbg_batch_finite_source_device(batch.handle, offsets_ptr, times_ptr, n_epochs,
                                              mag_ptr, method_ptr, stream),
              "lcbg_batch_finite_source_device")
        return None
    ms = (ctypes.c_float * 4)()
    qc = (ctypes.c_uint32 * 2)()
    fl = ctypes.c_uint32()
    check(lcbg_batch_finite_source_profile(batch.handle, offsets_ptr, times_ptr, n_epochs, mag_ptr,
                                           method_ptr, stream, ms, qc, ctypes.byref(fl)),
          "lcbg_batch_finite_source_profile")
    return list(ms), list(qc), fl.value


lcbg_batch_debug_epoch = _sig(
    "lcbg_batch_debug_epoch", ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
    ctypes.c_int32, c_double_p, ctypes.c_int32, ctypes.c_int32, c_int32_p, c_double_p)
lcbg_batch_debug_critical_sources = _sig(
    "lcbg_batch_debug_critical_sources", ctypes.c_int, ctypes.c_void_p, ctypes.c_int32, c_double_p)


def debug_critical_sources(batch: Batch, curve: int):
    """The cached critical sources of the seed scan of one curve, (1400, 4, 2)."""
    import numpy as np
    out = np.zeros(1400 * 4 * 2)
    check(lcbg_batch_debug_critical_sources(batch.handle, curve, out.ctypes.data_as(c_double_p)),
          "lcbg_batch_debug_critical_sources")
    return out.reshape(1400, 4, 2)


def debug_epoch(batch: Batch, epoch: int, method: int, seeds=None, capacity: int = 1024,
                last_phase: int = -1):
    """Seeds and inverse-ray magnification of one epoch of the batch's last
    finite-source run.  seeds=None searches them on the device (seed phases up
    to last_phase, -1 = all); otherwise the given (n, 2) seeds are used.
    Returns (seeds (n, 2), magnification)."""
    import numpy as np
    n_out = ctypes.c_int32()
    mag = ctypes.c_double()
    if seeds is None:
        buf = np.zeros(2 * capacity)
        n_in = -1
    else:
        buf = np.ascontiguousarray(np.asarray(seeds, dtype=np.float64).ravel())
        n_in = len(buf) // 2
        capacity = n_in
    check(lcbg_batch_debug_epoch(batch.handle, epoch, method, last_phase,
                                 buf.ctypes.data_as(c_double_p),
                                 capacity, n_in, ctypes.byref(n_out), ctypes.byref(mag)),
          "lcbg_batch_debug_epoch")
    n = n_out.value
    out = buf[:2 * max(min(n, capacity), 0)].reshape(-1, 2).copy()
    return out, mag.value

lcbg_batch_work_counts = _sig("lcbg_batch_work_counts", ctypes.c_int, ctypes.c_void_p,
                              ctypes.POINTER(ctypes.c_uint64))


def work_counts(batch: Batch):
    """Inverse-ray work of the last run: dict of quintic, quartic, cart,
    cart_guarded, polar counts."""
    out = (ctypes.c_uint64 * 8)()
    check(lcbg_batch_work_counts(batch.handle, out), "lcbg_batch_work_counts")
    return dict(zip(("quintic", "quartic", "cart", "cart_guarded", "polar", "seed_ticks",
                     "area_ticks", "rows"), list(out)))

class Dynamics(ctypes.Structure):
    """lcbg_dynamics: annual and terrestrial parallax, event coordinates, orbital
    motion and xallarap."""
    _fields_ = ([(k, ctypes.c_double) for k in ("piEN", "piEE", "ra", "dec", "tfix", "g1", "g2", "g3",
                                                "lom_szs", "lom_ar", "obs_lat", "obs_lon")] +
                [("xallarap_type", ctypes.c_int32), ("reserved", ctypes.c_int32)] +
                [(k, ctypes.c_double) for k in ("xi_1", "xi_2", "omega_xa", "inc_xa", "phi_xa", "piEN_xa",
                                                "piEE_xa", "period_xa", "ecc_xa", "peri_xa")])

    XALLARAP_NONE, XALLARAP_ANGULAR_VELOCITY, XALLARAP_ORBITAL_ELEMENTS = 0, 1, 2

    def __init__(self, **kw):
        kw.setdefault("lom_ar", 1.0)
        super().__init__(**kw)


ORBIT_STATIC, ORBIT_CIRCULAR, ORBIT_KEPLER = 0, 1, 2

lcbg_load_earth_ephemeris = _sig("lcbg_load_earth_ephemeris", ctypes.c_int, ctypes.c_char_p)
lcbg_dynamic_sources = _sig("lcbg_dynamic_sources", ctypes.c_int, ctypes.POINTER(Curve),
                            ctypes.POINTER(Dynamics), ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                            ctypes.c_void_p)
lcbg_batch_light_curves_dynamic = _sig("lcbg_batch_light_curves_dynamic", ctypes.c_int, ctypes.c_void_p,
                                       ctypes.POINTER(Dynamics), ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p)


lcbg_dynamic_sources_jacobian = _sig("lcbg_dynamic_sources_jacobian", ctypes.c_int, ctypes.POINTER(Curve),
                                     ctypes.POINTER(Dynamics), ctypes.c_int32, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p)
lcbg_batch_light_curves_dynamic_grad = _sig("lcbg_batch_light_curves_dynamic_grad", ctypes.c_int,
                                            ctypes.c_void_p, ctypes.POINTER(Dynamics), ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p)
lcbg_batch_light_curves_dynamic_device = _sig("lcbg_batch_light_curves_dynamic_device", ctypes.c_int,
                                              ctypes.c_void_p, ctypes.POINTER(Dynamics), ctypes.c_int32,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p)
lcbg_batch_dynamic_sources_device = _sig("lcbg_batch_dynamic_sources_device", ctypes.c_int, ctypes.c_void_p,
                                         ctypes.POINTER(Dynamics), ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
TRAJ_PARAMS = ("t0", "tE", "u0", "alpha", "s", "piEN", "piEE", "g1", "g2")
DYN_GRAD_PARAMS = ("t0", "tE", "u0", "alpha", "s", "q", "rho", "ld_c", "piEN", "piEE", "g1", "g2")


def dynamic_sources_jacobian(curve: Curve, dyn: Dynamics, orbit_mode: int, times):
    """Host trajectory and its Jacobian: ((n, 3), (n, 3, 9)) over TRAJ_PARAMS."""
    t = np.ascontiguousarray(times, dtype=np.float64)
    out = np.empty((len(t), 3))
    jac = np.empty((len(t), 3, 9))
    check(lcbg_dynamic_sources_jacobian(ctypes.byref(curve), ctypes.byref(dyn), int(orbit_mode),
                                        t.ctypes.data, len(t), out.ctypes.data, jac.ctypes.data),
          "lcbg_dynamic_sources_jacobian")
    return out, jac


def light_curves_dynamic_grad(batch: "Batch", dyn, orbit_mode: int, offsets, times):
    """(mag, method, mag_aa, dmag (n, 12) over DYN_GRAD_PARAMS), host arrays."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    t = np.ascontiguousarray(times, dtype=np.float64)
    d = (Dynamics * len(dyn))(*dyn)
    mag = np.empty(len(t))
    meth = np.empty(len(t), dtype=np.int32)
    aa = np.empty(len(t))
    dm = np.empty((len(t), 12))
    check(lcbg_batch_light_curves_dynamic_grad(batch.handle, d, int(orbit_mode), off.ctypes.data,
                                               t.ctypes.data, len(t), mag.ctypes.data, meth.ctypes.data,
                                               aa.ctypes.data, dm.ctypes.data),
          "lcbg_batch_light_curves_dynamic_grad")
    return mag, meth, aa, dm


def load_earth_ephemeris(path=None):
    """Reads the Horizons-format Earth table (None: LCBININT_EARTH_EPHEMERIS)."""
    check(lcbg_load_earth_ephemeris(None if path is None else str(path).encode()),
          "lcbg_load_earth_ephemeris")


def dynamic_sources(curve: Curve, dyn: Dynamics, orbit_mode: int, times):
    """Host trajectory: (n, 3) array of source x, y and separation s(t)."""
    t = np.ascontiguousarray(times, dtype=np.float64)
    out = np.empty((len(t), 3))
    check(lcbg_dynamic_sources(ctypes.byref(curve), ctypes.byref(dyn), int(orbit_mode),
                               t.ctypes.data, len(t), out.ctypes.data), "lcbg_dynamic_sources")
    return out


def light_curves_dynamic(batch: "Batch", dyn, orbit_mode: int, offsets, times):
    """Magnifications and methods of the batch's curves with parallax / orbital
    motion (host arrays; dyn: one Dynamics per curve)."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    t = np.ascontiguousarray(times, dtype=np.float64)
    d = (Dynamics * len(dyn))(*dyn)
    mag = np.empty(len(t))
    meth = np.empty(len(t), dtype=np.int32)
    check(lcbg_batch_light_curves_dynamic(batch.handle, d, int(orbit_mode), off.ctypes.data, t.ctypes.data,
                                          len(t), mag.ctypes.data, meth.ctypes.data),
          "lcbg_batch_light_curves_dynamic")
    return mag, meth


class Schedule(ctypes.Structure):
    """lcbg_schedule: launch schedule of the inverse-ray stage (0 = default)."""
    _fields_ = [("seed_waves_per_cu", ctypes.c_int32), ("cart_waves_per_cu", ctypes.c_int32),
                ("polar_waves_per_cu", ctypes.c_int32), ("heavy_budget", ctypes.c_int32),
                ("lead_items", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


lcbg_batch_set_schedule = _sig("lcbg_batch_set_schedule", ctypes.c_int, ctypes.c_void_p,
                               ctypes.POINTER(Schedule))
lcbg_batch_get_schedule = _sig("lcbg_batch_get_schedule", ctypes.c_int, ctypes.c_void_p,
                               ctypes.POINTER(Schedule))
lcbg_batch_autotune = _sig("lcbg_batch_autotune", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Schedule),
                           ctypes.POINTER(ctypes.c_float))


def autotune(batch: "Batch", offsets_ptr: int, times_ptr: int, n: int, mag_ptr: int, method_ptr: int,
             stream: int = 0, max_runs: int = 24):
    """Runs the autotuner on this workload; returns (Schedule, best run ms)."""
    best = Schedule()
    ms = ctypes.c_float()
    check(lcbg_batch_autotune(batch.handle, offsets_ptr, times_ptr, n, mag_ptr, method_ptr, stream,
                              max_runs, ctypes.byref(best), ctypes.byref(ms)), "lcbg_batch_autotune")
    return best, ms.value


def set_schedule(batch: "Batch", schedule: Schedule):
    check(lcbg_batch_set_schedule(batch.handle, ctypes.byref(schedule)), "lcbg_batch_set_schedule")


lcbg_batch_capture_finite_source = _sig("lcbg_batch_capture_finite_source", ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p)
lcbg_batch_launch_graph = _sig("lcbg_batch_launch_graph", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)


def capture_finite_source(batch: "Batch", offsets_ptr: int, times_ptr: int, n: int, mag_ptr: int,
                          method_ptr: int, stream: int = 0):
    """Records one finite-source run (device pointers) into the batch's HIP graph."""
    check(lcbg_batch_capture_finite_source(batch.handle, offsets_ptr, times_ptr, n, mag_ptr, method_ptr,
                                           stream), "lcbg_batch_capture_finite_source")


def launch_graph(batch: "Batch", stream: int = 0):
    """Replays the batch's captured finite-source run on `stream`."""
    check(lcbg_batch_launch_graph(batch.handle, stream), "lcbg_batch_launch_graph")


lcbg_batch_ir_kernel_ms = _sig("lcbg_batch_ir_kernel_ms", ctypes.c_int, ctypes.c_void_p,
                               ctypes.POINTER(ctypes.c_float))
IR_LAUNCHES = ("seeds", "cart_big", "polar_big", "polar_small", "cart_small", "cart_retry",
               "polar_retry")


def ir_kernel_ms(batch: Batch):
    """Per-launch durations (ms) of the inverse-ray kernels of the last
    profiled run, keyed by IR_LAUNCHES."""
    out = (ctypes.c_float * 7)()
    check(lcbg_batch_ir_kernel_ms(batch.handle, out), "lcbg_batch_ir_kernel_ms")
    return dict(zip(IR_LAUNCHES, list(out)))

lcbg_batch_finite_source_grad_device = _sig(
    "lcbg_batch_finite_source_grad_device", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_void_p)
lcbg_batch_chi2_grad_device = _sig(
    "lcbg_batch_chi2_grad_device", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
    ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)

GRAD_PARAMS = ("t0", "tE", "u0", "alpha", "s", "q", "rho", "ld_c")


def finite_source_grad(batch: Batch, offsets_ptr: int, times_ptr: int, n_epochs: int, mag_ptr: int,
                       method_ptr: int, mag_aa_ptr: int, dmag_ptr: int, stream: int = 0):
    """Values, methods, differentiated values and d mag / d GRAD_PARAMS (n x 8)."""
    check(lcbg_batch_finite_source_grad_device(batch.handle, offsets_ptr, times_ptr, n_epochs,
                                               mag_ptr, method_ptr, mag_aa_ptr, dmag_ptr, stream),
          "lcbg_batch_finite_source_grad_device")


def chi2_grad(batch: Batch, offsets_ptr: int, times_ptr: int, n_epochs: int, flux_ptr: int,
              sigma_ptr: int, fs: float, fb: float, mag_ptr: int, method_ptr: int,
              chi2_ptr: int, grad_ptr: int, mag_aa_ptr=None, dmag_ptr=None, stream: int = 0):
    """chi^2 per curve and its gradient with respect to GRAD_PARAMS (n_curves x 8)."""
    check(lcbg_batch_chi2_grad_device(batch.handle, offsets_ptr, times_ptr, n_epochs, flux_ptr,
                                      sigma_ptr, fs, fb, mag_ptr, method_ptr, mag_aa_ptr, dmag_ptr,
                                      chi2_ptr, grad_ptr, stream),
          "lcbg_batch_chi2_grad_device")


lcbg_light_curve_results = _sig(
    "lcbg_light_curve_results", ctypes.c_int, ctypes.POINTER(Curve), ctypes.POINTER(Options),
    c_double_p, ctypes.c_int64, ctypes.c_void_p, c_int64_p)
lcbg_batch_last_run_status = _sig(
    "lcbg_batch_last_run_status", ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32),
    c_int64_p)
lcbg_batch_epoch_results_device = _sig(
    "lcbg_batch_epoch_results_device", ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
    ctypes.c_void_p)


def light_curve_results(curve: Curve, times, options: Options | None = None):
    """(records, status, first failed epoch) of one curve on the host path:
    records is a structured array of RESULT_DTYPE, status the lcbg_status."""
    t = np.ascontiguousarray(times, dtype=np.float64)
    out = np.zeros(len(t), dtype=RESULT_DTYPE)
    first = ctypes.c_int64(-1)
    st = lcbg_light_curve_results(ctypes.byref(curve), ctypes.byref(options) if options is not None else None,
                                  t.ctypes.data_as(c_double_p), len(t), out.ctypes.data, ctypes.byref(first))
    return out, st, first.value


lcbg_triple_light_curve_point_source = _sig(
    "lcbg_triple_light_curve_point_source", ctypes.c_int, ctypes.POINTER(Triple), ctypes.c_void_p,
    ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p)
lcbg_triple_point_source_device = _sig(
    "lcbg_triple_point_source_device", ctypes.c_int, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)


def triple_light_curve_point_source(curve: Triple, times):
    """(magnifications, image counts, status) of one triple-lens curve, host arrays."""
    t = np.ascontiguousarray(times, dtype=np.float64)
    mag = np.empty(len(t))
    cnt = np.empty(len(t), dtype=np.int32)
    st = lcbg_triple_light_curve_point_source(ctypes.byref(curve), t.ctypes.data, len(t), mag.ctypes.data,
                                              cnt.ctypes.data)
    return mag, cnt, st


lcbg_batch_light_curves_dynamic_results = _sig(
    "lcbg_batch_light_curves_dynamic_results", ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(Dynamics),
    ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p)


def light_curves_dynamic_results(batch: "Batch", dyn, orbit_mode: int, offsets, times):
    """(records, status) of the batch's curves on the LensModel path: records is
    a structured array of RESULT_DTYPE (host), status the lcbg_status."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    t = np.ascontiguousarray(times, dtype=np.float64)
    d = (Dynamics * len(dyn))(*dyn)
    out = np.zeros(len(t), dtype=RESULT_DTYPE)
    st = lcbg_batch_light_curves_dynamic_results(batch.handle, d, int(orbit_mode), off.ctypes.data,
                                                 t.ctypes.data, len(t), out.ctypes.data)
    return out, st


def last_run_status(batch: "Batch"):
    """(status, error flags, first failed epoch or -1) of the batch's last run."""
    flags = ctypes.c_uint32()
    first = ctypes.c_int64()
    st = lcbg_batch_last_run_status(batch.handle, ctypes.byref(flags), ctypes.byref(first))
    return st, flags.value, first.value


def epoch_results(batch: "Batch", n_epochs: int, results_ptr: int, stream: int = 0):
    """Enqueues the per-epoch records of the batch's last run into a device buffer
    of n_epochs * RESULT_DTYPE.itemsize bytes."""
    check(lcbg_batch_epoch_results_device(batch.handle, n_epochs, results_ptr, stream),
          "lcbg_batch_epoch_results_device")

lcbg_batch_set_work_budget = _sig("lcbg_batch_set_work_budget", ctypes.c_int, ctypes.c_void_p, ctypes.c_int64)

lcbg_light_curve_device_noalloc = _sig(
    "lcbg_light_curve_device_noalloc", ctypes.c_int, ctypes.POINTER(Curve), ctypes.POINTER(Options),
    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
