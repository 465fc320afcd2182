"""ctypes bindings of the two CPU checkers.

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.
"""
import ctypes

import numpy as np

D = ctypes.c_double
P = ctypes.POINTER
I32 = ctypes.c_int


def _f(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class RefSettings(ctypes.Structure):
    """Layout of RefSettings in oracle/ref_shim.cpp (= FiniteSourceSettings fields)."""
    _fields_ = [
        ("source_bins", ctypes.c_int), ("caustic_bins", ctypes.c_int),
        ("grid_ratio", D), ("polar_source_bins", ctypes.c_int), ("polar_grid_ratio", D),
        ("finite_mode", ctypes.c_int), ("kinji_threshold", D), ("hex_threshold", D),
        ("adaptive_hex_threshold", D), ("limb_darkening_c", D), ("limb_darkening_d", D),
        ("adaptive_source_bins", ctypes.c_int), ("max_source_bins", ctypes.c_int),
    ]


def default_settings(**kw):
    s = RefSettings(50, 1400, 4.0, 0, 0.0, 4, 20.0, 3.0, 0.001, 0.0, 0.0, 0, 400)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


class LcbiParams(ctypes.Structure):
    """lcbi_params of the reference's C API (include/lcbinint/lcbinint.h:20-77)."""
    _fields_ = ([(k, D) for k in ("t0", "tE", "umin", "q", "sep", "theta", "rho", "omega", "piEN", "piEE",
                                  "piEN_xa", "piEE_xa", "ra_xa", "dec_xa", "period_xa", "ecc_xa", "peri_xa",
                                  "v_sep", "q2", "sep2", "ang", "ra", "dec", "earth_axis", "tfix", "obs_lat",
                                  "obs_lon", "limb_darkening_c", "limb_darkening_d")] +
                [("orbital_motion_mode", ctypes.c_int)] +
                [(k, D) for k in ("g1", "g2", "g3", "lom_szs", "lom_ar", "xi_1", "xi_2", "omega_xa", "inc_xa",
                                  "phi_xa")])


class LcbiOptions(ctypes.Structure):
    """lcbi_options of the reference's C API (include/lcbinint/lcbinint.h:85-104)."""
    _fields_ = ([(k, ctypes.c_int) for k in ("parallax_mode", "orbit_pair", "center_of_mass", "caustic_bins",
                                             "source_bins", "mode", "vbm_compatible", "xallarap_param_type")] +
                [("grid_ratio", D), ("polar_source_bins", ctypes.c_int), ("polar_grid_ratio", D),
                 ("point_source_threshold", D), ("hexadecapole_threshold", D), ("adaptive_hex_threshold", D),
                 ("adaptive_source_bins", ctypes.c_int), ("max_source_bins", ctypes.c_int),
                 ("finite_source_tol", D), ("finite_source_reltol", D)])


class LcbiResult(ctypes.Structure):
    """lcbi_result of the reference's C API (include/lcbinint/lcbinint.h:106-116)."""
    _fields_ = [("magnification", D), ("point_source_magnification", D), ("finite_source_magnification", D),
                ("source_x", D), ("source_y", D), ("image_count", ctypes.c_int),
                ("finite_source_error_estimate", D), ("finite_source_refinement_level", ctypes.c_int),
                ("finite_source_converged", ctypes.c_int)]


class Oracle:
    def __init__(self, path):
        L = ctypes.CDLL(str(path))
        self.lib = L
        V = ctypes.c_void_p
        self._mag0 = _f(L, "lcbo_binary_mag0", D, D, D, D, D, P(I32))
        self._mag0_many = _f(L, "lcbo_binary_mag0_many", None, V, V, V, V, ctypes.c_long, V, V)
        self._mag0_der = _f(L, "lcbo_binary_mag0_with_derivatives", D, D, D, D, D, P(I32), P(D))
        self._cands = _f(L, "lcbo_binary_image_candidates", I32, D, D, D, D, V)
        self._roots = _f(L, "lcbo_cmplx_roots_gen", None, V, I32, I32, I32, V)
        self._lc_ps = _f(L, "lcbo_light_curve_ps", None, V, ctypes.c_long, D, D, D, D, D, D, V)
        self._ld_table = _f(L, "lcbo_ld_table", None, D, D, V)
        self._ld_lookup = _f(L, "lcbo_ld_lookup", D, V, D)
        self._hex = _f(L, "lcbo_hexadecapole", D, D, D, D, D, D, D, D, D, P(D))
        self._crit = _f(L, "lcbo_critical_curve", None, D, D, D, D, V, V)
        self.flop_count = _f(L, "lcbo_flop_count", D)
        self.flop_reset = _f(L, "lcbo_flop_reset", None)

    def binary_mag0(self, s, q, x, y):
        n = I32()
        a = self._mag0(s, q, x, y, ctypes.byref(n))
        return a, n.value

    def binary_mag0_many(self, s, q, x, y):
        s, q, x, y = (np.ascontiguousarray(v, dtype=np.float64) for v in (s, q, x, y))
        out = np.empty(len(s))
        ic = np.empty(len(s), dtype=np.int32)
        self._mag0_many(ptr(s), ptr(q), ptr(x), ptr(y), len(s), ptr(out), ptr(ic))
        return out, ic

    def binary_mag0_with_derivatives(self, s, q, x, y):
        n = I32()
        ind = D()
        a = self._mag0_der(s, q, x, y, ctypes.byref(n), ctypes.byref(ind))
        return a, n.value, ind.value

    def image_candidates(self, s, q, x, y):
        out = np.zeros((5, 5))
        k = self._cands(s, q, x, y, ptr(out))
        return out[:k]

    def roots_gen(self, coeffs, polish=True, start=None):
        coeffs = np.asarray(coeffs, dtype=np.complex128)
        deg = len(coeffs) - 1
        c = np.ascontiguousarray(np.stack([coeffs.real, coeffs.imag], axis=1).ravel())
        r = np.zeros(2 * deg) if start is None else np.ascontiguousarray(
            np.stack([np.real(start), np.imag(start)], axis=1).ravel())
        self._roots(ptr(c), deg, int(polish), int(start is not None), ptr(r))
        return r[0::2] + 1j * r[1::2]

    def light_curve_ps(self, times, t0, tE, u0, alpha, s, q):
        t = np.ascontiguousarray(times, dtype=np.float64)
        out = np.empty(len(t))
        self._lc_ps(ptr(t), len(t), t0, tE, u0, alpha, s, q, ptr(out))
        return out

    def ld_table(self, c, d=0.0):
        t = np.empty(20001)
        self._ld_table(c, d, ptr(t))
        return t

    def ld_lookup(self, table, r2):
        return self._ld_lookup(ptr(table), r2)

    def critical_curve(self, s, q, cos_phi, sin_phi):
        ca = np.zeros(8)
        cr = np.zeros(8)
        self._crit(s, q, cos_phi, sin_phi, ptr(ca), ptr(cr))
        return ca.reshape(4, 2), cr.reshape(4, 2)

    def hexadecapole(self, s, q, x, y, rho, c, d, a0):
        e = D()
        m = self._hex(s, q, x, y, rho, c, d, a0, ctypes.byref(e))
        return m, e.value


class Ref:
    def __init__(self, path):
        L = ctypes.CDLL(str(path))
        self.lib = L
        V = ctypes.c_void_p
        self._mag0 = _f(L, "ref_binary_mag0", D, D, D, D, D, P(I32))
        self._mag0_many = _f(L, "ref_binary_mag0_many", None, V, V, V, V, ctypes.c_long, V, V)
        self._mag0_der = _f(L, "ref_binary_mag0_with_derivatives", D, D, D, D, D, P(I32), P(D))
        self._cands = _f(L, "ref_binary_image_candidates", I32, D, D, D, D, V)
        self._roots = _f(L, "ref_cmplx_roots_gen", None, V, I32, I32, I32, V)
        self._poly = _f(L, "ref_poly_solve", I32, V, I32, V)
        self._lens_eq = _f(L, "ref_binary_lens_equation", None, D, D, D, D, V)
        self._fs = _f(L, "ref_fs_binary_mag", D, P(RefSettings), D, D, D, D, D, D,
                      P(I32), P(D), P(I32))
        self._lc = _f(L, "ref_light_curve_static", I32, P(RefSettings), V, ctypes.c_long,
                      D, D, D, D, D, D, D, V, V)
        self._lci = _f(L, "ref_light_curve_static_info", I32, P(RefSettings), D, D, V, ctypes.c_long,
                       D, D, D, D, D, D, D, V)
        self._lcd = _f(L, "ref_light_curve_dynamic", I32, P(RefSettings), V, ctypes.c_long, V, I32,
                       V, V, V)
        self._ir = _f(L, "ref_binary_inverse_ray", D, P(RefSettings), D, D, D, D, D, P(I32))
        self._caus = _f(L, "ref_caustic_points_at_phase", I32, D, D, D, V)
        self._crits = _f(L, "ref_critical_sources_at_phase", I32, D, D, D, V)
        self._critp = _f(L, "ref_critical_points_at_phase", I32, D, D, D, V)
        self._geo_src = _f(L, "ref_source_for_geometry", I32, V, I32, V, ctypes.c_long, V)
        self._seeds = _f(L, "ref_image_seeds", I32, P(RefSettings), D, D, D, D, D, V, I32)
        self._cart = _f(L, "ref_cartesian_area", D, P(RefSettings), D, D, D, D, D, V, I32, V)
        self._polar = _f(L, "ref_polar_area", D, P(RefSettings), D, D, D, D, D, V, I32)
        self._hex = _f(L, "ref_hexadecapole", D, P(RefSettings), D, D, D, D, D, D, P(D))

    def binary_mag0(self, s, q, x, y):
        n = I32()
        a = self._mag0(s, q, x, y, ctypes.byref(n))
        return a, n.value

    def binary_mag0_many(self, s, q, x, y):
        s, q, x, y = (np.ascontiguousarray(v, dtype=np.float64) for v in (s, q, x, y))
        out = np.empty(len(s))
        ic = np.empty(len(s), dtype=np.int32)
        self._mag0_many(ptr(s), ptr(q), ptr(x), ptr(y), len(s), ptr(out), ptr(ic))
        return out, ic

    def binary_mag0_with_derivatives(self, s, q, x, y):
        n = I32()
        ind = D()
        a = self._mag0_der(s, q, x, y, ctypes.byref(n), ctypes.byref(ind))
        return a, n.value, ind.value

    def image_candidates(self, s, q, x, y):
        out = np.zeros((5, 5))
        k = self._cands(s, q, x, y, ptr(out))
        return out[:k]

    def roots_gen(self, coeffs, polish=True, start=None):
        coeffs = np.asarray(coeffs, dtype=np.complex128)
        deg = len(coeffs) - 1
        c = np.ascontiguousarray(np.stack([coeffs.real, coeffs.imag], axis=1).ravel())
        r = np.zeros(2 * deg) if start is None else np.ascontiguousarray(
            np.stack([np.real(start), np.imag(start)], axis=1).ravel())
        self._roots(ptr(c), deg, int(polish), int(start is not None), ptr(r))
        return r[0::2] + 1j * r[1::2]

    def poly_solve(self, coeffs):
        coeffs = np.asarray(coeffs, dtype=np.complex128)
        c = np.ascontiguousarray(np.stack([coeffs.real, coeffs.imag], axis=1).ravel())
        r = np.zeros(2 * len(coeffs))
        k = self._poly(ptr(c), len(coeffs), ptr(r))
        return (r[0::2] + 1j * r[1::2])[:k]

    def lens_equation(self, s, q, zx, zy):
        out = np.zeros(2)
        self._lens_eq(s, q, zx, zy, ptr(out))
        return out

    def fs_binary_mag(self, settings, s, q, x, y, rho, a_ps):
        m = I32()
        e = D()
        c = I32()
        a = self._fs(ctypes.byref(settings), s, q, x, y, rho, a_ps, ctypes.byref(m),
                     ctypes.byref(e), ctypes.byref(c))
        return a, m.value, e.value, c.value

    def light_curve(self, settings, times, t0, tE, u0, alpha, s, q, rho):
        t = np.ascontiguousarray(times, dtype=np.float64)
        out = np.empty(len(t))
        meth = np.empty(len(t), dtype=np.int32)
        st = self._lc(ctypes.byref(settings), ptr(t), len(t), t0, tE, u0, alpha, s, q, rho,
                      ptr(out), ptr(meth))
        return out, meth, st

    INFO_FIELDS = ("magnification", "point_source_magnification", "finite_source_magnification",
                   "source_x", "source_y", "image_count", "method", "error_estimate",
                   "refinement_level", "converged")

    def light_curve_info(self, settings, times, t0, tE, u0, alpha, s, q, rho, tol=0.0, reltol=0.0):
        """Per-epoch records of the static VBM info loop: dict of arrays keyed
        by INFO_FIELDS, and the status (0, or 2 if an epoch was not finite)."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        rec = np.full((len(t), 10), np.nan)
        st = self._lci(ctypes.byref(settings), tol, reltol, ptr(t), len(t), t0, tE, u0, alpha, s, q, rho,
                       ptr(rec))
        out = {k: rec[:, i] for i, k in enumerate(self.INFO_FIELDS)}
        for k in ("image_count", "method", "refinement_level", "converged"):
            out[k] = np.nan_to_num(out[k], nan=-1).astype(np.int32)
        return out, st

    DYNAMIC_KEYS = ("t0", "tE", "u0", "alpha", "s", "q", "rho", "piEN", "piEE", "ra", "dec", "tfix",
                    "g1", "g2", "g3", "lom_szs", "lom_ar")

    def lens_model_at(self, settings, s, q, x, y, rho):
        """(magnification, method) of the LensModel path's finite-source step at
        source (x, y) in the binary frame and separation s."""
        fn = self.lib.ref_lens_model_at
        fn.restype = D
        fn.argtypes = [P(RefSettings), D, D, D, D, D, P(I32)]
        m = I32()
        a = fn(ctypes.byref(settings), s, q, x, y, rho, ctypes.byref(m))
        return a, m.value

    def light_curve_dynamic(self, settings, times, params, orbit_mode=0):
        """LensModel path (parallax and/or orbital motion); params: dict with
        DYNAMIC_KEYS (missing keys: 0, lom_ar 1).  Returns (mag, method,
        source x/y and s(t) as an (n, 3) array, status).  The Earth table is the
        one LCBININT_EARTH_EPHEMERIS names when the first parallax epoch runs."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        p = np.array([params.get(k, 1.0 if k in ("lom_ar", "tE") else 0.0) for k in self.DYNAMIC_KEYS],
                     dtype=np.float64)
        out = np.empty(len(t))
        meth = np.empty(len(t), dtype=np.int32)
        src = np.empty((len(t), 3))
        st = self._lcd(ctypes.byref(settings), ptr(t), len(t), ptr(p), int(orbit_mode), ptr(out),
                       ptr(meth), ptr(src))
        return out, meth, src, st

    def caustic_points(self, s, q, phase):
        out = np.zeros(8)
        k = self._caus(s, q, phase, ptr(out))
        return out[:2 * k].reshape(-1, 2)

    def critical_points(self, s, q, phase):
        """The four roots of the critical-curve quartic at one phase, (k, 2)."""
        out = np.zeros(8)
        k = self._critp(s, q, phase, ptr(out))
        return out[:2 * k].reshape(-1, 2)

    def source_for_geometry(self, times, params, orbit_mode=0):
        """The geometry helpers' source positions, (n, 2); params as
        light_curve_dynamic takes them."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        p = np.array([params.get(k, 1.0 if k in ("lom_ar", "tE") else 0.0) for k in self.DYNAMIC_KEYS],
                     dtype=np.float64)
        out = np.empty((len(t), 2))
        self._geo_src(ptr(p), int(orbit_mode), ptr(t), len(t), ptr(out))
        return out

    def critical_sources(self, s, q, phase):
        out = np.zeros(8)
        k = self._crits(s, q, phase, ptr(out))
        return out[:2 * k].reshape(-1, 2)

    def image_seeds(self, settings, s, q, x, y, rho, cap=4096):
        out = np.zeros(2 * cap)
        k = self._seeds(ctypes.byref(settings), s, q, x, y, rho, ptr(out), cap)
        return out[:2 * min(k, cap)].reshape(-1, 2)

    def cartesian_area(self, settings, s, q, x, y, rho, seeds):
        sd = np.ascontiguousarray(seeds, dtype=np.float64).ravel()
        diag = np.zeros(8)
        a = self._cart(ctypes.byref(settings), s, q, x, y, rho, ptr(sd), len(sd) // 2, ptr(diag))
        return a, diag

    def polar_area(self, settings, s, q, x, y, rho, seeds):
        sd = np.ascontiguousarray(seeds, dtype=np.float64).ravel()
        return self._polar(ctypes.byref(settings), s, q, x, y, rho, ptr(sd), len(sd) // 2)

    def hexadecapole(self, settings, s, q, x, y, rho, a0):
        e = D()
        m = self._hex(ctypes.byref(settings), s, q, x, y, rho, a0, ctypes.byref(e))
        return m, e.value

    def inverse_ray(self, settings, s, q, x, y, rho):
        m = I32()
        a = self._ir(ctypes.byref(settings), s, q, x, y, rho, ctypes.byref(m))
        return a, m.value


class RefCApi:
    """The reference's own C API (src/lcbinint/lcbinint.cpp): lcbi_magnification_array
    runs LensModel::magnification over the epochs, the LensModel branch of the
    Python loop (python/lcbinint_pybind.cpp:1159-1173)."""

    def __init__(self, path):
        L = ctypes.CDLL(str(path))
        self._dp = L.lcbi_default_params
        self._dp.restype = LcbiParams
        self._do = L.lcbi_default_options
        self._do.restype = LcbiOptions
        self._arr = _f(L, "lcbi_magnification_array", ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                       ctypes.POINTER(LcbiParams), ctypes.POINTER(LcbiOptions), ctypes.c_void_p)

    def params(self, **kw):
        p = self._dp()
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def options(self, **kw):
        o = self._do()
        o.vbm_compatible = 1
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def curve(self, times, params, options):
        """Records of every epoch (dict of arrays of the lcbi_result fields) and
        the status (0, or that of the first failing epoch, where it stops)."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        res = (LcbiResult * len(t))()
        st = self._arr(t.ctypes.data_as(ctypes.c_void_p), len(t), ctypes.byref(params), ctypes.byref(options),
                       ctypes.cast(res, ctypes.c_void_p))
        out = {name: np.array([getattr(r, name) for r in res]) for name, _ in LcbiResult._fields_}
        return out, st


def load_oracle(path):
    return Oracle(path)


def load_ref(path):
    return Ref(path)


def ref_light_curves_worker(args):
    """Process-pool worker: the reference's static light curve of each
    (index, curve dict, nbin, times) item with the default settings at that
    nbin and the curve's linear limb darkening; returns (index, mag, method,
    status) per item.  args = (library path, items)."""
    lib, items = args
    ref = load_ref(lib)
    out = []
    for c, cv, nb, t in items:
        st = default_settings(source_bins=int(nb), limb_darkening_c=cv["ld_c"])
        m, k, status = ref.light_curve(st, t, cv["t0"], cv["tE"], cv["u0"], cv["alpha"], cv["s"],
                                       cv["q"], cv["rho"])
        out.append((c, m, k, status))
    return out
