"""MI355X (gfx950) binary-lens finite-source light curves.

Drop-in for the reference's Python API on its finite-source hot path::

    import lcbinint_amd as lcbinint
    lc = lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=50),
                             limb_darkening=lcbinint.LimbDarkening.linear(0.5))
    A = lc(times, {"t0": 0.0, "tE": 1.0, "u0": -0.01, "alpha": 0.5,
                   "s": 0.95, "q": 1e-2, "rho": 5e-3})

The evaluation runs in the HIP library (liblcbinint_amd.so).  If the extension
has not been built, every API name raises ImportError when it is used; the
build module stays importable (``python -m lcbinint_amd.build``).
"""
_API = ("LightCurve", "LightCurveFunc", "LightCurveInfo", "Options", "LimbDarkening",
        "OrbitalMotionMode", "FiniteSourceMethod", "EventCoordinates", "GeometryCurve", "GeometryBranches")

# torch first, so that the extension binds to torch's HIP runtime (one HIP
# runtime per process; see capi.py)
try:
    import torch  # noqa: F401
except ImportError:
    pass

try:
    from ._lcbinint import (  # noqa: F401
        EventCoordinates,
        FiniteSourceMethod,
        GeometryBranches,
        GeometryCurve,
        LightCurve,
        LightCurveFunc,
        LightCurveInfo,
        LimbDarkening,
        OrbitalMotionMode,
        Options,
    )
    _IMPORT_ERROR = None
except ImportError as exc:  # the extension is not built
    _IMPORT_ERROR = exc

    def __getattr__(name):
        if name in _API:
            raise ImportError(
                f"lcbinint_amd.{name} needs the built extension "
                f"(python -m lcbinint_amd.build): {_IMPORT_ERROR}")
        raise AttributeError(name)

__all__ = list(_API)
