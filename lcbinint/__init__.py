"""The reference's Python module name, ``lcbinint``, for this library.

The reference builds one extension module named ``lcbinint``
(python/lcbinint_pybind.cpp:2770); code written against it imports that
name::

    import lcbinint
    lc = lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=50),
                             limb_darkening=lcbinint.LimbDarkening.linear(0.5))
    A = lc(times, {"t0": 0.0, "tE": 1.0, "u0": -0.01, "alpha": 0.5,
                   "s": 0.95, "q": 1e-2, "rho": 5e-3})

Here the classes come from the HIP library's extension (lcbinint_amd), with
the same names, arguments, defaults and errors on the finite-source binary-lens
path.  Names of the reference outside that path are not provided.
"""
from lcbinint_amd import *  # noqa: F401,F403
from lcbinint_amd import __all__  # noqa: F401
