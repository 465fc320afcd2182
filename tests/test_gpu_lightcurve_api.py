"""The Python drop-in (LightCurve / Options / LimbDarkening) on the GPU,
against the reference compiled from source, through the dict and keyword paths."""
import numpy as np
import pytest

from oracle.bindings import default_settings

pytestmark = pytest.mark.gpu


def test_readme_light_curve(gpu, ref):
    import lcbinint  # the reference's module name
    t = np.linspace(-0.5, 0.5, 200)
    params = {"t0": 0.0, "tE": 1.0, "u0": -0.01, "alpha": 0.5, "s": 0.95, "q": 1e-2, "rho": 5e-3}
    lc = lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=50),
                             limb_darkening=lcbinint.LimbDarkening.linear(0.5))
    got = lc(t, params)
    assert got.dtype == np.float64 and got.shape == (200,)
    exp, meth, st = ref.light_curve(default_settings(limb_darkening_c=0.5), t, 0.0, 1.0, -0.01, 0.5,
                                    0.95, 1e-2, 5e-3)
    ps = meth <= 1
    assert np.array_equal(got[ps], exp[ps])
    assert np.max(np.abs(got - exp) / exp) <= 1e-9
    kw = lc(t, t0=0.0, tE=1.0, u0=-0.01, alpha=0.5, s=0.95, q=1e-2, rho=5e-3)
    assert np.array_equal(kw, got)


def test_point_source_curve_and_errors(gpu, ref):
    import lcbinint_amd as lcbinint
    t = np.linspace(-1.0, 1.0, 101)
    lc = lcbinint.LightCurve()
    got = lc(t, {"u0": 0.1, "alpha": 1.0, "s": 1.2, "q": 0.3})
    exp, _, _ = ref.light_curve(default_settings(), t, 0.0, 1.0, 0.1, 1.0, 1.2, 0.3, 0.0)
    assert np.array_equal(got, exp)
    with pytest.raises(ValueError):
        lc(np.zeros((3, 3)), {})
    # a finite source on a triple lens is outside this library (its point source is inside it)
    with pytest.raises(RuntimeError, match="unsupported"):
        lcbinint.LightCurve(lens="triple_lens")(t, {"u0": 0.1, "q2": 0.1, "sep2": 1.0, "ang": 0.5, "rho": 1e-3})
    with pytest.raises(RuntimeError, match="numerical error"):
        lc(t, {"tE": 0.0})


def test_info_matches_reference_info(gpu, ref):
    """LightCurve.info against evaluate_binary_light_curve_info
    (python/lcbinint_pybind.cpp:1176-1253), static VBM curve, with a relative
    tolerance that leaves some Cartesian epochs converged and some not."""
    import lcbinint
    t = np.linspace(-0.5, 0.5, 200)
    params = {"t0": 0.0, "tE": 1.0, "u0": -0.01, "alpha": 0.5, "s": 0.95, "q": 1e-2, "rho": 5e-3}
    lc = lcbinint.LightCurve(options=lcbinint.Options(coordinates="vbm", nbin=50, reltol=2e-3),
                             limb_darkening=lcbinint.LimbDarkening.linear(0.5))
    info = lc.info(t, params)
    exp, st = ref.light_curve_info(default_settings(limb_darkening_c=0.5), t, 0.0, 1.0, -0.01, 0.5,
                                   0.95, 1e-2, 5e-3, 0.0, 2e-3)
    assert st == 0
    meth = np.array(info.finite_source_methods)
    assert np.array_equal(meth, exp["method"])
    assert info.finite_source_method_names[int(np.argmax(meth == 2))] == "inverse_ray_cartesian"
    assert np.array_equal(np.array(info.image_counts), exp["image_count"])
    assert np.array_equal(np.array(info.point_source_magnifications), exp["point_source_magnification"])
    assert np.array_equal(np.array(info.source_x), exp["source_x"])
    assert np.array_equal(np.array(info.source_y), exp["source_y"])
    got = np.array(info.magnifications)
    assert np.max(np.abs(got - exp["magnification"]) / exp["magnification"]) <= 1e-9
    err = np.array(info.finite_source_error_estimates)
    cart = meth == 2
    assert np.array_equal(err[~cart], exp["error_estimate"][~cart])
    assert np.max(np.abs(err[cart] - exp["error_estimate"][cart]) / exp["error_estimate"][cart]) <= 1e-12
    conv = np.array(info.finite_source_converged, dtype=np.int32)
    assert np.array_equal(conv, exp["converged"])
    assert info.unconverged_indices == [int(i) for i in np.nonzero(conv == 0)[0]]
    assert info.all_converged == bool(conv.all())
    assert np.array_equal(np.array(lc(t, params)), got)
