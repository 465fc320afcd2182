"""Option and class surface of the drop-in Python API (no GPU calls).

Mirrors the property checks of tests/regression/test_vbm_consistency.py:1134-1196
in the reference: aliases, defaults and the inverse-ray grid names."""
import pytest


def test_options_defaults_and_aliases():
    import lcbinint_amd as lcbinint
    o = lcbinint.Options()
    assert (o.nbin, o.source_bins, o.caustic_bins, o.grid_ratio) == (50, 50, 1400, 4.0)
    assert o.inverse_ray_grid == "auto" and o._mode == 4
    assert o.coordinates == "vbm" and o.param_type == "vbm"
    assert o.polar_nbin is None and o.polar_grid_ratio is None
    assert (o.point_source_threshold, o.hexadecapole_threshold, o.hex_tol) == (20.0, 3.0, 0.001)
    o2 = lcbinint.Options(nbin=70, inverse_ray_grid="polar", polar_nbin=30, polar_grid_ratio=8.0,
                          hex_tol=2e-3)
    assert (o2.source_bins, o2._mode, o2.polar_source_bins, o2.polar_grid_ratio) == (70, 2, 30, 8.0)
    assert o2.adaptive_hex_threshold == 2e-3
    o2.inverse_ray_grid = "cartesian"
    assert o2._mode == 1
    with pytest.raises(ValueError):
        lcbinint.Options(inverse_ray_grid="hexagonal")
    with pytest.raises(ValueError):
        lcbinint.Options(coordinates="galactic")
    assert lcbinint.Options(coordinates="lcbinint").coordinates == "lcbinint"


def test_limb_darkening_and_light_curve_properties():
    import lcbinint_amd as lcbinint
    ld = lcbinint.LimbDarkening.linear(0.4)
    assert (ld.c, ld.d) == (0.4, 0.0)
    assert (lcbinint.LimbDarkening.quadratic(0.3, 0.2).d) == 0.2
    lc = lcbinint.LightCurve(limb_darkening=ld)
    assert lc.lens == "binary_lens" and lc.source == "single" and not lc.parallax
    with pytest.raises(ValueError):
        lcbinint.LightCurve(lens="quadruple")
