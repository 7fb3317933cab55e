"""Host-side parts of the catalog layer (no device needed): argument checks,
the PSF second-moment FWHM, the C layout of a catalog row."""
import numpy as np
import pytest

from tests._catalog_fields import stamp


def test_measure_psf_fwhm_closed_form():
    from delta_amd import catalog

    for sigma in (1.2, 2.0, 3.1):
        f = catalog.measure_psf_fwhm(stamp(41, sigma))
        assert f == pytest.approx(2.354820045030949 * sigma, rel=1e-4)
    with pytest.raises(ValueError):
        catalog.measure_psf_fwhm(-np.ones((5, 5)))


def test_argument_checks():
    from delta_amd import catalog

    s = np.zeros((10, 10), np.float32)
    with pytest.raises(ValueError):
        catalog.build_catalog(s, s)  # neither psf nor expected_fwhm
    with pytest.raises(ValueError):
        catalog.build_catalog(s, s, psf=stamp(5, 1.0), expected_fwhm=3.0)
    with pytest.raises(ValueError):
        catalog.build_catalog(s, s, expected_fwhm=3.0, fit_psf_shape=True)
    with pytest.raises(ValueError):
        catalog.build_catalog(s, s, psf=stamp(5, 1.0), fit_psf_shape=True)
    with pytest.raises(ValueError):
        catalog.build_catalog(s, s, expected_fwhm=3.0, polarity="up")


def test_row_layout_matches_the_header():
    import ctypes

    from delta_amd import _core

    assert _core.CATALOG_ENTRY.itemsize == 104
    assert _core.CATALOG_ENTRY.fields["peak_x"][1] == 80
    assert _core.CATALOG_ENTRY.fields["fwhm_consistent"][1] == 96
    assert ctypes.sizeof(_core.CatalogParams) == 80
