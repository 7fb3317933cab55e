"""Source catalog from the score image, the counterpart of the reference's
python/delta/catalog.py:19-233.

Coerces the inputs, resolves the expected PSF FWHM (second moments of a PSF
stamp, or a given value), calls ``_core.build_catalog`` (the device catalog)
once per polarity, and packs the columns into the reference's structured
dtype, with a ``sign`` column for ``polarity="both"``.
"""
from __future__ import annotations

import numpy as np

from . import _core

FWHM_TO_SIGMA = 1.0 / 2.354820045030949

CATALOG_DTYPE = np.dtype([
    ("x", np.float64), ("y", np.float64), ("peak_x", np.int32), ("peak_y", np.int32),
    ("peak_snr", np.float64), ("n_pix", np.int32), ("flux", np.float64),
    ("expected_n_pix", np.float64), ("fwhm_ratio", np.float64), ("fwhm_consistent", np.bool_),
    ("mask_flags", np.uint8), ("is_dipole", np.bool_), ("psf_chi2", np.float64),
    ("psf_chi2_dof", np.int32), ("psf_amplitude", np.float64),
    ("psf_amplitude_err", np.float64), ("shape_consistent", np.bool_),
    ("bright_star_dist", np.float64), ("near_bright_star", np.bool_), ("quality", np.uint8)])

QUALITY_FWHM_INCONSISTENT = 1 << 0
QUALITY_DIPOLE = 1 << 1
QUALITY_MASKED = 1 << 2
QUALITY_SHAPE_INCONSISTENT = 1 << 3
QUALITY_NEAR_BRIGHT_STAR = 1 << 4


def measure_psf_fwhm(psf) -> float:
    """Second-moment FWHM (px) of a PSF stamp's positive part
    (reference ``_measure_psf_fwhm``, catalog.py:61-76)."""
    psf = np.asarray(psf, dtype=np.float64)
    h, w = psf.shape
    ys, xs = np.mgrid[0:h, 0:w]
    weight = np.clip(psf, 0.0, None)
    total = weight.sum()
    if total <= 0:
        raise ValueError("PSF stamp has no positive flux")
    mx = (weight * xs).sum() / total
    my = (weight * ys).sum() / total
    varx = (weight * (xs - mx) ** 2).sum() / total
    vary = (weight * (ys - my) ** 2).sum() / total
    return float(np.sqrt(0.5 * (varx + vary)) / FWHM_TO_SIGMA)


def _pack(raw):
    n = raw["x"].shape[0]
    cat = np.empty(n, dtype=CATALOG_DTYPE)
    for name in CATALOG_DTYPE.names:
        cat[name] = raw[name]
    return cat


def build_catalog(score, difference, *, mask=None, psf=None, expected_fwhm=None,
                  threshold_sigma=5.0, threshold_sigma_dipole=3.0, fwhm_tolerance_lo=0.3,
                  fwhm_tolerance_hi=3.0, aperture_radius=0, exclude_bad_pixels=True,
                  fit_psf_shape=False, shape_chi2_tolerance=3.0, variance=None, bright_x=None,
                  bright_y=None, bright_star_radius=0.0, max_shape_fit=5000,
                  polarity="positive", as_table=False):
    """Catalog of the connected components of ``score`` (reference
    ``build_catalog``, catalog.py:79-233): exactly one of ``psf`` or
    ``expected_fwhm``; ``polarity`` "positive", "negative" or "both"."""
    if (psf is None) == (expected_fwhm is None):
        raise ValueError("supply exactly one of `psf` or `expected_fwhm`")
    if fit_psf_shape and psf is None:
        raise ValueError("fit_psf_shape=True needs `psf` (not `expected_fwhm`)")
    if fit_psf_shape and variance is None:
        raise ValueError("fit_psf_shape=True needs `variance`")
    if polarity not in ("positive", "negative", "both"):
        raise ValueError('polarity must be one of "positive", "negative", "both"')
    fwhm = measure_psf_fwhm(psf) if psf is not None else float(expected_fwhm or 0.0)
    psf_arr = np.ascontiguousarray(psf, dtype=np.float32) if psf is not None else None

    def run(return_negative):
        return _core.build_catalog(
            score, difference, mask=mask, threshold_sigma=threshold_sigma,
            threshold_sigma_dipole=threshold_sigma_dipole, expected_fwhm=fwhm,
            fwhm_tolerance_lo=fwhm_tolerance_lo, fwhm_tolerance_hi=fwhm_tolerance_hi,
            aperture_radius=aperture_radius, exclude_bad_pixels=exclude_bad_pixels,
            return_negative=return_negative, fit_psf_shape=fit_psf_shape,
            shape_chi2_tolerance=shape_chi2_tolerance, variance=variance,
            psf=psf_arr if fit_psf_shape else None, bright_star_radius=bright_star_radius,
            max_shape_fit=max_shape_fit, bright_x=bright_x, bright_y=bright_y)

    if polarity == "positive":
        cat = _pack(run(False))
    elif polarity == "negative":
        cat = _pack(run(True))
    else:
        pos, neg = _pack(run(False)), _pack(run(True))
        both = np.dtype([*CATALOG_DTYPE.descr, ("sign", np.int8)])
        cat = np.empty(len(pos) + len(neg), dtype=both)
        for name in CATALOG_DTYPE.names:
            cat[name] = np.concatenate([pos[name], neg[name]])
        cat["sign"] = np.concatenate([np.ones(len(pos), np.int8), -np.ones(len(neg), np.int8)])
    if as_table:
        from astropy.table import Table  # optional dependency, as in the reference

        return Table(cat)
    return cat
