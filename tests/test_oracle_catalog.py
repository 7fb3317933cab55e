"""Pins of the catalog-summary oracle (src/catalog.cpp:262-697).

* bit-exact against the reference's own catalog.cpp compiled into oracle/_ref
  and run on one thread (each component's pixels then reach the summary in
  raster order; on more threads the reference's own sums are reordered run to
  run) on score images with blobs of both signs, noise, plateaus, masks with
  exclusion on and off, frame-edge blobs, the negative catalog, the PSF-shape
  fit, bright-star distances and the max_shape_fit guard;
* the hand cases of the reference's tests/test_catalog.py, restated;
* closed forms: centroid and peak of a symmetric blob, n_pix of a block, the
  expected-pixel formula, aperture sums of a constant difference.
"""
import numpy as np
import pytest

import oracle
from tests._catalog_fields import CASES, blob as _blob, case_args as _case_args, stamp as _stamp

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
_BAD, _COSMIC = 1, 4


def _same(a, b):
    for k in oracle.CATALOG_COLUMNS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)




@needs_ref
@pytest.mark.parametrize("case", range(len(CASES)))
def test_catalog_matches_reference_build(case):
    s, d, kw = _case_args(case)
    a = oracle.build_catalog(s, d, **kw)
    b = oracle.ref_build_catalog(s, d, **kw)
    assert len(a["x"]) > 3
    _same(a, b)


@needs_ref
def test_catalog_enrichment_discriminates():
    """The PSF fit and bright-star columns are populated (so the bit-exact
    comparison above covers their arithmetic) and differ from the defaults."""
    s, d, kw = _case_args(6)
    a = oracle.build_catalog(s, d, **kw)
    assert (a["psf_chi2_dof"] > 0).sum() > 5
    assert np.isfinite(a["bright_star_dist"]).all()
    assert a["near_bright_star"].any() and not a["near_bright_star"].all()
    assert (a["shape_consistent"] == 0).any()
    s, d, kw = _case_args(8)
    a = oracle.build_catalog(s, d, **kw)
    assert len(a["x"]) > 3 and (a["psf_chi2_dof"] == 0).all() and np.isinf(a["bright_star_dist"]).all()


# --------------------------------------------------------------------------- reference cases
def test_clean_blob_spike_and_blend():
    score = _blob((80, 80), 40, 40, 8.0, 2.0)
    cat = oracle.build_catalog(score, score.copy(), expected_fwhm=2.0 * 2.354820045030949)
    assert len(cat["x"]) == 1 and cat["fwhm_consistent"][0] == 1
    assert cat["x"][0] == pytest.approx(40.0, abs=1e-9) and cat["y"][0] == pytest.approx(40.0, abs=1e-9)
    assert cat["peak_x"][0] == 40 and cat["peak_snr"][0] == np.float32(8.0)

    score = np.zeros((60, 60), np.float32)
    score[30, 30] = 8.0
    cat = oracle.build_catalog(score, score.copy(), expected_fwhm=4.0)
    assert cat["n_pix"][0] == 1 and cat["fwhm_consistent"][0] == 0 and cat["quality"][0] & 1
    sigma = 4.0 / 2.354820045030949
    assert cat["expected_n_pix"][0] == pytest.approx(np.pi * sigma ** 2 * np.log((8 / 5) ** 2))

    score = np.zeros((60, 60), np.float32)
    score[20:35, 20:35] = 6.0
    cat = oracle.build_catalog(score, score.copy(), expected_fwhm=2.0)
    assert cat["n_pix"][0] == 225 and cat["fwhm_ratio"][0] > 3.0
    assert (cat["peak_x"][0], cat["peak_y"][0]) == (20, 20)  # first of the plateau


def test_dipoles_masks_edges_apertures():
    score = np.zeros((60, 60), np.float32)
    score[28:33, 28:33] = 8.0
    score[30, 33] = -4.0
    cat = oracle.build_catalog(score, score.copy(), expected_fwhm=3.0)
    assert cat["is_dipole"][0] == 1 and cat["quality"][0] & 2
    neg = oracle.build_catalog(score, score.copy(), expected_fwhm=3.0, return_negative=True)
    assert len(neg["x"]) == 1 and neg["is_dipole"][0] == 1
    cat = oracle.build_catalog(score, score.copy(), expected_fwhm=3.0, threshold_sigma_dipole=5.0)
    assert cat["is_dipole"][0] == 0

    score = np.zeros((40, 40), np.float32)
    score[0:3, 0:3] = 8.0
    diff = np.ones((40, 40), np.float32)
    cat = oracle.build_catalog(score, diff, expected_fwhm=3.0, aperture_radius=10)
    # aperture about the rounded centroid (1, 1), clipped to the frame: a
    # 21x21 disc of radius 10 cut at x, y >= 0
    yy, xx = np.mgrid[-10:11, -10:11]
    disc = (xx ** 2 + yy ** 2 <= 100) & (xx + 1 >= 0) & (yy + 1 >= 0)
    assert cat["flux"][0] == disc.sum()

    score = np.zeros((40, 40), np.float32)
    score[10:15, 10:15] = 8.0
    mask = np.zeros((40, 40), np.uint8)
    mask[12, 12] = _BAD
    mask[9, 12] = _COSMIC  # boundary, not a member
    cat = oracle.build_catalog(score, score.copy(), mask=mask, expected_fwhm=3.0)
    # the excluded pixel is a hole whose flag still reaches the boundary OR
    assert cat["n_pix"][0] == 24 and cat["mask_flags"][0] == _BAD | _COSMIC
    assert cat["quality"][0] & 4
    cat = oracle.build_catalog(score, score.copy(), mask=mask, expected_fwhm=3.0,
                               exclude_bad_pixels=False)
    assert cat["n_pix"][0] == 25 and cat["mask_flags"][0] == _BAD | _COSMIC

    # aperture radius from expected_fwhm: max(1, lround(3 sigma))
    score = np.zeros((50, 50), np.float32)
    score[25, 25] = 8.0
    diff = np.ones((50, 50), np.float32)
    for fwhm, r in ((3.5, 4), (0.1, 1), (10.0, 13)):
        cat = oracle.build_catalog(score, diff, expected_fwhm=fwhm)
        yy, xx = np.mgrid[-r:r + 1, -r:r + 1]
        assert cat["flux"][0] == (xx ** 2 + yy ** 2 <= r * r).sum(), fwhm


def test_psf_shape_fit_and_bright_stars():
    psf = _stamp(11, 1.5)
    score = np.zeros((60, 60), np.float32)
    diff = np.zeros((60, 60), np.float32)
    diff[25:36, 25:36] = 500.0 * psf
    score[28:33, 28:33] = 9.0
    var = np.ones((60, 60), np.float32)
    cat = oracle.build_catalog(score, diff, psf=psf, variance=var, fit_psf_shape=True,
                               expected_fwhm=3.5)
    assert cat["psf_amplitude"][0] == pytest.approx(500.0, rel=1e-6)
    assert cat["psf_chi2"][0] == pytest.approx(0.0, abs=1e-6)
    assert cat["psf_chi2_dof"][0] == 120 and cat["shape_consistent"][0] == 1
    assert cat["psf_amplitude_err"][0] == pytest.approx(np.sqrt(1 / (psf.astype(np.float64) ** 2).sum()))
    spike = np.zeros((60, 60), np.float32)
    spike[30, 30] = 500.0
    cat = oracle.build_catalog(score, spike, psf=psf, variance=var, fit_psf_shape=True,
                               expected_fwhm=3.5)
    assert cat["shape_consistent"][0] == 0 and cat["quality"][0] & 8

    cat = oracle.build_catalog(score, diff, expected_fwhm=3.5, bright_x=[33.0, 80.0],
                               bright_y=[34.0, 80.0])
    assert cat["bright_star_dist"][0] == pytest.approx(np.hypot(3, 4))
    assert cat["near_bright_star"][0] == 1 and cat["quality"][0] & 16
    cat = oracle.build_catalog(score, diff, expected_fwhm=1.0, bright_x=[33.0], bright_y=[34.0])
    assert cat["near_bright_star"][0] == 0  # radius 3 * 1.0 < 5
    cat = oracle.build_catalog(score, diff, expected_fwhm=3.5)
    assert np.isinf(cat["bright_star_dist"][0])


def _butterfly(vals, slots=1024):
    """`slots` strided partials in order; within each run of 64 slots
    l + (l ^ off) for off = 1 .. 32; then the runs' totals in order."""
    acc = [0.0] * slots
    for k, v in enumerate(vals):
        acc[k % slots] = acc[k % slots] + v
    total = None
    for r0 in range(0, slots, 64):
        run = acc[r0:r0 + 64]
        off = 1
        while off < 64:
            run = [run[l] + run[l ^ off] for l in range(64)]
            off *= 2
        total = run[0] if total is None else total + run[0]
    return total


def giant_component_frame():
    """One 110 x 120 component (13,200 px) of distinct scores above threshold."""
    rng = np.random.default_rng(17)
    s = np.zeros((160, 170), np.float32)
    s[20:130, 30:150] = rng.uniform(6.0, 400.0, (110, 120)).astype(np.float32)
    return s


def giant_component_frame_wide():
    """The same block with weights spread over 2^40, so that the centroid
    sums round differently in different orders."""
    rng = np.random.default_rng(0)
    s = np.zeros((160, 170), np.float32)
    s[20:130, 30:150] = (rng.uniform(6.0, 400.0, (110, 120))
                         * 2.0 ** rng.integers(0, 40, (110, 120))).astype(np.float32)
    return s


def test_giant_component_order_discriminates():
    s = giant_component_frame_wide()
    cat = oracle.build_catalog(s, s.copy(), expected_fwhm=3.0)
    assert len(cat["x"]) == 1 and cat["n_pix"][0] == 13200
    ys, xs = np.nonzero(s >= 5.0)
    wt = s[ys, xs].astype(np.float64)

    def centroid(f):
        tot = f(wt.tolist())
        return f((wt * xs).tolist()) / tot, f((wt * ys).tolist()) / tot

    def sequential(vals):
        t = 0.0
        for v in vals:
            t += v
        return t

    got = (cat["x"][0], cat["y"][0])
    assert got == centroid(_butterfly)
    # the pin tells the order apart: 64 slots, or one running sum, differ
    assert got != centroid(lambda v: _butterfly(v, 64))
    assert got != centroid(sequential)


def test_giant_component_fixed_order():
    s = giant_component_frame()
    cat = oracle.build_catalog(s, s.copy(), expected_fwhm=3.0)
    assert len(cat["x"]) == 1 and cat["n_pix"][0] == 13200
    ys, xs = np.nonzero(s >= 5.0)  # raster order
    wt = s[ys, xs].astype(np.float64)
    # products wt * x are exact in double, so the fma partials equal plain adds
    tot = _butterfly(wt.tolist())
    assert cat["x"][0] == _butterfly((wt * xs).tolist()) / tot
    assert cat["y"][0] == _butterfly((wt * ys).tolist()) / tot
    exact = float(np.sum(wt * xs, dtype=np.float64)) / float(np.sum(wt))
    assert cat["x"][0] == pytest.approx(exact, rel=1e-14)
    if oracle.ref() is not None:  # the reference's one-thread raster order
        ref = oracle.ref_build_catalog(s, s.copy(), expected_fwhm=3.0)
        assert cat["x"][0] == pytest.approx(ref["x"][0], rel=1e-13)
        for k in ("n_pix", "peak_x", "peak_y", "peak_snr", "flux", "mask_flags"):
            assert cat[k][0] == ref[k][0], k
