"""CPU: the numpy restatements behind tests/golden/caller_numpy.npz.

The score-PSF estimate (python/delta/pipeline.py:285-302) and the variance
rescale factor (pipeline.py:66-86) are pinned here by what their definitions
fix -- a unit-sum profile survives any amplitude and sky level, sources off
the frame and cutouts with a non-positive sum are dropped, the median of an
even count averages its two middle elements, a chi-square scale of 1 gives a
factor of 1 -- and the committed vectors are checked to be the restatements'
outputs on the committed cases.
"""
import importlib.util
import os

import numpy as np
import pytest

from tests import _caller_cases as cc

HERE = os.path.dirname(os.path.abspath(__file__))


def _mg():
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_score_psf_closed_forms():
    mg = _mg()
    r = 4
    y, x = np.mgrid[-r:r + 1, -r:r + 1]
    prof = np.exp(-(x ** 2 + y ** 2) / 4.0)
    prof /= prof.sum()
    img = np.full((60, 70), 100.0, np.float32)
    pos = [(10, 10), (30, 20), (50, 40), (2, 30)]   # the last leaves the frame
    for (px, py), amp in zip(pos, (500.0, 2000.0, 900.0, 1e6)):
        if px - r >= 0:
            img[py - r:py + r + 1, px - r:px + r + 1] += (amp * prof).astype(np.float32)
    psf = mg.score_psf_numpy(img, [p[0] for p in pos], [p[1] for p in pos], r)
    assert psf.shape == (2 * r + 1,) * 2 and psf.dtype == np.float32
    # every cutout is sky + a scaled profile: less its median and over its sum,
    # each is (prof - median(prof)) / sum(prof - median(prof)); floored and renormalised
    want = prof - np.median(prof)
    want = np.clip(want / want.sum(), 0, None)
    np.testing.assert_allclose(psf, want / want.sum(), rtol=2e-6, atol=1e-9)
    assert abs(float(psf.sum()) - 1) < 1e-6
    # a source leaving the frame on any side is skipped, not clipped
    for bad in ((r - 1, 30), (66, 30), (30, r - 1), (30, 56)):
        only = mg.score_psf_numpy(img, [10, bad[0]], [10, bad[1]], r)
        np.testing.assert_array_equal(only, mg.score_psf_numpy(img, [10], [10], r))
    # a cutout whose background-subtracted sum is not positive is dropped
    neg = img.copy()
    neg[20 - r:20 + r + 1, 40 - r:40 + r + 1] = 100.0
    neg[20, 40] = 50.0
    np.testing.assert_array_equal(mg.score_psf_numpy(neg, [10, 40], [10, 20], r),
                                  mg.score_psf_numpy(neg, [10], [10], r))
    # two usable cutouts: the per-pixel median is their mean
    a = mg.score_psf_numpy(img, [10], [10], r).astype(np.float64)
    b = mg.score_psf_numpy(img, [30], [20], r).astype(np.float64)
    two = mg.score_psf_numpy(img, [10, 30], [10, 20], r)
    np.testing.assert_allclose(two, (a + b) / (a + b).sum(), rtol=1e-6, atol=1e-9)
    with pytest.raises(ValueError, match="no usable sources"):
        mg.score_psf_numpy(np.full((60, 70), 5.0, np.float32), [10], [10], r)


def test_rescale_closed_forms():
    mg = _mg()
    rng = np.random.default_rng(1)
    # d = sqrt(v) * z with |z| = the chi2_1 median's square root -> factor exactly 1
    v = rng.uniform(1, 4, (40, 50)).astype(np.float32)
    z = np.sqrt(mg.MEDIAN_CHI2_1)
    d = (np.sqrt(v.astype(np.float64)) * z).astype(np.float32)
    assert abs(mg.rescale_numpy(d, v) - 1.0) < 1e-6
    # scaling d by s scales the factor by s^2; variance by t divides it by t
    assert abs(mg.rescale_numpy(2 * d, v) / mg.rescale_numpy(d, v) - 4) < 1e-5
    assert abs(mg.rescale_numpy(d, 4 * v) * 4 / mg.rescale_numpy(d, v) - 1) < 1e-6
    # an even count averages the two middle ratios; masked / bad pixels ignored
    dd = np.array([1, 2, 3, 4, 100, np.nan], np.float32)
    vv = np.array([1, 1, 1, 1, 0, 1], np.float32)
    assert mg.rescale_numpy(dd, vv) == ((4.0 + 9.0) / 2) / mg.MEDIAN_CHI2_1
    m = np.array([0, 0, 0, 1, 0, 0], np.uint8)
    assert mg.rescale_numpy(dd, vv, m) == 4.0 / mg.MEDIAN_CHI2_1
    assert mg.rescale_numpy(dd, np.zeros_like(vv)) == 1.0


def test_committed_vectors_are_the_restatements_outputs():
    mg = _mg()
    g = np.load(os.path.join(HERE, "golden", "caller_numpy.npz"))
    for n_src, radius in cc.PSF_CASES:
        img, xs, ys = cc.psf_case(n_src, radius)
        key = f"psf_{n_src}_{radius}"
        assert str(g[key + "_digest"]) == cc.digest(img, xs, ys)
        np.testing.assert_array_equal(g[key], mg.score_psf_numpy(img, xs, ys, radius))
    for shape, masked in cc.RESCALE_CASES:
        d, v, m = cc.rescale_case(shape, masked)
        key = f"rescale_{shape[0]}_{shape[1]}_{int(masked)}"
        assert str(g[key + "_digest"]) == cc.digest(d, v, m)
        assert float(g[key]) == mg.rescale_numpy(d, v, m)
