"""Pins of the stamp-selection oracle (src/detect.cpp:84-314).

* bit-exact against the reference's own detect.cpp compiled into oracle/_ref
  (this container only; skipped where it is absent) on random star fields with
  masks, NaNs, saturation, frames above the 2^18-pixel sampling limit, every
  parameter knob, and the catalog path;
* the cases of the reference's tests/test_detect.py, restated;
* closed forms: the median/MAD of a known sample, the FWHM of a sampled
  Gaussian, the empty and all-masked frames.

Equal-flux candidates (plateaus) are ordered by raster index in the oracle;
the reference's std::sort leaves that order open, so the plateau case checks
stamps agree to the plateau's extent rather than bit for bit.
"""
import numpy as np
import pytest

import oracle
from tests._stamp_fields import (FWHM_PER_SIGMA, add_star, peak_flux, star_field,
                                 threshold_straddle_case)

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")


def _same_stamps(a, b, keys):
    for k in keys:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)


# --------------------------------------------------------------------------- closed forms
def test_background_closed_form():
    v = np.arange(1, 102, dtype=np.float32)  # 101 samples, median 51
    img = v.reshape(1, 101)
    med, sig = oracle.estimate_background(img)
    assert med == 51.0
    # |v - 51| = 0..50, each twice but 0 once: 101 values, element 50 is 25
    assert sig == pytest.approx(1.4826 * 25.0, rel=0, abs=1e-12)
    # even count: element n/2 (upper median), NaN and masked samples dropped
    img2 = np.array([[4, 1, np.nan, 3, 2, 100]], np.float32)
    m = np.array([[0, 0, 0, 0, 0, 1]], np.uint8)
    assert oracle.estimate_background(img2, m)[0] == 3.0
    assert oracle.estimate_background(np.full((4, 4), np.nan, np.float32)) == (0.0, 0.0)
    assert oracle.estimate_background(np.ones((3, 3), np.float32),
                                      np.ones((3, 3), np.uint8)) == (0.0, 0.0)


def test_background_strided_sample_large_frame():
    rng = np.random.default_rng(7)
    h, w = 700, 611  # 427,700 px > 2^18: step 1
    img = rng.normal(50, 3, (h, w)).astype(np.float32)
    img[rng.random((h, w)) < 0.01] = np.inf
    mask = (rng.random((h, w)) < 0.05).astype(np.uint8)
    step = img.size // (1 << 18)
    flat, mflat = img.ravel()[::step], mask.ravel()[::step]
    s = np.sort(flat[np.isfinite(flat) & (mflat == 0)])
    med = s[s.size // 2]
    dev = np.sort(np.abs(s - med))
    assert oracle.estimate_background(img, mask) == (float(med), 1.4826 * float(dev[dev.size // 2]))


def test_fwhm_of_sampled_gaussian():
    for sigma in (1.5, 2.5, 4.0):
        img = np.zeros((81, 81), np.float32)
        add_star(img, 40, 40, sigma, 1e5)
        f = oracle.estimate_fwhm(img, 40, 40, 30, 0.0)
        assert f == pytest.approx(FWHM_PER_SIGMA * sigma, rel=1e-3)
    # no positive flux -> NaN; window clipped at the frame edge stays finite
    assert np.isnan(oracle.estimate_fwhm(np.zeros((9, 9), np.float32), 4, 4, 3, 0.0))
    img = np.zeros((40, 40), np.float32)
    add_star(img, 1, 1, 2.0, 1e4)
    assert np.isfinite(oracle.estimate_fwhm(img, 1, 1, 7, 0.0))


# --------------------------------------------------------------------------- reference test cases
def test_reference_cases_background_and_detect():
    rng = np.random.default_rng(1)
    img = (100.0 + rng.normal(0.0, 2.0, size=(256, 256))).astype(np.float32)
    med, sig = oracle.estimate_background(img)
    assert abs(med - 100.0) < 0.3 and abs(sig - 2.0) < 0.3

    rng = np.random.default_rng(2)
    img = (100.0 + rng.normal(0.0, 1.0, size=(200, 200))).astype(np.float32)
    pos = [(40, 50), (120, 60), (80, 150), (160, 140)]
    for x0, y0 in pos:
        add_star(img, x0, y0, sigma=2.0, flux=peak_flux(400.0, 2.0))
    st = oracle.detect_stamps(img, stamp_radius=12, threshold_sigma=5.0)
    found = list(zip(st["x"].tolist(), st["y"].tolist()))
    assert len(found) == len(pos)
    for x0, y0 in pos:
        assert any(abs(x0 - fx) <= 1 and abs(y0 - fy) <= 1 for fx, fy in found)


def test_reference_cases_edge_saturation_fwhm():
    rng = np.random.default_rng(3)
    img = (100.0 + rng.normal(0.0, 1.0, size=(200, 200))).astype(np.float32)
    add_star(img, 100, 100, sigma=2.0, flux=peak_flux(300.0, 2.0))
    add_star(img, 4, 100, sigma=2.0, flux=peak_flux(300.0, 2.0))
    add_star(img, 150, 60, sigma=2.0, flux=peak_flux(900.0, 2.0))
    st = oracle.detect_stamps(img, stamp_radius=12, threshold_sigma=5.0, saturation=500.0)
    found = list(zip(st["x"].tolist(), st["y"].tolist()))
    assert any(abs(100 - fx) <= 1 and abs(100 - fy) <= 1 for fx, fy in found)
    assert all(fx > 12 for fx, _ in found)
    assert not any(abs(150 - fx) <= 2 and abs(60 - fy) <= 2 for fx, fy in found)

    img = np.full((200, 200), 100.0, dtype=np.float32)
    add_star(img, 100, 100, sigma=2.5, flux=peak_flux(500.0, 2.5))
    st = oracle.detect_stamps(img, stamp_radius=16, threshold_sigma=5.0)
    assert len(st["x"]) == 1
    assert abs(st["fwhm"][0] - FWHM_PER_SIGMA * 2.5) / (FWHM_PER_SIGMA * 2.5) < 0.2


def _pair(seed, s_sigma, r_sigma, pos, peak=400.0, noise=True, size=200):
    rng = np.random.default_rng(seed)
    if noise:
        sci = (100.0 + rng.normal(0, 1, (size, size))).astype(np.float32)
        ref = (100.0 + rng.normal(0, 1, (size, size))).astype(np.float32)
    else:
        sci = np.full((size, size), 100.0, np.float32)
        ref = np.full((size, size), 100.0, np.float32)
    for x0, y0 in pos:
        add_star(sci, x0, y0, sigma=s_sigma, flux=peak_flux(peak, s_sigma))
        add_star(ref, x0, y0, sigma=r_sigma, flux=peak_flux(peak, r_sigma))
    return sci, ref


def test_reference_cases_select():
    pos = [(50, 50), (140, 60), (90, 150)]
    sci, ref = _pair(4, 3.0, 1.5, pos)
    sel = oracle.select_stamps(sci, ref, stamp_radius=12)
    assert len(sel["x"]) == 3
    assert sel["median_fwhm_science"] > sel["median_fwhm_reference"]
    assert sel["direction"] == "reference"
    sci, ref = _pair(5, 1.5, 3.0, [(60, 60), (130, 70), (100, 150)])
    assert oracle.select_stamps(sci, ref, stamp_radius=12)["direction"] == "science"

    pos = [(70, 70), (130, 110)]
    sci, ref = _pair(0, 3.0, 1.5, pos, peak=500.0, noise=False)
    cx = np.array([p[0] for p in pos], np.int32)
    cy = np.array([p[1] for p in pos], np.int32)
    sel = oracle.select_stamps(sci, ref, catalog_x=cx, catalog_y=cy, stamp_radius=16)
    assert len(sel["x"]) == 2 and sel["direction"] == "reference"
    assert sel["median_fwhm_science"] > sel["median_fwhm_reference"]

    cat = pos + [(40, 160), (160, 40), (100, 30)]
    cx = np.array([p[0] for p in cat], np.int32)
    cy = np.array([p[1] for p in cat], np.int32)
    sel = oracle.select_stamps(sci, ref, catalog_x=cx, catalog_y=cy, stamp_radius=16)
    assert np.isfinite(sel["median_fwhm_science"]) and np.isfinite(sel["median_fwhm_reference"])
    assert np.isnan(sel["fwhm_science"][2:]).all()
    assert sel["direction"] == "reference"


def test_direction_and_empty_selection():
    blank = np.full((64, 64), 10.0, np.float32)
    sel = oracle.select_stamps(blank, blank, stamp_radius=5)
    assert len(sel["x"]) == 0
    assert np.isnan(sel["median_fwhm_science"]) and np.isnan(sel["median_fwhm_reference"])
    assert sel["direction"] == "science"  # NaN <= NaN is false
    with pytest.raises(ValueError):
        oracle.detect_stamps(blank, stamp_radius=0)


# --------------------------------------------------------------------------- reference build
_FIELDS = [
    # (seed, h, w, n_stars, kwargs of star_field, detect kwargs)
    (11, 300, 260, 60, {}, {}),
    (12, 256, 256, 120, {"nan_frac": 0.001, "mask_frac": 0.002}, {"stamp_radius": 8}),
    (13, 520, 530, 300, {}, {"stamp_radius": 10, "max_stamps": 25}),  # sampled background
    (14, 200, 333, 80, {}, {"stamp_radius": 6, "saturation": 900.0, "isolation_radius": 9,
                            "border": 20, "fwhm_radius": 4, "threshold_sigma": 3.0}),
    (15, 180, 170, 400, {"sigma": (0.6, 1.2)}, {"stamp_radius": 4, "max_stamps": 1000}),
    (16, 96, 700, 50, {"noise": 0.0, "peak": (5.0, 50.0)}, {"stamp_radius": 7}),
]


@needs_ref
@pytest.mark.parametrize("case", range(len(_FIELDS)))
def test_detect_matches_reference_build(case):
    seed, h, w, n, fk, dk = _FIELDS[case]
    img, mask, _ = star_field(seed, h, w, n, **fk)
    assert oracle.estimate_background(img, mask) == oracle.ref_estimate_background(img, mask)
    a = oracle.detect_stamps(img, mask, **dk)
    b = oracle.ref_detect_stamps(img, mask, **dk)
    assert len(a["x"]) > 0
    _same_stamps(a, b, ("x", "y", "flux", "snr", "fwhm"))


@needs_ref
def test_fwhm_matches_reference_build():
    img, _, stars = star_field(21, 128, 128, 20)
    for x, y, _, _ in stars[:10]:
        for r in (3, 7, 40):
            a = oracle.estimate_fwhm(img, int(x), int(y), r, 100.0)
            b = oracle.ref_estimate_fwhm(img, int(x), int(y), r, 100.0)
            assert a == b or (np.isnan(a) and np.isnan(b))


@needs_ref
@pytest.mark.parametrize("shift,match", [((0.0, 0.0), 2), ((1.3, -0.8), 1), ((2.5, 2.5), 3)])
def test_select_matches_reference_build(shift, match):
    sci, sm, _ = star_field(31, 260, 240, 90, sigma=(2.0, 3.0), mask_frac=0.001)
    ref, rm, _ = star_field(31, 260, 240, 90, sigma=(1.2, 1.8), shift=shift)
    kw = {"stamp_radius": 9, "match_radius": match}
    a = oracle.select_stamps(sci, ref, sm, rm, **kw)
    b = oracle.ref_select_stamps(sci, ref, sm, rm, **kw)
    _same_stamps(a, b, ("x", "y", "fwhm_science", "fwhm_reference"))
    for k in ("median_fwhm_science", "median_fwhm_reference", "direction"):
        assert a[k] == b[k] or (np.isnan(a[k]) and np.isnan(b[k])), k


@needs_ref
def test_select_catalog_matches_reference_build():
    sci, sm, stars = star_field(41, 200, 210, 40, mask_frac=0.0005)
    ref, rm, _ = star_field(41, 200, 210, 40, sigma=(1.0, 1.5))
    rng = np.random.default_rng(43)
    cx = np.concatenate([stars[:, 0].round(), rng.integers(-20, 230, 30)]).astype(np.int32)
    cy = np.concatenate([stars[:, 1].round(), rng.integers(-20, 220, 30)]).astype(np.int32)
    kw = {"stamp_radius": 7, "saturation": 1500.0}
    a = oracle.select_stamps(sci, ref, sm, rm, cx, cy[:-3], **kw)
    b = oracle.ref_select_stamps(sci, ref, sm, rm, cx, cy[:-3], **kw)
    assert len(a["x"]) > 5
    _same_stamps(a, b, ("x", "y", "fwhm_science", "fwhm_reference"))
    for k in ("median_fwhm_science", "median_fwhm_reference", "direction"):
        assert a[k] == b[k], k


@needs_ref
def test_plateau_ties_agree_to_the_plateau():
    img, _, _ = star_field(51, 160, 160, 30, flat_tops=8)
    a = oracle.detect_stamps(img, stamp_radius=6)
    b = oracle.ref_detect_stamps(img, stamp_radius=6)
    assert len(a["x"]) == len(b["x"])
    np.testing.assert_array_equal(a["flux"], b["flux"])
    assert np.abs(a["x"] - b["x"]).max() <= 2 and np.abs(a["y"] - b["y"]).max() <= 2


@needs_ref
def test_threshold_is_unfused_as_the_reference_compiles_it():
    """detect.cpp:164-165 compiles to median + round(threshold_sigma * sigma)
    under the reference's default flags (the conditional operand keeps GCC
    from fusing it, unlike the FWHM sums); a peak placed between the fused and
    unfused thresholds tells the two apart."""
    img, kw, plain = threshold_straddle_case()
    ref = oracle.ref_detect_stamps(img, **kw)
    got = oracle.detect_stamps(img, **kw)
    assert (31 in ref["x"][ref["y"] == 31]) == plain
    _same_stamps(got, ref, ("x", "y", "flux", "snr", "fwhm"))
