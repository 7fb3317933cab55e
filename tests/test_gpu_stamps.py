"""GPU parity of stamp selection (src/detect.cpp:84-314) through the C ABI.

Bar: bit-identical to the oracle -- background median/sigma, stamp positions
and order, flux, snr and FWHM (FP64 moments in the reference's order and
fusion), cross-matched selections, medians and direction.  Equal-flux
candidates go by value then raster index in both.  Sizes cover the exact
(<= 2^18 px) and sampled backgrounds, masks, NaN/inf, saturation, every knob,
the overflow path of the peak list (> 2^18 peaks), sorts across LDS blocks,
more kept stamps than the isolation kernel mirrors in LDS, and the catalog
path with off-frame, masked and flux-free positions.
"""
import numpy as np
import pytest

import oracle
from tests._stamp_fields import add_star, peak_flux, star_field, threshold_straddle_case

pytestmark = pytest.mark.gpu


def _core():
    from delta_amd import _core
    return _core


def _eq(a, b, keys):
    for k in keys:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)


def _same_scalar(a, b):
    assert a == b or (np.isnan(a) and np.isnan(b)), (a, b)


# --------------------------------------------------------------------------- background
@pytest.mark.parametrize("h,w,bad", [(64, 80, False), (300, 260, True), (700, 611, True),
                                     (1024, 1024, False), (3, 5, True)])
def test_background_bit_exact(gpu, h, w, bad):
    rng = np.random.default_rng(h + w)
    img = rng.normal(200.0, 7.0, (h, w)).astype(np.float32)
    mask = None
    if bad:
        img[rng.random((h, w)) < 0.02] = np.nan
        img[rng.random((h, w)) < 0.01] = -np.inf
        mask = (rng.random((h, w)) < 0.1).astype(np.uint8)
    print(f"background {h}x{w} bad={bad}", flush=True)
    assert _core().estimate_background(img, mask) == oracle.estimate_background(img, mask)


def test_background_degenerate(gpu):
    c = _core()
    assert c.estimate_background(np.full((9, 9), np.nan, np.float32)) == (0.0, 0.0)
    assert c.estimate_background(np.ones((4, 4), np.float32), np.ones((4, 4), np.uint8)) == (0, 0)
    assert c.estimate_background(np.full((30, 30), 5.0, np.float32)) == (5.0, 0.0)
    # huge values: |v - median| overflows to inf and leaves the MAD sample
    img = np.array([[-3e38, 3e38, 3e38, 1.0, 2.0]], np.float32)
    assert c.estimate_background(img) == oracle.estimate_background(img)


# --------------------------------------------------------------------------- detection
_FIELDS = [
    # (seed, h, w, n_stars, star_field kwargs, detect kwargs)
    (11, 300, 260, 60, {}, {}),
    (12, 256, 256, 120, {"nan_frac": 0.001, "mask_frac": 0.002}, {"stamp_radius": 8}),
    (13, 520, 530, 300, {}, {"stamp_radius": 10, "max_stamps": 25}),
    (14, 200, 333, 80, {}, {"stamp_radius": 6, "saturation": 900.0, "isolation_radius": 9,
                            "border": 20, "fwhm_radius": 4, "threshold_sigma": 3.0}),
    (15, 180, 170, 400, {"sigma": (0.6, 1.2)}, {"stamp_radius": 4, "max_stamps": 1000}),
    (16, 96, 700, 50, {"noise": 0.0, "peak": (5.0, 50.0)}, {"stamp_radius": 7}),
    (17, 160, 160, 30, {"flat_tops": 8}, {"stamp_radius": 6}),
    (18, 2048, 2048, 3000, {"mask_frac": 1e-5}, {}),
]


def _field(case):
    seed, h, w, n, fk, dk = _FIELDS[case]
    img, mask, _ = star_field(seed, h, w, n, **fk)
    return img, mask, dk


@pytest.mark.parametrize("case", range(len(_FIELDS)))
def test_detect_bit_exact(gpu, case):
    img, mask, dk = _field(case)
    print(f"detect case {case} {img.shape} {dk}", flush=True)
    got = _core().detect_stamps(img, mask, **dk)
    want = oracle.detect_stamps(img, mask, **dk)
    assert len(want["x"]) > 0
    _eq(got, want, ("x", "y", "flux", "snr", "fwhm"))
    assert got["x"].dtype == np.int32 and got["fwhm"].dtype == np.float64


def test_detect_many_peaks_and_kept(gpu):
    """Threshold below the median: every noise maximum is a peak (> 2^18 of
    them: the list grows and the sort spans many LDS blocks), and more stamps
    are kept than the isolation kernel holds in LDS."""
    rng = np.random.default_rng(5)
    img = rng.normal(0.0, 1.0, (1700, 1600)).astype(np.float32)
    kw = {"stamp_radius": 2, "threshold_sigma": -10.0, "max_stamps": 9000,
          "isolation_radius": 3}
    print("detect many peaks", flush=True)
    got = _core().detect_stamps(img, **kw)
    want = oracle.detect_stamps(img, **kw)
    assert len(want["x"]) == 9000
    _eq(got, want, ("x", "y", "flux", "snr", "fwhm"))


def test_detect_reference_cases(gpu):
    rng = np.random.default_rng(2)
    img = (100.0 + rng.normal(0.0, 1.0, size=(200, 200))).astype(np.float32)
    pos = [(40, 50), (120, 60), (80, 150), (160, 140)]
    for x0, y0 in pos:
        add_star(img, x0, y0, sigma=2.0, flux=peak_flux(400.0, 2.0))
    st = _core().detect_stamps(img, stamp_radius=12, threshold_sigma=5.0)
    assert len(st["x"]) == 4
    for x0, y0 in pos:
        assert any(abs(x0 - fx) <= 1 and abs(y0 - fy) <= 1 for fx, fy in zip(st["x"], st["y"]))

    rng = np.random.default_rng(3)
    img = (100.0 + rng.normal(0.0, 1.0, size=(200, 200))).astype(np.float32)
    add_star(img, 100, 100, sigma=2.0, flux=peak_flux(300.0, 2.0))
    add_star(img, 4, 100, sigma=2.0, flux=peak_flux(300.0, 2.0))
    add_star(img, 150, 60, sigma=2.0, flux=peak_flux(900.0, 2.0))
    st = _core().detect_stamps(img, stamp_radius=12, threshold_sigma=5.0, saturation=500.0)
    _eq(st, oracle.detect_stamps(img, stamp_radius=12, threshold_sigma=5.0, saturation=500.0),
        ("x", "y", "flux", "snr", "fwhm"))
    assert all(fx > 12 for fx in st["x"])


def test_detect_threshold_rounding(gpu):
    """A peak between the fused and unfused thresholds: detected exactly when
    the unfused one (the reference's) lets it through."""
    img, kw, plain = threshold_straddle_case()
    got = _core().detect_stamps(img, **kw)
    assert (31 in got["x"][got["y"] == 31]) == plain
    _eq(got, oracle.detect_stamps(img, **kw), ("x", "y", "flux", "snr", "fwhm"))


def test_detect_edge_knobs(gpu):
    c = _core()
    img, mask, _ = star_field(3, 120, 140, 40)
    for kw in ({"max_stamps": 0}, {"max_stamps": -3}, {"border": 70},
               {"stamp_radius": -2, "border": 3}, {"threshold_sigma": 1e9},
               {"isolation_radius": 1, "stamp_radius": 3, "max_stamps": 5000}):
        print(f"detect knobs {kw}", flush=True)
        _eq(c.detect_stamps(img, mask, **kw), oracle.detect_stamps(img, mask, **kw),
            ("x", "y", "flux", "snr", "fwhm"))
    with pytest.raises(ValueError):
        c.detect_stamps(img, stamp_radius=0)
    with pytest.raises(RuntimeError):
        c.detect_stamps(img, np.zeros((3, 3), np.uint8))


# --------------------------------------------------------------------------- selection
@pytest.mark.parametrize("shift,match", [((0.0, 0.0), 2), ((1.3, -0.8), 1), ((2.5, 2.5), 3)])
def test_select_bit_exact(gpu, shift, match):
    sci, sm, _ = star_field(31, 260, 240, 90, sigma=(2.0, 3.0), mask_frac=0.001)
    ref, rm, _ = star_field(31, 260, 240, 90, sigma=(1.2, 1.8), shift=shift)
    kw = {"stamp_radius": 9, "match_radius": match}
    print(f"select shift {shift} match {match}", flush=True)
    got = _core().select_stamps(sci, ref, sm, rm, **kw)
    want = oracle.select_stamps(sci, ref, sm, rm, **kw)
    assert len(want["x"]) > 3
    _eq(got, want, ("x", "y", "fwhm_science", "fwhm_reference"))
    for k in ("median_fwhm_science", "median_fwhm_reference"):
        _same_scalar(got[k], want[k])
    assert got["direction"] == want["direction"]


def test_select_catalog_bit_exact(gpu):
    sci, sm, stars = star_field(41, 200, 210, 40, mask_frac=0.0005)
    ref, _, _ = star_field(41, 200, 210, 40, sigma=(1.0, 1.5))
    rng = np.random.default_rng(43)
    cx = np.concatenate([stars[:, 0].round(), rng.integers(-20, 230, 30), [50, 50]]).astype(np.int32)
    cy = np.concatenate([stars[:, 1].round(), rng.integers(-20, 220, 30), [60, 60]]).astype(np.int32)
    for kw in ({"stamp_radius": 7, "saturation": 1500.0}, {"stamp_radius": 16},
               {"stamp_radius": 5, "fwhm_radius": 30}):
        print(f"select catalog {kw}", flush=True)
        got = _core().select_stamps(sci, ref, sm, None, cx, cy[:-1], **kw)
        want = oracle.select_stamps(sci, ref, sm, None, cx, cy[:-1], **kw)
        _eq(got, want, ("x", "y", "fwhm_science", "fwhm_reference"))
        for k in ("median_fwhm_science", "median_fwhm_reference"):
            _same_scalar(got[k], want[k])
        assert got["direction"] == want["direction"]


def test_select_reference_cases(gpu):
    c = _core()
    img_sci = np.full((200, 200), 100.0, dtype=np.float32)
    img_ref = np.full((200, 200), 100.0, dtype=np.float32)
    real = [(70, 70), (130, 110)]
    for x0, y0 in real:
        add_star(img_sci, x0, y0, sigma=3.0, flux=peak_flux(500.0, 3.0))
        add_star(img_ref, x0, y0, sigma=1.5, flux=peak_flux(500.0, 1.5))
    cat = real + [(40, 160), (160, 40), (100, 30)]
    cx = np.array([p[0] for p in cat], np.int32)
    cy = np.array([p[1] for p in cat], np.int32)
    sel = c.select_stamps(img_sci, img_ref, catalog_x=cx, catalog_y=cy, stamp_radius=16)
    assert np.isfinite(sel["median_fwhm_science"]) and np.isfinite(sel["median_fwhm_reference"])
    assert sel["median_fwhm_science"] > sel["median_fwhm_reference"]
    assert sel["direction"] == "reference"

    rng = np.random.default_rng(5)
    sci = (100.0 + rng.normal(0, 1, (200, 200))).astype(np.float32)
    ref = (100.0 + rng.normal(0, 1, (200, 200))).astype(np.float32)
    for x0, y0 in [(60, 60), (130, 70), (100, 150)]:
        add_star(sci, x0, y0, sigma=1.5, flux=peak_flux(400.0, 1.5))
        add_star(ref, x0, y0, sigma=3.0, flux=peak_flux(400.0, 3.0))
    sel = c.select_stamps(sci, ref, stamp_radius=12)
    assert len(sel["x"]) == 3 and sel["direction"] == "science"

    blank = np.full((64, 64), 10.0, np.float32)
    sel = c.select_stamps(blank, blank, stamp_radius=5)
    assert len(sel["x"]) == 0 and sel["direction"] == "science"
    assert np.isnan(sel["median_fwhm_science"])
    sel = c.select_stamps(blank, blank, catalog_x=np.zeros(0, np.int32),
                          catalog_y=np.zeros(0, np.int32))
    assert len(sel["x"]) == 0 and np.isnan(sel["median_fwhm_reference"])


def test_device_entries_match_host(gpu):
    import torch

    c = _core()
    dev = torch.device("cuda:0")
    sci, sm, stars = star_field(61, 400, 380, 150, sigma=(2.0, 2.6), mask_frac=0.001)
    ref, _, _ = star_field(61, 400, 380, 150, sigma=(1.3, 1.6))
    ts, tr = torch.from_numpy(sci).to(dev), torch.from_numpy(ref).to(dev)
    tm = torch.from_numpy(sm).to(dev)
    assert c.estimate_background_device(ts, tm) == c.estimate_background(sci, sm)
    got = c.detect_stamps_device(ts, tm, stamp_radius=8)
    want = c.detect_stamps(sci, sm, stamp_radius=8)
    _eq({k: v.cpu().numpy() for k, v in got.items()}, want, ("x", "y", "flux", "snr", "fwhm"))
    got = c.select_stamps_device(ts, tr, tm, None, stamp_radius=8)
    want = c.select_stamps(sci, ref, sm, None, stamp_radius=8)
    _eq({k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in got.items()}, want,
        ("x", "y", "fwhm_science", "fwhm_reference", "median_fwhm_science",
         "median_fwhm_reference", "direction"))
    cx = torch.from_numpy(stars[:, 0].round().astype(np.int32)).to(dev)
    cy = torch.from_numpy(stars[:, 1].round().astype(np.int32)).to(dev)
    got = c.select_stamps_device(ts, tr, tm, None, cx, cy, stamp_radius=8)
    want = c.select_stamps(sci, ref, sm, None, cx.cpu().numpy(), cy.cpu().numpy(),
                           stamp_radius=8)
    _eq({k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in got.items()}, want,
        ("x", "y", "fwhm_science", "fwhm_reference", "direction"))
