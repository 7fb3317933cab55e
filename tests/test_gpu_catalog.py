"""GPU parity of the catalog rows (src/catalog.cpp:262-697) through the C ABI.

Bar: every column bit-identical to the oracle (which is bit-identical to the
reference's compiled catalog.cpp on one thread), except expected_n_pix and
fwhm_ratio, which go through log() and are held to 4 ulp (the device log and
glibc's may round the last bit apart).  Cases: blobs of both signs, plateaus,
masks with exclusion on and off, the negative catalog, the PSF-shape fit,
bright-star distances, the max_shape_fit guard, a 2048^2 frame whose sort
spans many LDS blocks, ragged frame sizes, components above 8,192 pixels, the
capacity protocol and the device entry.
"""
import numpy as np
import pytest

import oracle
from tests._catalog_fields import CASES, blob, case_args, score_field, stamp

pytestmark = pytest.mark.gpu
_LOG_COLUMNS = ("expected_n_pix", "fwhm_ratio")


def _core():
    from delta_amd import _core
    return _core


def _same(got, want):
    for k in oracle.CATALOG_COLUMNS:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        if k in _LOG_COLUMNS:
            fin = np.isfinite(w)
            np.testing.assert_array_equal(np.isfinite(g), fin, err_msg=k)
            np.testing.assert_array_equal(g[~fin], w[~fin], err_msg=k)
            np.testing.assert_array_max_ulp(g[fin], w[fin], maxulp=4)
        else:
            np.testing.assert_array_equal(g, w, err_msg=k)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_catalog_bit_exact(gpu, case):
    s, d, kw = case_args(case)
    print(f"catalog case {case} {CASES[case]}", flush=True)
    got = _core().build_catalog(s, d, **kw)
    want = oracle.build_catalog(s, d, **kw)
    assert len(want["x"]) > 3
    _same(got, want)


def test_catalog_large_frame(gpu):
    s, d, v, m, pos = score_field(7, 2048, 2048, 900, 300, mask_frac=0.001)
    kw = dict(mask=m, threshold_sigma=3.0, variance=v, psf=stamp(11, 1.8), fit_psf_shape=True,
              bright_x=pos[::5, 0], bright_y=pos[::5, 1], max_shape_fit=20000)
    print("catalog 2048^2", flush=True)
    got = _core().build_catalog(s, d, **kw)
    want = oracle.build_catalog(s, d, **kw)
    assert len(want["x"]) > 1000 and (want["psf_chi2_dof"] > 0).sum() > 1000
    _same(got, want)


@pytest.mark.parametrize("h,w", [(151, 173), (4097, 9)])
def test_catalog_ragged_frame(gpu, h, w):
    """Frames whose pixel count is not a multiple of 4 or of the 4,096-pixel
    chunks the key compaction works in (the scalar tails of its passes)."""
    s, d, v, m, pos = score_field(11 + h, h, w, max(4, h * w // 2000), 3, mask_frac=0.01)
    kw = dict(mask=m, threshold_sigma=3.0, variance=v, psf=stamp(7, 1.5), fit_psf_shape=True)
    print(f"catalog ragged {h}x{w}", flush=True)
    got = _core().build_catalog(s, d, **kw)
    want = oracle.build_catalog(s, d, **kw)
    assert len(want["x"]) > 3
    _same(got, want)


def test_catalog_giant_component(gpu):
    from tests.test_oracle_catalog import giant_component_frame, giant_component_frame_wide

    for make in (giant_component_frame, giant_component_frame_wide):
        s = make()
        print("catalog giant component", make.__name__, flush=True)
        _same(_core().build_catalog(s, s.copy(), expected_fwhm=3.0),
              oracle.build_catalog(s, s.copy(), expected_fwhm=3.0))


def test_capacity_and_device_entry(gpu):
    import torch

    c = _core()
    s, d, kw = case_args(6)
    want = c.build_catalog(s, d, **kw)
    n = len(want["x"])
    dev = torch.device("cuda:0")
    t = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    dkw = dict(kw)
    dkw.update(mask=t(kw["mask"]), variance=t(kw["variance"]), psf=t(kw["psf"]),
               bright_x=t(np.asarray(kw["bright_x"], np.float64)),
               bright_y=t(np.asarray(kw["bright_y"], np.float64)))
    for cap in (n, n + 5, 3, 0):
        rows, total = c.build_catalog_device(t(s), t(d), capacity=cap, **dkw)
        assert total == n
        got = c.catalog_rows_to_dict(rows, min(cap, n))
        for k in oracle.CATALOG_COLUMNS:
            np.testing.assert_array_equal(got[k], want[k][:min(cap, n)], err_msg=f"{k} cap {cap}")


@pytest.mark.parametrize("case", [0, 6])
def test_labelled_entry_equals_own_labelling(gpu, case):
    """Rows from the caller's labels (dg_build_catalog_labelled_device) equal
    the rows of the catalog's own labelling of the same seeds, positive and
    negative, with and without a seed mask."""
    import torch

    c = _core()
    s, d, kw = case_args(case)
    dev = torch.device("cuda:0")
    t = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    dkw = dict(kw)
    dkw.update(mask=t(kw.get("mask")), variance=t(kw.get("variance")), psf=t(kw.get("psf")))
    for key in ("bright_x", "bright_y"):
        if kw.get(key) is not None:
            dkw[key] = t(np.asarray(kw[key], np.float64))
    for neg in (False, True):
        dkw["return_negative"] = neg
        thr = dkw.get("threshold_sigma_dipole", 3.0) if neg else dkw.get("threshold_sigma", 5.0)
        seed_mask = dkw["mask"] if dkw.get("exclude_bad_pixels", True) else None
        labels, nl = c.label_components_device(t(s), thr, -1 if neg else 1, mask=seed_mask)
        print(f"labelled catalog case {case} negative={neg}: {nl} components", flush=True)
        want, n_want = c.build_catalog_device(t(s), t(d), capacity=4096, **dkw)
        got, n_got = c.build_catalog_device(t(s), t(d), capacity=4096, labels=labels,
                                            n_labels=nl, **dkw)
        assert n_got == n_want == nl and nl > 0
        assert torch.equal(got[:n_got], want[:n_want])
        # the same labels from a buffer that is not 16-byte aligned (the
        # passes over them then load one pixel at a time)
        shifted = torch.empty(labels.numel() + 1, dtype=labels.dtype, device=dev)[1:]
        shifted.copy_(labels.reshape(-1))
        got2, n_got2 = c.build_catalog_device(t(s), t(d), capacity=4096,
                                              labels=shifted.view(labels.shape), n_labels=nl,
                                              **dkw)
        assert n_got2 == nl and torch.equal(got2[:n_got2], want[:n_want])


def test_reference_catalog_cases(gpu):
    from delta_amd import catalog

    score = blob((80, 80), 40, 40, 8.0, 2.0)
    cat = catalog.build_catalog(score, score.copy(), expected_fwhm=2.0 * 2.354820045030949)
    assert len(cat) == 1 and cat["fwhm_consistent"][0]
    assert abs(cat["x"][0] - 40) < 1.0 and cat["peak_snr"][0] > 7.5
    score = np.zeros((60, 60), np.float32)
    score[28:33, 28:33] = 8.0
    score[30, 33] = -4.0
    both = catalog.build_catalog(score, score.copy(), expected_fwhm=3.0, polarity="both")
    assert list(both["sign"]) == [1, -1] and both["is_dipole"].all()
    neg = catalog.build_catalog(score, score.copy(), expected_fwhm=3.0, polarity="negative")
    assert len(neg) == 1 and neg["peak_snr"][0] == np.float32(-4.0)
    psf = stamp(11, 1.5)
    diff = np.zeros((60, 60), np.float32)
    diff[25:36, 25:36] = 500.0 * psf
    cat = catalog.build_catalog(score, diff, psf=psf, variance=np.ones((60, 60), np.float32),
                                fit_psf_shape=True, bright_x=[33.0], bright_y=[34.0])
    assert cat["psf_amplitude"][0] == pytest.approx(500.0, rel=1e-6)
    assert cat["near_bright_star"][0] and cat["bright_star_dist"][0] == pytest.approx(5.0)
    rng = np.random.default_rng(7)
    noise = rng.normal(0.0, 1.0, (300, 300)).astype(np.float32)
    assert len(catalog.build_catalog(noise, np.zeros_like(noise), expected_fwhm=3.5)) <= 5
    with pytest.raises(RuntimeError):
        _core().build_catalog(score, np.zeros((5, 5), np.float32))
