"""GPU: the Subtractor / subtract() / apply() layer with the reference's knobs
(python/delta/pipeline.py:305-940) and the caller arithmetic it runs on
device.

Bars:
* synth_variance (python/delta/_inputs.py:75-86) bit-equal to the reference's
  numpy expression, NaN / negative / zero pixels included;
* variance_rescale_factor (pipeline.py:66-86) bit-equal to the reference's
  numpy expression (odd and even good counts, masks, non-finite pixels, no
  good pixel -> 1.0);
* Subtractor on host arrays equals the device functions on the same options,
  bit for bit, and returns host arrays; its orchestration choices are the
  reference's: cv_folds=5 by default, the output mask at the lowest positive
  of saturation and bright_mask (the oracle's mask of the same subtraction,
  bit for bit), knots thinned by spatial_scale, variances synthesised from
  gain, the rescale applied to Var(D) and to the whitened variance, and
  apply() of the returned solution reproducing subtract() bit for bit.
"""
import os

import numpy as np
import pytest

import oracle
from tests import _caller_cases as cc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "caller_numpy.npz")


def _core():
    from delta_amd import _core
    return _core


def _golden(key, *inputs):
    """The stored output of the caller's numpy arithmetic for these inputs
    (tests/golden/make_golden.py); the digest proves they are the inputs the
    vector was made from."""
    g = np.load(GOLDEN)
    assert str(g[key + "_digest"]) == cc.digest(*inputs), f"{key}: inputs drifted"
    return g[key]


def test_synth_variance_bit_exact(gpu):
    rng = np.random.default_rng(0)
    data = (rng.standard_normal((300, 257)) * 500).astype(np.float32)
    data[0, :4] = [np.nan, -0.0, 0.0, np.inf]
    for gain, rn in ((1.5, 4.0), (2.3, 0.0), (0.7, 11.25)):
        want = np.ascontiguousarray(np.clip(data, 0.0, None) / gain + float(rn) ** 2,
                                    dtype=np.float32)
        got = _core().synth_variance(data, gain, rn)
        np.testing.assert_array_equal(got, want)
    with pytest.raises(ValueError, match="gain"):
        _core().synth_variance(data, 0.0)


@pytest.mark.parametrize("shape,masked", cc.RESCALE_CASES)
def test_variance_rescale_factor_bit_exact(gpu, shape, masked):
    """The device factor equals the reference's numpy expression
    (python/delta/pipeline.py:66-86) bit for bit, from stored vectors."""
    d, v, m = cc.rescale_case(shape, masked)
    key = f"rescale_{shape[0]}_{shape[1]}_{int(masked)}"
    print(key, flush=True)
    assert _core().variance_rescale_factor(d, v, m) == float(_golden(key, d, v, m))
    # an odd and an even good count
    for keep in (11, 12):
        dk, vk = cc.rescale_count_case(shape, keep)
        want = float(_golden(f"{key}_keep{keep}", dk, vk))
        assert _core().variance_rescale_factor(dk, vk) == want
    assert _core().variance_rescale_factor(d, np.zeros(shape, np.float32)) == 1.0


def _pair(size=512, seed=3):
    from delta_amd import synth

    cfg = synth.PairConfig(size=size, n_stars=int(600 * (size / 512) ** 2), seed=seed,
                           n_transients=6, transient_snr=(20.0, 40.0))
    p = synth.make_pair(cfg)
    rng = np.random.default_rng(seed)
    smask = np.zeros((size, size), np.uint8)
    smask[:, rng.integers(0, size, 3)] = 1
    rmask = np.zeros((size, size), bool)
    rmask[rng.integers(0, size, 2), :] = True
    return p, smask, rmask


def test_subtractor_host_equals_device_functions(gpu):
    import torch

    from delta_amd import Subtractor, pipeline

    p, smask, rmask = _pair()
    kw = dict(n_max=3, n_knots=3, stamp_radius=12, saturation=6e4, bright_mask=3e4,
              decorrelate=True, score=True, block=128, rescale_variance=True)
    print("Subtractor 512^2 host", flush=True)
    res = Subtractor(**kw).subtract(p["science"], p["reference"], science_var=p["science_var"],
                                    reference_var=p["reference_var"], science_mask=smask,
                                    reference_mask=rmask)
    assert isinstance(res.difference, np.ndarray) and isinstance(res.mask, np.ndarray)
    assert res.config["DLTCVF"] == 5 and res.elapsed > 0
    assert res.solution.n_stamps_used > 0 and res.variance_scale is not None
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev)
         for k in ("science", "reference", "science_var", "reference_var")}
    d = pipeline.subtract_pair(
        t["science"], t["reference"], t["science_var"], t["reference_var"], n_knots=3, n_max=3,
        stamp_radius=12, saturation=6e4, bright_mask=3e4, block=128, threshold=5.0,
        cv_folds=5, rescale_variance=True,
        science_mask=torch.from_numpy(smask).to(dev),
        reference_mask=torch.from_numpy(rmask.astype(np.uint8)).to(dev))
    for name in ("difference", "variance", "mask", "score"):
        np.testing.assert_array_equal(getattr(res, name), getattr(d, name).cpu().numpy(),
                                      err_msg=name)
    assert res.variance_scale == d.variance_scale
    np.testing.assert_array_equal(res.solution.theta, d.solution.theta)
    assert len(res.source_catalog) == len(d.source_catalog)
    # the output mask: the oracle's subtraction mask at min(saturation, bright_mask)
    sol = res.solution
    tgt, conv = ((p["science"], p["reference"]) if sol.direction == "reference"
                 else (p["reference"], p["science"]))
    tm, cm = ((smask, rmask) if sol.direction == "reference" else (rmask, smask))
    o = oracle.subtract(tgt, conv, sol.knots, sol.theta, sol.beta, sol.n_max,
                        radius=sol.radius, saturation=3e4,
                        science_mask=tm.astype(np.uint8), reference_mask=cm.astype(np.uint8))
    np.testing.assert_array_equal(res.mask, o["mask"])
    # apply() of the solution reproduces the products
    app = Subtractor(**kw).apply(sol, p["science"], p["reference"], science_var=p["science_var"],
                                 reference_var=p["reference_var"], science_mask=smask,
                                 reference_mask=rmask)
    for name in ("difference", "variance", "mask", "score"):
        np.testing.assert_array_equal(getattr(app, name), getattr(res, name), err_msg=name)


def test_subtractor_choices(gpu):
    """cv_folds=5 by default; spatial_scale thins the knots; gain synthesises
    the variances (bit-equal to numpy's); the rescale factor is the reference
    formula on the returned Var(D) before scaling; subtract() splits
    configuration from call keywords; no score without score=True."""
    from delta_amd import Subtractor, subtract

    p, smask, rmask = _pair(seed=4)
    print("Subtractor choices", flush=True)
    base = dict(n_max=3, n_knots=5, stamp_radius=12)
    res = subtract(p["science"], p["reference"], gain=1.5, read_noise=4.0,
                   spatial_scale=250.0, rescale_variance=True, **base)
    assert res.score is None and res.source_catalog is None
    assert res.solution.knots.shape[0] == 9         # round(511/250)+1 = 3 per axis
    fit = res.fit
    assert np.isfinite(fit["gcv_curve"]).sum() < fit["gcv_curve"].size   # CV, coarse-to-fine
    # the same call with the variances given explicitly as numpy's synth_variance
    sv = np.ascontiguousarray(np.clip(p["science"], 0.0, None) / 1.5 + 16.0, dtype=np.float32)
    rv = np.ascontiguousarray(np.clip(p["reference"], 0.0, None) / 1.5 + 16.0, dtype=np.float32)
    res2 = Subtractor(spatial_scale=250.0, rescale_variance=True, **base).subtract(
        p["science"], p["reference"], science_var=sv, reference_var=rv)
    np.testing.assert_array_equal(res.difference, res2.difference)
    np.testing.assert_array_equal(res.variance, res2.variance)
    # rescale: the factor (pinned bit-exact above) of the unscaled Var(D)
    plain = Subtractor(spatial_scale=250.0, **base).subtract(
        p["science"], p["reference"], science_var=sv, reference_var=rv)
    want = _core().variance_rescale_factor(plain.difference, plain.variance, plain.mask)
    assert res2.variance_scale == want
    np.testing.assert_array_equal(res2.variance,
                                  (plain.variance * want).astype(np.float32))


@pytest.mark.parametrize("n_src,radius", cc.PSF_CASES)
def test_estimate_psf_bit_exact(gpu, n_src, radius):
    """The score PSF on device equals the reference's numpy _estimate_psf
    (python/delta/pipeline.py:285-302) bit for bit, from stored vectors: odd
    and even source counts, sources off the frame edge, one cutout with a NaN,
    another whose background-subtracted sum is negative."""
    img, xs, ys = cc.psf_case(n_src, radius)
    print(f"estimate_psf n={n_src} r={radius}", flush=True)
    got = _core().estimate_psf(img, xs, ys, radius)
    np.testing.assert_array_equal(got, _golden(f"psf_{n_src}_{radius}", img, xs, ys))
    with pytest.raises(ValueError, match="no usable sources"):
        _core().estimate_psf(np.full((64, 64), 3.0, np.float32), [20, 30], [20, 40], 4)


def test_weighted_mean(gpu):
    """_core.weighted_mean (src/bindings.cpp:107-125): sum(d/v)/sum(1/v) over
    unmasked pixels with v > 0, 0 when none; a parallel FP64 sum, equal to
    numpy's to 1e-12."""
    rng = np.random.default_rng(6)
    d = (100 + 10 * rng.standard_normal((300, 257))).astype(np.float32)
    v = rng.uniform(0.5, 3.0, d.shape).astype(np.float32)
    v[::9] = 0.0
    v[5, 5] = -1.0
    m = (rng.random(d.shape) < 0.05).astype(np.uint8)
    good = (m == 0) & (v > 0)
    w = 1.0 / v[good].astype(np.float64)
    want = float((w * d[good].astype(np.float64)).sum() / w.sum())
    got = _core().weighted_mean(d, v, m)
    assert abs(got - want) <= 1e-12 * abs(want)
    assert _core().weighted_mean(d, np.zeros_like(v), m) == 0.0
