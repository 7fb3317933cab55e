"""GPU test of the device-resident pipeline (Subtractor._fit/_subtract counterpart).

Each stage's output is checked against the oracle applied to the previous
stage's GPU output: the score bit-exactly (oracle matched filter of the GPU
whitened difference, PSF and variance), the labels bit-exactly (oracle
labelling of the GPU score); the decorrelation of the GPU difference to the
decorrelation tolerance.  Bright injected transients must be recovered.
"""
import numpy as np
import pytest

import oracle
from oracle import noise

pytestmark = pytest.mark.gpu


def test_pipeline_stages_and_recovery(gpu):
    import torch

    from delta_amd import _core, grid_knots, pipeline, synth

    cfg = synth.PairConfig(size=768, n_stars=900, seed=5, n_transients=12,
                           transient_snr=(15.0, 40.0))
    p = synth.make_pair(cfg)
    sx, sy = synth.pick_stamps(p["stars"], cfg.size, 60, 12)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(p[k]).to(dev) for k in ("science", "reference", "science_var",
                                                      "reference_var")}
    kw = dict(n_knots=3, n_max=3, radius=9, stamp_radius=12, block=128, threshold=5.0,
              fwhm=(cfg.ref_fwhm, 0.5 * (cfg.sci_fwhm0 + cfg.sci_fwhm1)))
    print("pipeline 768^2", flush=True)
    res = pipeline.subtract_pair(t["science"], t["reference"], t["science_var"],
                                 t["reference_var"], sx, sy, **kw)
    fit = res.fit
    beta = synth.default_beta(*kw["fwhm"])
    knots = grid_knots(0.0, 0.0, cfg.size - 1.0, cfg.size - 1.0, 3, 3)
    # decorrelation of the GPU difference image
    sub = _core.subtract(p["science"], p["reference"], knots, fit["theta"], beta, 3, 9, 0.0,
                         p["science_var"], p["reference_var"])
    want = noise.decorrelate(sub["difference"], knots, fit["theta"], beta, 3, p["science_var"],
                             p["reference_var"], block=128, radius=9)
    got = res.difference.cpu().numpy()
    assert np.abs(got - want["difference"]).max() <= 2e-5 * np.std(want["difference"])
    np.testing.assert_allclose(res.variance.cpu().numpy(), want["variance"], rtol=1e-6)
    # score and labels
    score = res.score.cpu().numpy()
    np.testing.assert_array_equal(
        score, oracle.matched_filter(got, res.psf, res.variance.cpu().numpy()))
    lab, n = oracle.label(score, 5.0)
    assert n == res.n_candidates
    np.testing.assert_array_equal(res.labels.cpu().numpy(), lab)
    found = 0
    for x, y in p["transients"][:, :2]:
        ix, iy = int(round(x)), int(round(y))
        found += bool(lab[iy - 2:iy + 3, ix - 2:ix + 3].any())
    print(f"recovered {found}/{len(p['transients'])}, candidates {n}")
    assert found == len(p["transients"])
    # the source catalog of the score, against the oracle's build_catalog
    from delta_amd.catalog import measure_psf_fwhm

    cat = res.source_catalog
    want = oracle.build_catalog(score, got, threshold_sigma=5.0,
                                expected_fwhm=measure_psf_fwhm(res.psf))
    assert len(cat) == n == len(want["x"])
    for k in oracle.CATALOG_COLUMNS:
        if k in ("expected_n_pix", "fwhm_ratio"):
            np.testing.assert_array_max_ulp(cat[k], want[k], maxulp=4)
        else:
            np.testing.assert_array_equal(cat[k], want[k].astype(cat[k].dtype), err_msg=k)


def test_pipeline_selects_stamps_and_direction(gpu):
    """No stamp positions: the pair goes through select_stamps on device
    (bit-exact with the oracle), the selection's direction and FWHM medians set
    the target and beta; with the frames swapped and direction "science" the
    same fit comes back negated, as the reference's sign convention has it."""
    import torch

    from delta_amd import pipeline, synth

    cfg = synth.PairConfig(size=768, n_stars=900, seed=6, n_transients=8,
                           transient_snr=(15.0, 40.0))
    p = synth.make_pair(cfg)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(p[k]).to(dev) for k in ("science", "reference", "science_var",
                                                      "reference_var")}
    kw = dict(n_knots=3, n_max=3, radius=9, stamp_radius=12, block=128, threshold=5.0)
    print("pipeline 768^2 with stamp selection", flush=True)
    res = pipeline.subtract_pair(t["science"], t["reference"], t["science_var"],
                                 t["reference_var"], **kw)
    fit = res.fit
    sel = fit["selection"]
    want = oracle.select_stamps(p["science"], p["reference"], stamp_radius=12)
    for k in ("x", "y", "fwhm_science", "fwhm_reference"):
        np.testing.assert_array_equal(sel[k].cpu().numpy(), want[k])
    assert fit["direction"] == want["direction"] == "reference"
    assert fit["beta"] == synth.default_beta(want["median_fwhm_science"],
                                             want["median_fwhm_reference"])
    lab = res.labels.cpu().numpy()
    found = sum(bool(lab[int(round(y)) - 2:int(round(y)) + 3,
                         int(round(x)) - 2:int(round(x)) + 3].any())
                for x, y in p["transients"][:, :2])
    assert found == len(p["transients"])

    print("pipeline 768^2 swapped frames, direction science", flush=True)
    sx, sy = sel["x"].cpu().numpy(), sel["y"].cpu().numpy()
    a = pipeline.subtract_pair(t["science"], t["reference"], t["science_var"],
                               t["reference_var"], sx, sy, beta=fit["beta"], score=False, **kw)
    b = pipeline.subtract_pair(t["reference"], t["science"], t["reference_var"],
                               t["science_var"], sx, sy, beta=fit["beta"], direction="science",
                               score=False, **kw)
    np.testing.assert_array_equal(b.difference.cpu().numpy(), -a.difference.cpu().numpy())
    np.testing.assert_array_equal(b.variance.cpu().numpy(), a.variance.cpu().numpy())
