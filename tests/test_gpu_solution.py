"""GPU: photometric scale, the solution round trip and the apply path.

Bars: the photometric scale image bit-identical to the oracle (same coarse
lattice, FP32 ramp and fused float accumulation); the at-points scale within
1e-12 of numpy's design @ coefficients @ sums; a frozen solution applied to
the pair it was fitted on gives the fit-then-subtract difference, variance,
score and catalog bit for bit (as the reference's test_apply_matches_subtract);
a saved and reloaded solution gives the same; a mismatched frame raises.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _core():
    from delta_amd import _core
    return _core


@pytest.mark.parametrize("h,w,nk,nc", [(32, 40, 3, 1), (40, 48, 3, 4), (400, 400, 3, 6),
                                       (1000, 900, 5, 28), (257, 513, 4, 10)])
def test_photometric_scale_bit_exact(gpu, h, w, nk, nc):
    rng = np.random.default_rng(h + w + nc)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, nk, nk)
    theta = 0.1 * rng.standard_normal((nc + 1) * knots.shape[0])
    sums = rng.standard_normal(nc)
    print(f"photometric_scale {h}x{w} knots {nk} nc {nc}", flush=True)
    got = _core().photometric_scale(knots, theta, sums, h, w)
    np.testing.assert_array_equal(got, oracle.photometric_scale(knots, theta, sums, h, w))


def test_photometric_scale_at_and_device(gpu):
    import torch

    rng = np.random.default_rng(2)
    h, w = 50, 50
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 4, 4)
    k = knots.shape[0]
    sums = rng.standard_normal(6)
    theta = 0.2 * rng.standard_normal(7 * k)
    pts = rng.uniform([0, 0], [w - 1, h - 1], size=(40, 2))
    got = _core().photometric_scale_at(knots, theta, sums, pts)
    c = theta.reshape(7, k).T
    want = (oracle.tps_design(knots, pts) @ c)[:, :6] @ sums
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    dev = _core().photometric_scale_device(knots, theta, sums, h, w, torch.device("cuda:0"))
    np.testing.assert_array_equal(dev.cpu().numpy(), _core().photometric_scale(knots, theta,
                                                                                sums, h, w))
    with pytest.raises(RuntimeError, match="theta size mismatch"):
        _core().photometric_scale(knots, theta[:-1], sums, h, w)


def _pair():
    import torch

    from delta_amd import synth

    cfg = synth.PairConfig(size=512, n_stars=400, seed=8, n_transients=6,
                           transient_snr=(20.0, 40.0))
    p = synth.make_pair(cfg)
    dev = torch.device("cuda:0")
    return p, {k: torch.from_numpy(p[k]).to(dev) for k in ("science", "reference",
                                                             "science_var", "reference_var")}


def test_apply_matches_subtract_and_roundtrip(gpu, tmp_path):
    from delta_amd import pipeline
    from delta_amd.solution import KernelSolution

    p, t = _pair()
    kw = dict(n_knots=3, n_max=3, radius=9, stamp_radius=12, block=128)
    print("subtract then apply, 512^2", flush=True)
    res = pipeline.subtract_pair(t["science"], t["reference"], t["science_var"],
                                 t["reference_var"], **kw)
    sol = res.solution
    assert sol.direction == res.fit["direction"] and sol.shape == (512, 512)
    np.testing.assert_array_equal(sol.theta, res.fit["theta"])
    app = pipeline.apply_solution(sol, t["science"], t["reference"], t["science_var"],
                                  t["reference_var"], block=128, stamp_radius=12)
    assert app.solution is sol
    for a, b in ((app.difference, res.difference), (app.variance, res.variance),
                 (app.score, res.score), (app.labels, res.labels)):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    np.testing.assert_array_equal(app.source_catalog, res.source_catalog)

    path = str(tmp_path / "sol.npz")
    sol.save(path)
    back = KernelSolution.load(path)
    np.testing.assert_array_equal(back.photometric_scale(), sol.photometric_scale())
    app2 = pipeline.apply_solution(back, t["science"], t["reference"], t["science_var"],
                                   t["reference_var"], block=128, stamp_radius=12)
    np.testing.assert_array_equal(app2.difference.cpu().numpy(), res.difference.cpu().numpy())
    # the photometric scale at the stamps is the kernel's flux ratio, near 1
    # for this pair's unit-flux PSFs
    scale = sol.photometric_scale_at(np.column_stack([sol.stamp_x, sol.stamp_y]).astype(float))
    assert np.all((scale > 0.85) & (scale < 1.15))

    with pytest.raises(ValueError, match="solution shape"):
        pipeline.apply_solution(sol, t["science"][:100, :100].contiguous(),
                                t["reference"][:100, :100].contiguous())
