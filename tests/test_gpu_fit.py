"""GPU parity of the kernel fit and the GLS/GCV seam against the oracle.

The oracle is delta's algorithm: FP32 B_n, FP64 normal equations, FP64
Cholesky/GCV.  The penalised normal matrix is ill-conditioned (cond ~ 1e10 at
the selected lambda on a 4096^2 frame), so how M = X^T W X is accumulated
decides how closely the fit reproduces the reference:
  * "fp64" (FP64 MFMA, the default): identical lambda, stamp selection, pixel
    counts and IRLS history; GCV curve within 1e-9 relative; the difference
    image made with the GPU theta within 1e-6 sigma (rms) of the oracle's;
  * "fp32" / "mixed" (split-BF16, three products): the same pair-product
    GEMM with FP32 MFMA accumulation flushed to FP64 every 256 rows, held to
    the looser bars below (reduced chi^2 to 1e-3; the difference-image effect
    of the theta change measured at rms 6.9e-4 / 3.3e-3 sigma on this pair,
    bars about 4x that).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

# precision -> (GCV curve rtol, fitted-values tolerance, diff-image rms bound)
BARS = {"fp64": (1e-9, 1e-9, 1e-6), "fp32": (1e-5, 1e-4, 3e-3), "mixed": (1e-3, 1e-2, 1.5e-2)}


def _core():
    from delta_amd import _core
    return _core


def build_design(knots, points, bn):
    d = oracle.tps_design(knots, points)
    return np.hstack([d * bn[:, [c]] for c in range(bn.shape[1])] + [d])


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
def test_solve_gls_gcv_matches_oracle(gpu, precision):
    ctol, ftol, _ = BARS[precision]
    knots = oracle.grid_knots(0.0, 0.0, 100.0, 100.0, 5, 5)
    rng = np.random.default_rng(5)
    n = 6000
    pts = rng.uniform(0, 100, (n, 2))
    bn = rng.standard_normal((n, 3))
    x = build_design(knots, pts, bn)
    tgt = x @ rng.standard_normal(x.shape[1]) + 0.2 * rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n)
    grid = np.logspace(-6, 6, 25)
    print(f"solve_gls_gcv precision={precision}", flush=True)
    got = _core().solve_gls_gcv(knots, pts, tgt, w, bn, grid, precision=precision)
    ref = oracle.solve_gls_gcv(knots, pts, tgt, w, bn, grid)
    np.testing.assert_allclose(got["gcv_curve"], ref["gcv_curve"], rtol=ctol)
    if precision == "fp64":
        assert got["lambda"] == ref["lambda"]
        np.testing.assert_allclose(got["effective_dof"], ref["effective_dof"], rtol=1e-9)
    fit_g, fit_r = x @ got["theta"], x @ ref["theta"]
    assert np.abs(fit_g - fit_r).max() <= ftol * np.abs(fit_r).max()


@pytest.mark.parametrize("nx,ny,ncomp", [(1, 1, 3), (3, 1, 6), (4, 3, 15), (7, 7, 10)])
def test_solve_gls_gcv_knot_counts(gpu, nx, ny, ncomp):
    """FP64 path across spatial models: the constant field (k=1, no penalty),
    the plain affine trend (k=3, penalty rank 0) and larger TPS/component
    counts that change the pair-GEMM tile shape."""
    knots = oracle.grid_knots(0.0, 0.0, 100.0, 100.0, nx, ny)
    rng = np.random.default_rng(nx * 10 + ny)
    n = 4000
    pts = rng.uniform(0, 100, (n, 2))
    bn = rng.standard_normal((n, ncomp))
    x = build_design(knots, pts, bn)
    tgt = x @ rng.standard_normal(x.shape[1]) + 0.2 * rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n)
    grid = np.logspace(-4, 4, 9)
    print(f"solve_gls_gcv knots {nx}x{ny} components {ncomp}", flush=True)
    got = _core().solve_gls_gcv(knots, pts, tgt, w, bn, grid)
    ref = oracle.solve_gls_gcv(knots, pts, tgt, w, bn, grid)
    np.testing.assert_allclose(got["gcv_curve"], ref["gcv_curve"], rtol=1e-9)
    assert got["lambda"] == ref["lambda"]
    np.testing.assert_allclose(got["effective_dof"], ref["effective_dof"], rtol=1e-9)
    fit_g, fit_r = x @ got["theta"], x @ ref["theta"]
    assert np.abs(fit_g - fit_r).max() <= 1e-9 * np.abs(fit_r).max()


@pytest.mark.parametrize("nx,ny,ncomp", [(1, 1, 3), (3, 1, 6), (4, 3, 15), (5, 5, 6), (7, 7, 10),
                                         (9, 9, 4), (5, 5, 38)])
@pytest.mark.parametrize("descending", [False, True])
def test_reduced_rss_is_the_direct_residual_sum(gpu, nx, ny, ncomp, descending):
    """With a reduced-precision M the rss of every grid point comes from
    direct FP64 residuals (one MFMA chain per component over the knots, 16
    grid points per block).  The reported rss must equal sum w (t - x theta)^2
    at the returned theta to FP64 rounding, whatever M's precision; the knot
    counts reach each chain length (1, 3, 12, 25, 49 and 81 knots), the
    descending grid puts the pick in the second block of 16, and 38
    components make the solutions too large for LDS."""
    knots = oracle.grid_knots(0.0, 0.0, 100.0, 100.0, nx, ny)
    rng = np.random.default_rng(nx * 100 + ny * 10 + ncomp)
    n = 4000
    pts = rng.uniform(0, 100, (n, 2))
    bn = rng.standard_normal((n, ncomp))
    x = build_design(knots, pts, bn)
    tgt = x @ rng.standard_normal(x.shape[1]) + 0.2 * rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n)
    grid = np.logspace(-6, 6, 25)
    if descending:
        grid = grid[::-1].copy()
    print(f"reduced rss knots {nx}x{ny} components {ncomp} descending {descending}", flush=True)
    got = _core().solve_gls_gcv(knots, pts, tgt, w, bn, grid, precision="fp32")
    r = tgt - x @ got["theta"]
    rss = float(np.sum(w * r * r))
    assert abs(got["rss"] - rss) <= 1e-10 * rss
    best = int(np.flatnonzero(grid == got["lambda"])[0])
    assert got["gcv_curve"][best] == got["gcv"]
    edf = got["effective_dof"]
    assert got["gcv"] == pytest.approx(n * got["rss"] / (n - edf) ** 2, rel=1e-12)


def _pair(seed=3, size=256, n_stars=60):
    from delta_amd import synth

    cfg = synth.PairConfig(size=size, n_stars=n_stars, flux_min=3e3, flux_max=3e5, seed=seed)
    p = synth.make_pair(cfg)
    st = p["stars"]
    keep = (st[:, 0] > 20) & (st[:, 0] < size - 21) & (st[:, 1] > 20) & (st[:, 1] < size - 21)
    st = st[keep][:24]
    return p, np.round(st[:, 0]).astype(np.int32), np.round(st[:, 1]).astype(np.int32)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
def test_fit_kernel_matches_oracle(gpu, precision):
    _, _, dbar = BARS[precision]
    p, sx, sy = _pair()
    n = p["science"].shape[0]
    knots = oracle.grid_knots(0.0, 0.0, n - 1.0, n - 1.0, 3, 3)
    grid = np.logspace(-6, 6, 25)
    kw = dict(knots=knots, stamp_x=sx, stamp_y=sy, stamp_radius=10, beta=1.0, n_max=4,
              lambda_grid=grid, radius=10, science_var=p["science_var"],
              reference_var=p["reference_var"])
    print(f"fit_kernel precision={precision}", flush=True)
    got = _core().fit_kernel(p["science"], p["reference"], precision=precision, **kw)
    ref = oracle.fit_kernel(p["science"], p["reference"], bits=32, **kw)
    sub = dict(knots=knots, beta=1.0, n_max=4, radius=10, science_var=p["science_var"],
               reference_var=p["reference_var"])
    d_g = oracle.subtract(p["science"], p["reference"], theta=got["theta"], bits=64, **sub)
    d_r = oracle.subtract(p["science"], p["reference"], theta=ref["theta"], bits=64, **sub)
    e = (d_g["difference"] - d_r["difference"]) / np.sqrt(d_r["variance"])
    rms = float(np.sqrt((e ** 2).mean()))
    print(f"fit {precision}: theta effect on the difference image max {np.abs(e).max():.3e} "
          f"sigma, rms {rms:.3e}; reduced chi2 {got['reduced_chi2']:.6f} vs "
          f"{ref['reduced_chi2']:.6f}; lambda {got['lambda']:.3g} vs {ref['lambda']:.3g}")
    assert rms < dbar
    np.testing.assert_allclose(got["reduced_chi2"], ref["reduced_chi2"], rtol=1e-3)
    if precision == "fp64":
        for key in ("n_pixels", "n_stamps_total", "n_stamps_used", "n_stamps_rejected",
                    "irls_passes"):
            assert got[key] == ref.get(key, ref.get("passes")), key
        np.testing.assert_array_equal(got["stamp_accepted"], ref["stamp_accepted"])
        assert got["lambda"] == ref["lambda"]
        np.testing.assert_allclose(got["gcv_curve"], ref["gcv_curve"], rtol=1e-8)
        np.testing.assert_allclose(got["stamp_chi2"], ref["stamp_chi2"], rtol=1e-8)
        np.testing.assert_allclose(got["effective_dof"], ref["effective_dof"], rtol=1e-8)


def test_fit_kernel_errors(gpu):
    p, sx, sy = _pair()
    n = p["science"].shape[0]
    knots = oracle.grid_knots(0.0, 0.0, n - 1.0, n - 1.0, 3, 3)
    with pytest.raises(RuntimeError, match="no usable stamp pixels"):
        _core().fit_kernel(p["science"], p["reference"], knots, [-100], [-100], 5, 1.0, 2,
                           np.logspace(-3, 3, 5))
    with pytest.raises(RuntimeError, match="lambda_grid is empty"):
        _core().fit_kernel(p["science"], p["reference"], knots, sx, sy, 5, 1.0, 2, [])
    with pytest.raises(RuntimeError, match="lambda_grid is empty"):
        _core().fit_kernel(p["science"], p["reference"], knots, sx, sy, 5, 1.0, 2, [],
                           cv_folds=5)


def test_fit_kernel_with_masks_matches_oracle(gpu):
    """Bad science columns and masked reference pixels: the kept stamp pixels,
    the clip decisions, the grid point and the GCV curve follow the FP64
    build of the oracle (src/fit.cpp:117-229 pixel predicate)."""
    p, sx, sy = _pair()
    n = p["science"].shape[0]
    rng = np.random.default_rng(17)
    smask = np.zeros((n, n), np.uint8)
    smask[:, rng.choice(n, 6, replace=False)] = 1
    rmask = np.zeros((n, n), np.uint8)
    rmask[rng.integers(0, n, 40), rng.integers(0, n, 40)] = 2
    knots = oracle.grid_knots(0.0, 0.0, n - 1.0, n - 1.0, 3, 3)
    kw = dict(knots=knots, stamp_x=sx, stamp_y=sy, stamp_radius=10, beta=1.0, n_max=4,
              lambda_grid=np.logspace(-6, 6, 25), radius=10, science_var=p["science_var"],
              reference_var=p["reference_var"], science_mask=smask, reference_mask=rmask)
    print("fit_kernel with masks, fp64", flush=True)
    got = _core().fit_kernel(p["science"], p["reference"], precision="fp64", **kw)
    ref = oracle.fit_kernel(p["science"], p["reference"], bits=32, **kw)
    assert got["n_pixels"] == ref["n_pixels"]
    np.testing.assert_array_equal(got["stamp_accepted"], ref["stamp_accepted"])
    assert got["lambda"] == ref["lambda"]
    np.testing.assert_allclose(got["gcv_curve"], ref["gcv_curve"], rtol=1e-8)
    np.testing.assert_allclose(got["reduced_chi2"], ref["reduced_chi2"], rtol=1e-8)


def test_fit_kernel_constant_kernel_matches_oracle(gpu):
    """One knot (spatially constant kernel): no penalty, so every grid point
    takes the pivoted LDLT route (src/solve.cpp:177-181), batched on the GPU."""
    p, sx, sy = _pair()
    n = p["science"].shape[0]
    knots = np.array([[0.5 * (n - 1), 0.5 * (n - 1)]])
    kw = dict(knots=knots, stamp_x=sx, stamp_y=sy, stamp_radius=10, beta=1.0, n_max=4,
              lambda_grid=np.logspace(-3, 3, 7), radius=10, science_var=p["science_var"],
              reference_var=p["reference_var"])
    print("fit_kernel constant kernel, fp64", flush=True)
    got = _core().fit_kernel(p["science"], p["reference"], precision="fp64", **kw)
    ref = oracle.fit_kernel(p["science"], p["reference"], bits=32, **kw)
    assert got["n_pixels"] == ref["n_pixels"]
    np.testing.assert_array_equal(got["stamp_accepted"], ref["stamp_accepted"])
    np.testing.assert_allclose(got["theta"], ref["theta"], rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(got["effective_dof"], ref["effective_dof"], rtol=1e-9)
    np.testing.assert_allclose(got["reduced_chi2"], ref["reduced_chi2"], rtol=1e-8)


def test_fit_graph_replays_match_direct_launches(gpu, monkeypatch):
    """The GCV solve of the FP64 fit is replayed from a captured HIP graph
    once its arguments recur.  Five fits in a row, with other device work and
    synchronisations between them, give results bit-equal to each other and
    to a run in a fresh process with graphs off (DELTA_AMD_GRAPHS=0)."""
    import json
    import os
    import subprocess
    import sys

    import torch

    from delta_amd import synth

    p = synth.make_kernel_pair(size=512, seed=9)
    sx, sy = synth.pick_stamps(p["stars"], 512, 60, 10)
    knots = oracle.grid_knots(0.0, 0.0, 511.0, 511.0, 3, 3)
    kw = dict(stamp_radius=10, beta=float(p["params"][-3]), n_max=3, radius=9,
              lambda_grid=np.logspace(-4, 4, 9), science_var=p["science_var"],
              reference_var=p["reference_var"])
    runs = []
    for i in range(5):
        runs.append(_core().fit_kernel(p["science"], p["reference"], knots, sx, sy, **kw))
        x = torch.randn(1 << 20, device=gpu)   # unrelated eager work between replays
        float((x * x).sum())
        torch.cuda.synchronize()
    for r in runs[1:]:
        np.testing.assert_array_equal(r["theta"], runs[0]["theta"])
        np.testing.assert_array_equal(r["gcv_curve"], runs[0]["gcv_curve"])
    code = (
        "import json, numpy as np, sys; sys.path.insert(0, %r)\n"
        "from delta_amd import _core, synth\n"
        "import oracle\n"
        "p = synth.make_kernel_pair(size=512, seed=9)\n"
        "sx, sy = synth.pick_stamps(p['stars'], 512, 60, 10)\n"
        "knots = oracle.grid_knots(0.0, 0.0, 511.0, 511.0, 3, 3)\n"
        "r = _core.fit_kernel(p['science'], p['reference'], knots, sx, sy, stamp_radius=10,"
        " beta=float(p['params'][-3]), n_max=3, radius=9, lambda_grid=np.logspace(-4, 4, 9),"
        " science_var=p['science_var'], reference_var=p['reference_var'])\n"
        "print(json.dumps({'theta': r['theta'].tolist(), 'curve': r['gcv_curve'].tolist()}))\n"
    ) % os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, DELTA_AMD_GRAPHS="0")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    direct = json.loads(out.stdout.strip().splitlines()[-1])
    np.testing.assert_array_equal(np.array(direct["theta"]), runs[0]["theta"])
    np.testing.assert_array_equal(np.array(direct["curve"]), runs[0]["gcv_curve"])


def test_device_entry_points_reject_bad_tensors(gpu):
    """fit_kernel_device and build_catalog_device check every tensor (dtype,
    shape, contiguity, device) and raise ValueError before any launch."""
    import torch

    from delta_amd import _core

    h, w = 64, 72
    f32 = lambda *s: torch.zeros(s, dtype=torch.float32, device=gpu)  # noqa: E731
    sci, ref = f32(h, w), f32(h, w)
    kn = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    args = (kn, [30], [30], 8, 1.5, 2, np.logspace(-6, 6, 5))
    bad = [dict(reference=f32(h, w + 1)),
           dict(reference=torch.zeros((h, w), dtype=torch.float64, device=gpu)),
           dict(reference=torch.zeros((h, w), dtype=torch.float32)),          # host tensor
           dict(reference=f32(w, h).t()),                                     # not contiguous
           dict(science_var=f32(h + 1, w)),
           dict(reference_var=torch.zeros((h, w), dtype=torch.float16, device=gpu)),
           dict(science_mask=f32(h, w)),
           dict(reference_mask=torch.zeros((h, w + 1), dtype=torch.uint8, device=gpu))]
    for kw in bad:
        r = kw.pop("reference", ref)
        with pytest.raises(ValueError):
            _core.fit_kernel_device(sci, r, *args, **kw)
    score, diff = f32(h, w), f32(h, w)
    with pytest.raises(ValueError, match="psf"):
        _core.build_catalog_device(score, diff,
                                   psf=torch.ones((5, 5), dtype=torch.float64, device=gpu))
    bx = torch.zeros(3, dtype=torch.float32, device=gpu)
    with pytest.raises(ValueError, match="bright_x"):
        _core.build_catalog_device(score, diff, bright_x=bx, bright_y=bx.double())
