"""GPU: leave-stamp-out CV lambda selection (src/solve.cpp:271-633,
src/fit.cpp:313-383) against the oracle.

Bars: the same selected lambda and the same set of evaluated grid points;
the CV-error curve within 1e-7 of the oracle's where evaluated; the
difference image of the fit within 2e-4 sigma of the oracle's everywhere
(the FP64 solves of a system with condition ~1e10 differ in rounding only;
the GCV path's own difference error against the oracle peaks at 8e-5 sigma); cv_exact_design_bytes as
the oracle's.  Both builds: per stamp (knots >= 16 stamp widths apart) and per
row (forced with DELTA_STAMP_APPROX=0, as the reference honours it), cold and
warm-started across IRLS passes, with masks.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _core():
    from delta_amd import _core
    return _core


def _case():
    from delta_amd import synth

    p = synth.make_kernel_pair(size=1024, seed=5)
    sx, sy = synth.pick_stamps(p["stars"], 1024, 80, 12)
    knots = oracle.grid_knots(0.0, 0.0, 1023.0, 1023.0, 3, 3)
    return p, sx, sy, knots, float(p["params"][-3])


def _compare(got, want, p, knots, beta, n_max):
    assert got["lambda"] == want["lambda"]
    ev = np.isfinite(want["gcv_curve"])
    np.testing.assert_array_equal(np.isfinite(got["gcv_curve"]), ev)
    np.testing.assert_allclose(got["gcv_curve"][ev], want["gcv_curve"][ev], rtol=1e-7)
    assert got["cv_exact_design_bytes"] == want["cv_exact_design_bytes"]
    assert got["n_pixels"] == want["n_pixels"] and got["irls_passes"] == want["passes"]
    np.testing.assert_array_equal(got["stamp_accepted"], want["stamp_accepted"])
    sub = dict(knots=knots, beta=beta, n_max=n_max, radius=15, science_var=p["science_var"],
               reference_var=p["reference_var"], bits=64)
    dg = oracle.subtract(p["science"], p["reference"], theta=got["theta"], **sub)
    dw = oracle.subtract(p["science"], p["reference"], theta=want["theta"], **sub)
    sig = np.sqrt(dw["variance"])
    assert np.abs(dg["difference"] - dw["difference"]).max() / sig.min() < 2e-4


@pytest.mark.parametrize("stamped", [True, False])
def test_fit_cv_matches_oracle(gpu, monkeypatch, stamped):
    p, sx, sy, knots, beta = _case()
    if not stamped:
        monkeypatch.setenv("DELTA_STAMP_APPROX", "0")
    kw = dict(stamp_radius=12, beta=beta, n_max=4, radius=15, lambda_grid=np.logspace(-6, 4, 11),
              science_var=p["science_var"], reference_var=p["reference_var"], cv_folds=5)
    print(f"fit cv stamped={stamped}", flush=True)
    got = _core().fit_kernel(p["science"], p["reference"], knots, sx, sy, **kw)
    want = oracle.fit_kernel(p["science"], p["reference"], knots, sx, sy,
                             stamp_approx=-1 if stamped else 0, bits=64, **kw)
    assert (got["cv_exact_design_bytes"] == 0) == stamped
    _compare(got, want, p, knots, beta, 4)


def test_fit_cv_masks_and_clipping(gpu):
    p, sx, sy, knots, beta = _case()
    rng = np.random.default_rng(1)
    smask = np.zeros(p["science"].shape, np.uint8)
    smask[:, rng.integers(0, 1024, 20)] = 1
    kw = dict(stamp_radius=12, beta=beta, n_max=3, radius=15, lambda_grid=np.logspace(-6, 4, 9),
              science_var=p["science_var"], reference_var=p["reference_var"], cv_folds=4,
              science_mask=smask, clip_sigma=2.0)
    got = _core().fit_kernel(p["science"], p["reference"], knots, sx, sy, **kw)
    want = oracle.fit_kernel(p["science"], p["reference"], knots, sx, sy, bits=64, **kw)
    _compare(got, want, p, knots, beta, 3)


def _rows(seed=11, n=6000, nk=5, nc=3, size=100.0):
    knots = oracle.grid_knots(0.0, 0.0, size, size, nk, nk)
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, size, size=(n, 2))
    bn = rng.standard_normal((n, nc))
    d = oracle.tps_design(knots, pts)
    x = np.hstack([d * bn[:, c:c + 1] for c in range(nc)] + [d])
    tgt = x @ rng.standard_normal(x.shape[1]) + 0.5 * rng.standard_normal(n)
    return knots, pts, bn, x, tgt, rng.uniform(0.5, 2.0, n)


@pytest.mark.parametrize("n_groups", [3, 5, 8])
def test_solve_gls_cv_matches_oracle_and_numpy(gpu, n_groups):
    """The row seam (src/bindings.cpp:553-566): the same lambda, evaluated set
    and curve (1e-7) as the oracle, the curve against numpy's fit-on-the-rest /
    score-the-fold (1e-6), and the fitted rows (1e-8 of their scale)."""
    knots, pts, bn, x, tgt, w = _rows(seed=n_groups)
    grid = np.logspace(-5, 5, 11)
    group = (np.random.default_rng(n_groups).permutation(len(tgt)) % n_groups).astype(np.int32)
    print(f"solve_gls_cv n_groups={n_groups}", flush=True)
    got = _core().solve_gls_cv(knots, pts, tgt, w, bn, grid, group, n_groups)
    want = oracle.solve_gls_cv(knots, pts, tgt, w, bn, grid, group, n_groups)
    assert got["lambda"] == want["lambda"]
    ev = np.isfinite(want["gcv_curve"])
    np.testing.assert_array_equal(np.isfinite(got["gcv_curve"]), ev)
    np.testing.assert_allclose(got["gcv_curve"][ev], want["gcv_curve"][ev], rtol=1e-7)
    pen = oracle.tps_setup(knots)[2]
    pmat = np.kron(np.eye(bn.shape[1] + 1), pen)
    for lam, cv in zip(grid[ev], got["gcv_curve"][ev]):
        tot = 0.0
        for g in range(n_groups):
            tr, ho = group != g, group == g
            a = x[tr].T @ (w[tr, None] * x[tr]) + lam * pmat
            th = np.linalg.solve(a, x[tr].T @ (w[tr] * tgt[tr]))
            tot += np.sum(w[ho] * (tgt[ho] - x[ho] @ th) ** 2)
        np.testing.assert_allclose(cv, tot, rtol=1e-6)
    fit_g, fit_w = x @ got["theta"], x @ want["theta"]
    assert np.abs(fit_g - fit_w).max() <= 1e-8 * np.abs(fit_w).max()
    np.testing.assert_allclose(got["effective_dof"], want["effective_dof"], rtol=1e-8)


def test_solve_gls_cv_edges(gpu):
    """n_groups < 2 is the GCV solve; a group id out of range and a group
    array of the wrong length raise, as the reference does."""
    knots, pts, bn, x, tgt, w = _rows(seed=2, n=2000, nk=3, nc=2)
    grid = np.logspace(-3, 3, 7)
    group = (np.arange(len(tgt)) % 4).astype(np.int32)
    one = _core().solve_gls_cv(knots, pts, tgt, w, bn, grid, group, 1)
    gcv = _core().solve_gls_gcv(knots, pts, tgt, w, bn, grid)
    np.testing.assert_array_equal(one["theta"], gcv["theta"])
    np.testing.assert_array_equal(one["gcv_curve"], gcv["gcv_curve"])
    with pytest.raises(RuntimeError, match="group id out of range"):
        _core().solve_gls_cv(knots, pts, tgt, w, bn, grid, group, 3)
    with pytest.raises(RuntimeError, match="group.size"):
        _core().solve_gls_cv(knots, pts, tgt, w, bn, grid, group[:-1], 4)


def test_solve_gls_fixed_lambda(gpu):
    """solve_gls (src/bindings.cpp:528-537): the normal-equation solution at
    the given lambda, against numpy (1e-8 of the fitted rows' scale)."""
    knots, pts, bn, x, tgt, w = _rows(seed=4, n=3000, nk=4, nc=3)
    pmat = np.kron(np.eye(bn.shape[1] + 1), oracle.tps_setup(knots)[2])
    for lam in (1e-3, 10.0):
        got = _core().solve_gls(knots, pts, tgt, w, bn, lam)
        a = x.T @ (w[:, None] * x) + lam * pmat
        th = np.linalg.solve(a, x.T @ (w * tgt))
        assert got["lambda"] == lam
        assert got["lambda_grid"].size == 0 and got["gcv_curve"].size == 0
        assert np.abs(x @ got["theta"] - x @ th).max() <= 1e-8 * np.abs(x @ th).max()
