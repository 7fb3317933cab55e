"""KernelSolution file format and the photometric-scale oracle (no device).

The .npz keys, dtypes and shapes are the ones the reference's
KernelSolution.save writes (python/delta/solution.py:60-91); a file in that
layout, with or without the later optional keys, loads.  The oracle's
photometric scale is pinned by the reference's tests/test_fit.py:120-166 and
tests/test_subtract.py:45-85 cases: a constant kernel gives amp * S exactly, the
exact-fallback small frame matches the at-points definition, and the coarse
lattice of a large frame stays within 5e-3 of the field's range.
"""
import numpy as np
import pytest

import oracle

_KEYS = {"beta", "n_max", "radius", "knots", "theta", "n_components", "n_spatial", "lam",
         "gcv", "effective_dof", "rss", "component_sums", "direction", "shape", "reduced_chi2",
         "n_stamps_used", "n_stamps_rejected", "stamp_x", "stamp_y", "stamp_chi2",
         "stamp_accepted"}


def _solution(**kw):
    from delta_amd.solution import KernelSolution

    rng = np.random.default_rng(0)
    knots = oracle.grid_knots(0.0, 0.0, 99.0, 79.0, 3, 3)
    base = dict(beta=1.7, n_max=2, radius=0, knots=knots, theta=rng.normal(size=7 * 9),
                n_components=6, n_spatial=9, lam=3.2e-2, gcv=1.5, effective_dof=40.2,
                rss=12.5, component_sums=rng.normal(size=6), direction="science",
                shape=(80, 100), reduced_chi2=1.02, n_stamps_used=30, n_stamps_rejected=2,
                stamp_x=np.arange(32, dtype=np.int32), stamp_y=np.arange(32, dtype=np.int32),
                stamp_chi2=rng.random(32), stamp_accepted=np.ones(32, np.uint8))
    base.update(kw)
    return KernelSolution(**base)


def test_npz_layout_and_roundtrip(tmp_path):
    from delta_amd.solution import KernelSolution

    sol = _solution()
    path = str(tmp_path / "sol.npz")
    sol.save(path)
    d = np.load(path, allow_pickle=False)
    assert set(d.files) == _KEYS
    assert d["direction"].dtype.kind == "U" and d["shape"].tolist() == [80, 100]
    assert d["stamp_x"].dtype == np.int32 and d["stamp_accepted"].dtype == np.uint8
    back = KernelSolution.load(path)
    for f in ("beta", "n_max", "radius", "lam", "gcv", "effective_dof", "rss", "direction",
              "shape", "reduced_chi2", "n_stamps_used", "n_stamps_rejected", "n_components",
              "n_spatial"):
        assert getattr(back, f) == getattr(sol, f), f
    for f in ("knots", "theta", "component_sums", "stamp_x", "stamp_y", "stamp_chi2",
              "stamp_accepted"):
        np.testing.assert_array_equal(getattr(back, f), getattr(sol, f))
    assert back.header_cards()["DLTCONV"] == "science"
    assert back.header_cards()["DLTNKNOT"] == 9


def test_absent_diagnostics_and_older_files(tmp_path):
    from delta_amd.solution import KernelSolution

    sol = _solution(stamp_x=None, stamp_y=None, stamp_chi2=None, stamp_accepted=None)
    path = str(tmp_path / "s.npz")
    sol.save(path)
    back = KernelSolution.load(path)
    assert back.stamp_x.size == 0 and back.stamp_x.dtype == np.int32
    # a file without the keys added later loads with their defaults
    d = dict(np.load(path, allow_pickle=False))
    for k in ("reduced_chi2", "n_stamps_used", "n_stamps_rejected", "stamp_x", "stamp_y",
              "stamp_chi2", "stamp_accepted"):
        d.pop(k)
    old = str(tmp_path / "old.npz")
    np.savez(old, **d)
    back = KernelSolution.load(old)
    assert np.isnan(back.reduced_chi2) and back.n_stamps_used == 0 and back.stamp_x is None


def _exact_scale(knots, theta, sums, pts):
    k = knots.shape[0]
    c = np.asarray(theta).reshape(len(sums) + 1, k).T
    return (oracle.tps_design(knots, pts) @ c)[:, :len(sums)] @ sums


def test_oracle_photometric_scale_pins():
    h, w = 32, 40
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    k = knots.shape[0]
    sums = np.array([2.37])
    theta = np.zeros(2 * k)
    theta[k - 3] = 0.7  # the intercept column of [sigma Z | 1 | u | v]
    np.testing.assert_allclose(oracle.photometric_scale(knots, theta, sums, h, w), 0.7 * 2.37,
                               rtol=1e-6)

    rng = np.random.default_rng(8)
    h, w = 40, 48  # small: the exact per-pixel fallback
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    sums = rng.standard_normal(4)
    theta = 0.1 * rng.standard_normal(5 * knots.shape[0])
    ys, xs = np.mgrid[0:h, 0:w]
    pts = np.column_stack([xs.ravel(), ys.ravel()]).astype(float)
    np.testing.assert_allclose(oracle.photometric_scale(knots, theta, sums, h, w),
                               _exact_scale(knots, theta, sums, pts).reshape(h, w),
                               rtol=1e-5, atol=1e-5)

    rng = np.random.default_rng(7)
    h, w = 400, 400  # the coarse lattice is active
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    sums = rng.standard_normal(6)
    theta = 0.1 * rng.standard_normal(7 * knots.shape[0])
    ys, xs = np.mgrid[0:h, 0:w]
    pts = np.column_stack([xs.ravel(), ys.ravel()]).astype(float)
    exact = _exact_scale(knots, theta, sums, pts).reshape(h, w)
    err = np.abs(oracle.photometric_scale(knots, theta, sums, h, w) - exact).max()
    assert 0.0 < err < 5e-3 * np.ptp(exact)
