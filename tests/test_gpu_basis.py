"""GPU: whole-image basis convolution, TPS design / evaluate / fit and the
single-block decorrelation QA functions, through the C ABI.

Bars:
* basis_convolve (src/bindings.cpp:209-227): bit-equal to the reference's own
  compiled basis_convolve (the frozen fixture) and to the oracle's restatement
  on frames that are not tile multiples, with NaN/inf pixels, across radii and
  orders; the device entry equals the host entry;
* tps_design / tps_evaluate: the oracle's design rows and design @ coeffs to
  1e-12 relative (FP64 rows on device; the reference's are Eigen's);
* tps_fit (src/spatial.cpp:131-140): numpy's solve of (D^T D + lam P) theta =
  D^T v, fitted values to 1e-9 of their scale;
* decorrelate_block / decorrelation_kernel (noise.cpp:98-161): within 2e-5
  of the oracle's PocketFFT result relative to the block's scale (the
  decorrelate bar; the FFTs differ in rounding only); a unit kernel gives the
  block back and a centred impulse.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import noise as onoise

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _core():
    from delta_amd import _core
    return _core


def test_basis_convolve_equals_reference_fixture(gpu):
    z = np.load(os.path.join(GOLDEN, "basis_convolve_ref.npz"))
    for i in range(int(z["n_cases"])):
        beta, n_max, radius = float(z[f"c{i}_beta"]), int(z[f"c{i}_n_max"]), int(z[f"c{i}_radius"])
        print(f"basis_convolve fixture case {i}", flush=True)
        got = _core().basis_convolve(z[f"c{i}_img"], beta, n_max, radius)
        np.testing.assert_array_equal(got, z[f"c{i}_bn"], err_msg=f"case {i}")


@pytest.mark.parametrize("shape,beta,n_max,radius", [
    ((130, 200), 1.3, 6, 15), ((64, 64), 0.8, 3, 4), ((77, 45), 2.2, 4, 20), ((9, 300), 1.0, 2, 1),
    ((257, 129), 1.7, 5, 0)])
def test_basis_convolve_equals_oracle(gpu, shape, beta, n_max, radius):
    rng = np.random.default_rng(sum(shape))
    img = (200 + 50 * rng.standard_normal(shape)).astype(np.float32)
    img[shape[0] // 2, shape[1] // 3] = np.nan
    img[0, -1] = np.inf
    print(f"basis_convolve {shape} beta={beta} n_max={n_max} radius={radius}", flush=True)
    got = _core().basis_convolve(img, beta, n_max, radius)
    want = oracle.basis_convolve(img, beta, n_max, radius)
    np.testing.assert_array_equal(got, want)


def test_basis_convolve_device_equals_host(gpu):
    import torch

    rng = np.random.default_rng(4)
    img = rng.standard_normal((300, 420)).astype(np.float32)
    host = _core().basis_convolve(img, 1.5, 4, 9)
    dev = _core().basis_convolve_device(torch.from_numpy(img).to(gpu), 1.5, 4, 9)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), host)


def test_basis_convolve_rejects_large_radius(gpu):
    with pytest.raises(ValueError, match="radius"):
        _core().basis_convolve(np.zeros((32, 32), np.float32), 1.0, 2, 21)


def test_tps_design_and_evaluate(gpu):
    rng = np.random.default_rng(7)
    for nx, ny in [(5, 5), (3, 1), (1, 1), (4, 3)]:
        knots = oracle.grid_knots(0.0, 0.0, 999.0, 599.0, nx, ny)
        pts = rng.uniform(-20, 1020, (3000, 2))
        d = _core().tps_design(knots, pts)
        want = oracle.tps_design(knots, pts)
        np.testing.assert_allclose(d, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        coef = rng.standard_normal(knots.shape[0])
        v = _core().tps_evaluate(knots, pts, coef)
        ref = want @ coef
        np.testing.assert_allclose(v, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    with pytest.raises(RuntimeError, match="coeffs"):
        _core().tps_evaluate(knots, pts, coef[:-1])


def test_tps_fit(gpu):
    """tps_fit against numpy; the reference's test_spatial.py cases: a field
    in the affine null space is recovered exactly at any lambda, and a large
    lambda leaves only the affine trend."""
    rng = np.random.default_rng(8)
    knots = oracle.grid_knots(0.0, 0.0, 500.0, 400.0, 4, 4)
    pts = rng.uniform(0, 500, (4000, 2)) * [1.0, 0.8]
    vals = np.sin(pts[:, 0] / 90.0) + 0.01 * pts[:, 1] + 0.05 * rng.standard_normal(len(pts))
    d = oracle.tps_design(knots, pts)
    pen = oracle.tps_setup(knots)[2]
    for lam in (1e-4, 1.0, 100.0):
        th = _core().tps_fit(knots, pts, vals, lam)
        want = np.linalg.solve(d.T @ d + lam * pen, d.T @ vals)
        assert np.abs(d @ th - d @ want).max() <= 1e-9 * np.abs(d @ want).max()
    affine = 3.0 + 0.02 * pts[:, 0] - 0.01 * pts[:, 1]
    th = _core().tps_fit(knots, pts, affine, 1e3)
    np.testing.assert_allclose(d @ th, affine, rtol=0, atol=1e-8)


@pytest.mark.parametrize("n,ks", [(64, 15), (256, 31), (32, 5), (8, 3)])
def test_decorrelate_block_matches_oracle(gpu, n, ks):
    rng = np.random.default_rng(n + ks)
    img = (10 * rng.standard_normal((n, n))).astype(np.float32)
    y, x = np.mgrid[-(ks // 2):ks // 2 + 1, -(ks // 2):ks // 2 + 1]
    kern = np.exp(-(x ** 2 + 1.3 * y ** 2) / (2 * (ks / 6.0) ** 2)).astype(np.float32)
    kern /= kern.sum()
    print(f"decorrelate_block n={n} ks={ks}", flush=True)
    got = _core().decorrelate_block(img, kern, 4.0, 2.5)
    want = onoise.decorrelate_block(img, kern, 4.0, 2.5)
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
    kg = _core().decorrelation_kernel(kern, 4.0, 2.5, n)
    kw = onoise.decorrelation_kernel(kern, 4.0, 2.5, n)
    assert np.abs(kg - kw).max() <= 2e-5 * np.abs(kw).max()


def test_decorrelation_unit_kernel(gpu):
    rng = np.random.default_rng(3)
    n = 128
    img = rng.standard_normal((n, n)).astype(np.float32)
    unit = np.zeros((3, 3), np.float32)
    unit[1, 1] = 1.0
    np.testing.assert_allclose(_core().decorrelate_block(img, unit, 2.0, 1.0), img, atol=1e-5)
    k = _core().decorrelation_kernel(unit, 2.0, 1.0, n)
    want = np.zeros((n, n), np.float32)
    want[n // 2, n // 2] = 1.0
    np.testing.assert_allclose(k, want, atol=1e-6)
    with pytest.raises(ValueError, match="power of two"):
        _core().decorrelate_block(np.zeros((48, 48), np.float32), unit, 1.0, 1.0)
    with pytest.raises(RuntimeError, match="square"):
        _core().decorrelate_block(np.zeros((32, 16), np.float32), unit, 1.0, 1.0)
