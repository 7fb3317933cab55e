"""Pins the CPU oracle's basis, spline, fields and subtraction.

Checks, each chosen so a plausible slip (wrong sign, index, transpose, dropped
term) fails one of them:
  * Gauss-Hermite samples == the reference's compiled basis.cpp (golden
    fixture, bit-exact) and == the closed form through numpy's hermval;
  * B_n == the reference's compiled basis_convolve (fixture, bit-exact);
  * TPS null space == LAPACK's complete QR (numpy.linalg.qr), penalty ==
    Z^T E Z, design == [U(r) Z | 1 u v] built here independently;
  * difference == S - b - sum a_n (phi_n (x) R) built with a direct numpy 2-D
    convolution (reference test_subtract.py:88-106), self-subtraction ~ 0
    (:109-125), constant-kernel variance closed form (:128-154), science-only
    variance passthrough (:157-171), mask dilation box and edge (:173-215);
  * the median fill == numpy's sort of the stride-7 sample, element n/2.
"""
import os

import numpy as np
import pytest
from numpy.polynomial.hermite import hermval

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def conv2d_same(img, kernel):
    """True 2-D convolution, zero padded, 'same' size."""
    kh, kw = kernel.shape
    ry, rx = kh // 2, kw // 2
    h, w = img.shape
    pad = np.zeros((h + 2 * ry, w + 2 * rx))
    pad[ry:ry + h, rx:rx + w] = img
    kf = kernel[::-1, ::-1]
    out = np.zeros((h, w))
    for i in range(kh):
        for j in range(kw):
            out += kf[i, j] * pad[i:i + h, j:j + w]
    return out


def gh_closed_form(beta, n_max, radius):
    from math import factorial, pi, sqrt

    x = np.arange(-radius, radius + 1) / beta
    out = []
    for n in range(n_max + 1):
        c = np.zeros(n + 1)
        c[n] = 1
        out.append(hermval(x, c) * np.exp(-x ** 2 / 2) / sqrt(2 ** n * factorial(n) * sqrt(pi) * beta))
    return np.array(out)


def kernels2d(beta, n_max, radius):
    g = oracle.gh_basis1d(beta, n_max, radius)
    return [np.outer(g[ny], g[nx]) for nx, ny in oracle.gh_orders(n_max)]


def fields_numpy(knots, theta, nc, h, w):
    ys, xs = np.mgrid[0:h, 0:w]
    pts = np.column_stack([xs.ravel(), ys.ravel()]).astype(float)
    d = oracle.tps_design(knots, pts)
    return (d @ theta.reshape(nc + 1, -1).T).reshape(h, w, nc + 1)


# ---------------------------------------------------------------- basis
def test_basis_matches_reference_fixture_bit_exact():
    z = np.load(os.path.join(GOLDEN, "basis_convolve_ref.npz"))
    for i in range(int(z["n_cases"])):
        b = oracle.gh_basis1d(float(z[f"c{i}_beta"]), int(z[f"c{i}_n_max"]), int(z[f"c{i}_radius"]))
        np.testing.assert_array_equal(b, z[f"c{i}_b1d"])


@pytest.mark.parametrize("beta,n_max,radius", [(1.3, 6, 15), (0.7, 4, 5), (2.5, 2, 12)])
def test_basis_closed_form(beta, n_max, radius):
    np.testing.assert_allclose(oracle.gh_basis1d(beta, n_max, radius),
                               gh_closed_form(beta, n_max, radius), rtol=1e-12, atol=1e-15)


def test_basis_near_orthonormal():
    # test_basis.py:54 -- the sampled basis is ~orthonormal for a wide footprint
    g = oracle.gh_basis1d(1.5, 5, 20)
    np.testing.assert_allclose(g @ g.T, np.eye(6), atol=1e-2)


def test_basis_convolve_matches_reference_fixture_bit_exact():
    z = np.load(os.path.join(GOLDEN, "basis_convolve_ref.npz"))
    for i in range(int(z["n_cases"])):
        img = z[f"c{i}_img"]
        beta, n_max, radius = float(z[f"c{i}_beta"]), int(z[f"c{i}_n_max"]), int(z[f"c{i}_radius"])
        bn = z[f"c{i}_bn"]
        h, w = img.shape
        knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
        for n in range(bn.shape[0]):
            theta = np.zeros((bn.shape[0] + 1) * 9)
            theta[n * 9 + 6] = 1.0
            d = oracle.subtract(np.zeros_like(img), img, knots, theta, beta, n_max,
                                radius=radius)["difference"]
            np.testing.assert_array_equal(-d, bn[n])


def test_basis_convolve_matches_direct_2d_convolution():
    rng = np.random.default_rng(5)
    img = rng.standard_normal((40, 45)).astype(np.float32)
    beta, n_max, radius = 1.4, 3, 6
    knots = oracle.grid_knots(0.0, 0.0, 44.0, 39.0, 3, 3)
    for n, k2 in enumerate(kernels2d(beta, n_max, radius)):
        theta = np.zeros(11 * 9)
        theta[n * 9 + 6] = 1.0
        d = oracle.subtract(np.zeros_like(img), img, knots, theta, beta, n_max,
                            radius=radius, bits=64)["difference"]
        np.testing.assert_allclose(-d, conv2d_same(img.astype(np.float64), k2), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- thin-plate spline
def test_tps_nullspace_penalty_design():
    knots = oracle.grid_knots(3.0, -2.0, 411.0, 297.0, 5, 4)
    norm, Z, pen = oracle.tps_setup(knots)
    k = knots.shape[0]
    c = knots.mean(axis=0)
    s = max(np.abs(knots - c).max(), 1e-12)
    np.testing.assert_allclose(norm, [c[0], c[1], s], rtol=1e-14)
    kn = (knots - c) / s
    q, _ = np.linalg.qr(np.column_stack([np.ones(k), kn]), mode="complete")
    np.testing.assert_allclose(Z, q[:, 3:], atol=1e-13)
    d2 = ((kn[:, None, :] - kn[None, :, :]) ** 2).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(d2 > 0, 0.5 * d2 * np.log(d2), 0.0)
    np.testing.assert_allclose(pen[:k - 3, :k - 3], Z.T @ e @ Z, atol=1e-12)
    assert np.all(pen[k - 3:] == 0) and np.all(pen[:, k - 3:] == 0)
    assert np.linalg.eigvalsh(pen).min() > -1e-10  # PSD (test_spatial.py:27)
    pts = np.array([[0.0, 0.0], [100.5, 33.0], [411.0, 297.0], [250.0, 10.0]])
    u = (pts - c) / s
    r2 = ((u[:, None, :] - kn[None, :, :]) ** 2).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        sig = np.where(r2 > 0, 0.5 * r2 * np.log(r2), 0.0)
    expect = np.column_stack([sig @ Z, np.ones(4), u])
    np.testing.assert_allclose(oracle.tps_design(knots, pts), expect, atol=1e-12)
    spacing = min(np.hypot(*(knots[i] - knots[j])) for i in range(k) for j in range(i + 1, k))
    assert abs(oracle.tps_min_spacing(knots) - spacing) < 1e-9


def test_affine_field_exact():
    # test_spatial.py:50 -- an affine field is reproduced exactly
    knots = oracle.grid_knots(0.0, 0.0, 99.0, 79.0, 3, 3)
    theta = np.zeros(2 * 9)
    theta[6:9] = [0.3, -1.2, 0.7]  # c_0 = 0.3 - 1.2 u + 0.7 v
    f, exact = oracle.fields(100, 80, knots, theta, 1, bits=64)
    norm, _, _ = oracle.tps_setup(knots)
    ys, xs = np.mgrid[0:80, 0:100]
    u, v = (xs - norm[0]) / norm[2], (ys - norm[1]) / norm[2]
    np.testing.assert_allclose(f[0], 0.3 - 1.2 * u + 0.7 * v, atol=1e-12)


def test_fields_exact_mode_matches_design():
    h, w = 40, 48
    rng = np.random.default_rng(8)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    theta = 0.1 * rng.standard_normal(5 * 9)
    f, exact = oracle.fields(w, h, knots, theta, 4, bits=32)
    assert exact
    ref = fields_numpy(knots, theta, 4, h, w)
    np.testing.assert_allclose(np.moveaxis(f, 0, -1), ref, rtol=1e-6, atol=1e-7)


def test_fields_coarse_lattice_bound():
    # test_subtract.py:45-70: coarse lattice + bilinear stays within 5e-3 ptp
    h, w = 400, 400
    rng = np.random.default_rng(7)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    theta = 0.1 * rng.standard_normal(7 * 9)
    f, exact = oracle.fields(w, h, knots, theta, 6, bits=64)
    assert not exact
    ref = np.moveaxis(fields_numpy(knots, theta, 6, h, w), -1, 0)
    err = np.abs(f - ref).max(axis=(1, 2))
    ptp = np.ptp(ref, axis=(1, 2))
    assert np.all(err < 5e-3 * ptp) and np.all(err > 0)
    # lattice nodes are exact
    assert np.allclose(f[:, ::12, ::12][:, :2, :2], ref[:, ::12, ::12][:, :2, :2], atol=1e-12)


# ---------------------------------------------------------------- subtraction
def test_difference_matches_numpy_model():
    h, w, beta, n_max = 48, 56, 2.0, 2
    rng = np.random.default_rng(0)
    ref = rng.standard_normal((h, w)).astype(np.float32)
    sci = rng.standard_normal((h, w)).astype(np.float32)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 4, 4)
    radius = oracle.default_radius(beta, n_max)
    ks = kernels2d(beta, n_max, radius)
    nc = len(ks)
    theta = 0.1 * rng.standard_normal((nc + 1) * 16)
    f = fields_numpy(knots, theta, nc, h, w)
    model = f[:, :, nc] + sum(f[:, :, n] * conv2d_same(ref.astype(np.float64), ks[n])
                              for n in range(nc))
    for bits, tol in ((32, 1e-4), (64, 1e-10)):
        res = oracle.subtract(sci, ref, knots, theta, beta, n_max, bits=bits)
        assert res["variance"] is None and res["mask"] is None
        np.testing.assert_allclose(res["difference"], sci - model, rtol=tol, atol=tol)


def test_self_subtraction_is_zero():
    h, w, beta, n_max = 40, 40, 1.8, 2
    rng = np.random.default_rng(1)
    ref = (5.0 + rng.standard_normal((h, w))).astype(np.float32)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    radius = oracle.default_radius(beta, n_max)
    ks = kernels2d(beta, n_max, radius)
    theta = 0.05 * rng.standard_normal((len(ks) + 1) * 9)
    f = fields_numpy(knots, theta, len(ks), h, w)
    model = f[:, :, -1] + sum(f[:, :, n] * conv2d_same(ref.astype(np.float64), k)
                              for n, k in enumerate(ks))
    res = oracle.subtract(model.astype(np.float32), ref, knots, theta, beta, n_max)
    np.testing.assert_allclose(res["difference"], 0.0, atol=1e-3)


def test_variance_constant_kernel_closed_form():
    h, w, beta = 50, 50, 2.0
    rng = np.random.default_rng(2)
    ref = rng.standard_normal((h, w)).astype(np.float32)
    sci = rng.standard_normal((h, w)).astype(np.float32)
    var_r = rng.uniform(0.5, 2.0, (h, w)).astype(np.float32)
    var_s = rng.uniform(0.5, 2.0, (h, w)).astype(np.float32)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    theta = np.zeros(18)
    theta[6] = 1.3
    res = oracle.subtract(sci, ref, knots, theta, beta, 0, science_var=var_s, reference_var=var_r)
    radius = oracle.default_radius(beta, 0)
    phi0 = kernels2d(beta, 0, radius)[0]
    expected = var_s + 1.3 ** 2 * conv2d_same(var_r.astype(np.float64), phi0 ** 2)
    np.testing.assert_allclose(res["variance"], expected, rtol=1e-3, atol=1e-3)


def test_variance_science_only_passthrough():
    rng = np.random.default_rng(3)
    ref = rng.standard_normal((30, 30)).astype(np.float32)
    sci = rng.standard_normal((30, 30)).astype(np.float32)
    var_s = rng.uniform(0.5, 2.0, (30, 30)).astype(np.float32)
    knots = oracle.grid_knots(0.0, 0.0, 29.0, 29.0, 3, 3)
    res = oracle.subtract(sci, ref, knots, np.zeros(4 * 9), 1.5, 1, science_var=var_s)
    np.testing.assert_array_equal(res["variance"], var_s)


def test_mask_growth_dilation_and_edge():
    h, w, beta, n_max = 41, 41, 2.0, 1
    rng = np.random.default_rng(4)
    ref = rng.standard_normal((h, w)).astype(np.float32)
    sci = rng.standard_normal((h, w)).astype(np.float32)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    rm = np.zeros((h, w), np.uint8)
    rm[20, 20] = 1
    sm = np.zeros((h, w), np.uint8)
    sm[10, 30] = 1
    res = oracle.subtract(sci, ref, knots, np.zeros(4 * 9), beta, n_max,
                          science_mask=sm, reference_mask=rm)
    m = res["mask"]
    r = oracle.default_radius(beta, n_max)
    assert (m[20 - r:20 + r + 1, 20 - r:20 + r + 1] & 1).all()
    assert (m & 1).sum() == (2 * r + 1) ** 2 + 1  # the box plus the science pixel
    assert m[10, 30] & 1
    assert (m[0, :] & 16).all() and (m[:, 0] & 16).all()
    assert (m[r:-r, r:-r] & 16 == 0).all()
    # exact-bad pixels (the two input pixels) are filled; grown ones keep D
    d_plain = oracle.subtract(sci, ref, knots, np.zeros(4 * 9), beta, n_max)["difference"]
    grown = (m != 0)
    grown[20, 20] = grown[10, 30] = False
    np.testing.assert_array_equal(res["difference"][grown], d_plain[grown])
    assert res["difference"][20, 20] == res["difference"][10, 30]


def test_saturation_grows_reference_only():
    h, w = 60, 60
    rng = np.random.default_rng(9)
    ref = rng.standard_normal((h, w)).astype(np.float32)
    sci = rng.standard_normal((h, w)).astype(np.float32)
    ref[30, 30] = 100.0
    sci[15, 40] = 100.0
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    m = oracle.subtract(sci, ref, knots, np.zeros(4 * 9), 1.0, 1, radius=4, saturation=50.0)["mask"]
    assert (m[26:35, 26:35] & 2).all() and not (m[25, 30] & 2) and not (m[30, 35] & 2)
    assert m[15, 40] & 2 and not (m[15, 41] & 2)


def test_fill_value_is_upper_median_of_stride7_sample():
    rng = np.random.default_rng(12)
    img = rng.standard_normal((97, 101)).astype(np.float32)
    img[3, 4] = np.nan
    mask = (rng.random((97, 101)) < 0.1).astype(np.uint8)
    flat, mflat = img.ravel(), mask.ravel()
    s = flat[::7][np.isfinite(flat[::7]) & (mflat[::7] == 0)]
    assert oracle.fill_value(img, mask) == np.sort(s)[s.size // 2]
    s2 = flat[::7][np.isfinite(flat[::7])]
    assert oracle.fill_value(img) == np.sort(s2)[s2.size // 2]
    assert oracle.fill_value(np.full((5, 5), np.nan, np.float32)) == 0.0
