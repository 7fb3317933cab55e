"""Pins the oracle of the differentiable constant-kernel model.

The model, its shift and its gradient have no counterpart in the reference,
so each piece is pinned by something independent of the oracle's own code:
  * zero-shift samples == gh_init (itself bit-exact against the reference's
    compiled basis.cpp) and == the Hermite closed form through numpy's
    hermval at x = i - r - shift;
  * the ladder derivatives dg/dx, dg/dbeta == central differences of that
    closed form;
  * D at zero shift == or_subtract with one knot (the constant-kernel mode,
    itself pinned against a numpy convolution in test_oracle_subtract.py);
  * an impulse reference gives D = -K around the impulse (orientation of the
    true convolution, sign of the shift);
  * the analytic gradient == central finite differences of chi^2 (step 1e-3)
    for every parameter, and == 0 where S is exactly the model.
"""
from math import factorial, pi, sqrt

import numpy as np
import pytest
from numpy.polynomial.hermite import hermval

import oracle


def gh_direct(beta, n_max, radius, shift):
    x = (np.arange(-radius, radius + 1) - shift) / beta
    out = []
    for n in range(n_max + 1):
        c = np.zeros(n + 1)
        c[n] = 1
        out.append(hermval(x, c) * np.exp(-x ** 2 / 2) /
                   sqrt(2 ** n * factorial(n) * sqrt(pi) * beta))
    return np.array(out)


def nc_of(n_max):
    return (n_max + 1) * (n_max + 2) // 2


def test_zero_shift_equals_basis():
    for beta, n_max, r in ((1.3, 6, 15), (2.1, 4, 10), (0.9, 8, 7)):
        g, _, _ = oracle.gh_shifted(beta, n_max, r, 0.0)
        np.testing.assert_array_equal(g, oracle.gh_basis1d(beta, n_max, r))


@pytest.mark.parametrize("shift", [0.3, -0.2, 0.77, -1.4])
def test_shifted_samples_closed_form(shift):
    g, _, _ = oracle.gh_shifted(1.3, 6, 15, shift)
    np.testing.assert_allclose(g, gh_direct(1.3, 6, 15, shift), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("beta,shift", [(1.3, 0.3), (2.0, -0.45), (0.8, 0.0)])
def test_ladder_derivatives(beta, shift):
    _, dgx, dgb = oracle.gh_shifted(beta, 6, 12, shift)
    h = 1e-5
    # g(i - r - shift): d/dx at fixed beta is -d/dshift
    fd_x = -(gh_direct(beta, 6, 12, shift + h) - gh_direct(beta, 6, 12, shift - h)) / (2 * h)
    fd_b = (gh_direct(beta + h, 6, 12, shift) - gh_direct(beta - h, 6, 12, shift)) / (2 * h)
    np.testing.assert_allclose(dgx, fd_x, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(dgb, fd_b, rtol=1e-7, atol=1e-9)


def random_params(rng, n_max, beta=1.3, dx=0.3, dy=-0.2, b=2.5):
    nc = nc_of(n_max)
    p = np.zeros(nc + 4)
    p[:nc] = 0.05 * rng.standard_normal(nc)
    p[0] = 0.2
    p[nc:] = (b, beta, dx, dy)
    return p


def test_model_equals_constant_kernel_subtract():
    rng = np.random.default_rng(11)
    h, w, n_max, r = 45, 52, 4, 6
    sci = (1000 + 40 * rng.standard_normal((h, w))).astype(np.float32)
    ref = (1000 + 40 * rng.standard_normal((h, w))).astype(np.float32)
    p = random_params(rng, n_max, dx=0.0, dy=0.0, beta=1.7)
    nc = nc_of(n_max)
    d = oracle.kernel_chi2(sci, ref, None, p, n_max, r, grad=False, diff=True)["difference"]
    knots = np.array([[w / 2.0, h / 2.0]])
    sub = oracle.subtract(sci, ref, knots, p[:nc + 1], p[nc + 1], n_max, radius=r, bits=64)
    np.testing.assert_allclose(d, sub["difference"], rtol=0, atol=1e-9)


def test_impulse_reference_gives_kernel():
    n_max, r = 4, 7
    h = w = 41
    ref = np.zeros((h, w), np.float32)
    ref[20, 20] = 1.0
    sci = np.zeros((h, w), np.float32)
    rng = np.random.default_rng(5)
    p = random_params(rng, n_max, dx=0.6, dy=-0.35, b=0.0)
    d = oracle.kernel_chi2(sci, ref, None, p, n_max, r, grad=False, diff=True)["difference"]
    k = oracle.shifted_kernel(p, n_max, r)
    np.testing.assert_allclose(-d[20 - r:21 + r, 20 - r:21 + r], k, rtol=0, atol=1e-15)
    # centroid follows the shift for a pure g_00 kernel (up to the sampling
    # error of a Gaussian of width 1.4 on the integer grid)
    p0 = np.zeros_like(p)
    p0[0], p0[nc_of(n_max) + 1:] = 1.0, (1.4, 0.6, -0.35)
    k0 = oracle.shifted_kernel(p0, n_max, r)
    u = np.arange(-r, r + 1)
    assert abs((k0.sum(0) * u).sum() / k0.sum() - 0.6) < 1e-4
    assert abs((k0.sum(1) * u).sum() / k0.sum() + 0.35) < 1e-4


@pytest.mark.parametrize("n_max,r", [(3, 5), (6, 9)])
def test_gradient_matches_central_differences(n_max, r):
    rng = np.random.default_rng(7 + n_max)
    h, w = 40, 48
    ref = (500 + 300 * rng.random((h, w))).astype(np.float32)
    sci = (500 + 300 * rng.random((h, w))).astype(np.float32)
    wt = rng.random((h, w)).astype(np.float32)
    wt[rng.random((h, w)) < 0.1] = 0.0
    p = random_params(rng, n_max)
    o = oracle.kernel_chi2(sci, ref, wt, p, n_max, r)
    step = 1e-3
    fd = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = step
        fd[i] = (oracle.kernel_chi2(sci, ref, wt, p + e, n_max, r, grad=False)["chi2"] -
                 oracle.kernel_chi2(sci, ref, wt, p - e, n_max, r, grad=False)["chi2"]) / (2 * step)
    err = np.abs(o["grad"] - fd) / np.abs(fd).max()
    assert err.max() < 1e-6, err


def test_gradient_vanishes_at_exact_model():
    rng = np.random.default_rng(3)
    n_max, r, h, w = 4, 6, 36, 30
    ref = (200 * rng.random((h, w))).astype(np.float64)
    p = random_params(rng, n_max)
    k = oracle.shifted_kernel(p, n_max, r)
    # S = b + K (x) R exactly, by a direct numpy convolution (zero padded)
    pad = np.zeros((h + 2 * r, w + 2 * r))
    pad[r:r + h, r:r + w] = ref
    model = np.zeros((h, w))
    for iy in range(2 * r + 1):
        for ix in range(2 * r + 1):
            model += k[iy, ix] * pad[2 * r - iy:2 * r - iy + h, 2 * r - ix:2 * r - ix + w]
    sci = p[nc_of(n_max)] + model
    # float32 inputs: evaluate where the float32 rounding of S is exact enough
    o = oracle.kernel_chi2(sci.astype(np.float32), ref.astype(np.float32), None, p, n_max, r)
    scale = oracle.kernel_chi2(sci.astype(np.float32), ref.astype(np.float32), None,
                               p + 0.01, n_max, r)
    assert o["chi2"] < 1e-9 * scale["chi2"]
    assert np.abs(o["grad"]).max() < 1e-4 * np.abs(scale["grad"]).max()
