"""Pins of the noise-decorrelation and matched-filter oracle.

The reference's own tests (tests/test_noise.py) pin this stage only
statistically; they are restated here against the oracle, together with
closed forms and brute force:
  * the single-block FFT apply equals a brute-force circular convolution
    with the real-space kernel ifft(Phi) (so the FFT conventions -- half
    spectrum, unnormalised forward, 1/n^2 inverse -- are the reference's);
  * Phi's closed form at chosen frequencies;
  * the matched filter equals a brute-force zero-padded correlation, with
    the reference's peak S/N identity.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import noise


def _autocorr_lag1(img):
    a = img - img.mean()
    var = a.var()
    cx = (a[:, :-1] * a[:, 1:]).mean()
    cy = (a[:-1, :] * a[1:, :]).mean()
    return 0.5 * (cx + cy) / var


def _correlated_difference(n, kernel, vs, vr, rng):
    ns = np.sqrt(vs) * rng.standard_normal((n, n))
    nr = np.sqrt(vr) * rng.standard_normal((n, n))
    ks = kernel.shape[0]
    r = ks // 2
    conv = np.zeros((n, n))
    kf = kernel[::-1, ::-1]
    padded = np.pad(nr, r, mode="wrap")
    for i in range(ks):
        for j in range(ks):
            conv += kf[i, j] * padded[i:i + n, j:j + n]
    return (ns + conv).astype(np.float32)


def _kernels(beta, n_max):
    return noise.phi_stack(beta, n_max, oracle.default_radius(beta, n_max))


def decorrelate_block(img, kern, vs, vr):
    n = img.shape[0]
    return noise.apply_filter_fft(img, noise.decorrelation_filter(kern, vs, vr, n))


def test_filter_closed_form():
    kern = _kernels(2.0, 0)[0]
    vs, vr, n = 1.5, 0.8, 64
    filt = noise.decorrelation_filter(kern, vs, vr, n)
    ks = kern.shape[0]
    r = ks // 2
    u = np.arange(ks) - r
    sumk2 = float((kern.astype(np.float64) ** 2).sum())
    for ky, kx in [(0, 0), (3, 5), (17, 31), (63, 32)]:
        ph = np.exp(-2j * np.pi * (ky * u[:, None] + kx * u[None, :]) / n)
        khat = (kern.astype(np.float64) * ph).sum()
        want = math.sqrt(vs + vr * sumk2) / math.sqrt(vs + vr * abs(khat) ** 2)
        assert filt[ky, kx] == pytest.approx(want, rel=2e-6)


def test_block_apply_is_circular_convolution():
    kern = _kernels(1.7, 1)[2]
    n = 32
    filt = noise.decorrelation_filter(kern, 1.0, 2.0, n)
    rng = np.random.default_rng(4)
    img = rng.standard_normal((n, n)).astype(np.float32)
    got = noise.apply_filter_fft(img, filt)
    # real-space kernel: inverse transform of the (real, even) filter
    phi = np.fft.irfft2(filt.astype(np.float64), s=(n, n))
    want = np.zeros((n, n))
    for dy in range(n):
        for dx in range(n):
            want += phi[dy, dx] * np.roll(img.astype(np.float64), (dy, dx), axis=(0, 1))
    np.testing.assert_allclose(got, want, atol=2e-5 * np.abs(want).max())


def test_decorrelation_flattens_autocorrelation():
    rng = np.random.default_rng(0)
    kern = _kernels(2.0, 0)[0]
    diff = _correlated_difference(128, kern, 1.0, 1.0, rng)
    white = decorrelate_block(diff, kern, 1.0, 1.0)
    assert _autocorr_lag1(diff) > 0.15
    assert abs(_autocorr_lag1(white)) < 0.03


def test_decorrelation_preserves_noise_variance():
    rng = np.random.default_rng(1)
    kern = _kernels(2.0, 0)[0]
    vs, vr = 1.5, 0.8
    sumk2 = float((kern.astype(np.float64) ** 2).sum())
    diff = _correlated_difference(128, kern, vs, vr, rng)
    white = decorrelate_block(diff, kern, vs, vr)
    expected = vs + vr * sumk2
    np.testing.assert_allclose(diff.var(), expected, rtol=0.1)
    np.testing.assert_allclose(white.var(), expected, rtol=0.1)


def _const_theta(nc, k):
    theta = np.zeros((nc + 1) * k)
    theta[k - 3] = 1.0  # a_0 = 1 through the affine intercept
    return theta


def test_decorrelate_spatially_varying_matches_constant_interior():
    n, beta, n_max = 192, 2.0, 1
    rng = np.random.default_rng(2)
    knots = oracle.grid_knots(0.0, 0.0, n - 1.0, n - 1.0, 3, 3)
    phi = _kernels(beta, n_max)
    theta = _const_theta(phi.shape[0], knots.shape[0])
    kern0 = phi[0]
    diff = _correlated_difference(n, kern0, 1.0, 1.0, rng)
    var = np.full((n, n), 1.0, np.float32)
    out = noise.decorrelate(diff, knots, theta, beta, n_max, var, var, block=64)
    single = decorrelate_block(diff, kern0, 1.0, 1.0)
    c = slice(48, n - 48)
    assert abs(_autocorr_lag1(out["difference"][c, c])) < 0.05
    np.testing.assert_allclose(out["difference"][c, c], single[c, c], atol=0.15)
    sumk2 = float((kern0.astype(np.float64) ** 2).sum())
    np.testing.assert_allclose(out["variance"][c, c], 1.0 + sumk2, rtol=0.05)


def test_decorrelate_kernel_cache_matches_exact():
    n, beta, n_max = 384, 2.0, 1
    rng = np.random.default_rng(7)
    knots = oracle.grid_knots(0.0, 0.0, n - 1.0, n - 1.0, 3, 3)
    k = knots.shape[0]
    phi = _kernels(beta, n_max)
    theta = _const_theta(phi.shape[0], k)
    theta[2 * k - 2] = 0.6
    diff = _correlated_difference(n, phi[0], 1.0, 1.0, rng)
    var = np.full((n, n), 1.0, np.float32)
    auto = noise.decorrelate(diff, knots, theta, beta, n_max, var, var, block=64)
    exact = noise.decorrelate(diff, knots, theta, beta, n_max, var, var, block=64,
                              kernel_cell_blocks=1)
    c = slice(48, n - 48)
    d = auto["difference"][c, c] - exact["difference"][c, c]
    assert np.sqrt(np.mean(d ** 2)) / np.std(exact["difference"][c, c]) < 0.02
    np.testing.assert_allclose(auto["variance"][c, c], exact["variance"][c, c], rtol=0.05)


def test_block_lattice_and_weights():
    assert noise.origins(200, 64, 32) == [0, 32, 64, 96, 128, 136]
    assert noise.origins(192, 64, 32) == [0, 32, 64, 96, 128]
    assert noise.origins(50, 64, 32) == [0]
    # the separable weight sum of a full lattice is positive everywhere
    h = noise.hann1d(64)
    tot = np.zeros(200)
    for o in noise.origins(200, 64, 32):
        tot[o:o + 64] += h[: min(64, 200 - o)]
    assert (tot > 0).all()


def test_block_variance_is_subsampled_median():
    rng = np.random.default_rng(3)
    var = rng.uniform(0.5, 2.0, (300, 300)).astype(np.float32)
    var[10:20, 10:20] = 1e30
    var[50, 50] = np.nan
    var[60, 60] = -1.0
    got = noise.block_variance(var, 0, 0, 256)
    samp = var[0:256:4, 0:256:4].ravel()
    samp = samp[np.isfinite(samp) & (samp > 0)]
    assert got == float(np.sort(samp)[samp.size // 2])
    # partially outside the frame
    got = noise.block_variance(var, 200, 200, 256)
    samp = var[200:300:4, 200:300:4].ravel()
    assert got == float(np.sort(samp)[samp.size // 2])


def test_whiten_psf_preserves_unit_sum_and_changes_profile():
    beta, n_max, n = 2.0, 1, 128
    knots = oracle.grid_knots(0.0, 0.0, n - 1.0, n - 1.0, 3, 3)
    phi = _kernels(beta, n_max)
    theta = _const_theta(phi.shape[0], knots.shape[0])
    psf = np.clip(phi[0], 0.0, None).astype(np.float32)
    psf /= psf.sum()
    white = noise.whiten_score_psf(psf, knots, theta, beta, n_max,
                                   np.full((n, n), 1.0, np.float32),
                                   np.full((n, n), 1.0, np.float32), block=64)
    assert white.shape == psf.shape
    np.testing.assert_allclose(white.sum(), 1.0, atol=1e-5)
    assert not np.allclose(white, psf, atol=1e-4)


def _brute_matched(img, psf, var):
    h, w = img.shape
    ps = psf.shape[0]
    r = ps // 2
    pad = np.pad(img.astype(np.float64), r)
    acc = np.zeros((h, w))
    for ly in range(ps):
        for lx in range(ps):
            acc += float(psf[ly, lx]) * pad[ly:ly + h, lx:lx + w]
    s2 = float((psf.astype(np.float64) ** 2).sum())
    den = np.sqrt(np.maximum(var, 1e-30) * s2)
    return np.where(var > 0, acc / den, 0.0)


def test_matched_filter_matches_brute_force_and_peak():
    rng = np.random.default_rng(8)
    img = rng.standard_normal((70, 90)).astype(np.float32)
    var = rng.uniform(0.5, 3.0, img.shape).astype(np.float32)
    var[3, 4] = 0.0
    psf = _kernels(2.0, 0)[0]
    got = oracle.matched_filter(img, psf, var)
    np.testing.assert_allclose(got, _brute_matched(img, psf, var), atol=2e-5)
    assert got[3, 4] == 0.0
    # peak S/N identity of the reference test (test_noise.py:163-181)
    n = 128
    psf64 = psf.astype(np.float64)
    r = psf.shape[0] // 2
    src = np.zeros((n, n))
    src[n // 2 - r:n // 2 + r + 1, n // 2 - r:n // 2 + r + 1] += 50.0 * psf64
    score = oracle.matched_filter(src.astype(np.float32), psf, np.ones((n, n), np.float32))
    want = 50.0 * np.sqrt((psf64 ** 2).sum())
    assert abs(score[n // 2, n // 2] - want) / want < 1e-3
    assert score.argmax() == (n // 2) * n + n // 2


def test_matched_filter_unit_variance_noise():
    rng = np.random.default_rng(3)
    img = (np.sqrt(2.3) * rng.standard_normal((200, 200))).astype(np.float32)
    var = np.full((200, 200), 2.3, dtype=np.float32)
    score = oracle.matched_filter(img, _kernels(2.0, 0)[0], var)
    np.testing.assert_allclose(score.std(), 1.0, rtol=0.05)
    np.testing.assert_allclose(score.mean(), 0.0, atol=0.05)
