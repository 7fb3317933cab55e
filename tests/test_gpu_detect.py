"""GPU parity of the noise decorrelation, the matched filter and the labelling.

Bars:
  * matched filter: bit-identical to the oracle (same float operation order:
    FMA taps ly-outer/lx-inner, zero taps skipped, IEEE sqrt and division);
  * labels: bit-identical (raster-first numbering, count);
  * decorrelate / whitened PSF: the FFTs differ from PocketFFT's only in
    rounding, so the whitened difference is held to 2e-5 of the frame's
    robust scale (max abs error) and the post-whitening variance to 1e-6
    relative; the whitened PSF to 1e-6 absolute (entries ~1e-2).
"""
import numpy as np
import pytest

import oracle
from oracle import noise

pytestmark = pytest.mark.gpu


def _core():
    from delta_amd import _core
    return _core


def _psf(beta, n_max, comp=0):
    return noise.phi_stack(beta, n_max, oracle.default_radius(beta, n_max))[comp]


# ------------------------------------------------------------------ matched filter
@pytest.mark.parametrize("h,w,ps", [(70, 90, 13), (257, 300, 17), (64, 64, 4), (40, 130, 37),
                                    (5, 7, 9)])
def test_matched_filter_bit_exact(gpu, h, w, ps):
    rng = np.random.default_rng(h * 1000 + w + ps)
    img = rng.standard_normal((h, w)).astype(np.float32)
    var = rng.uniform(0.3, 3.0, (h, w)).astype(np.float32)
    var[rng.random((h, w)) < 0.05] = 0.0
    psf = rng.random((ps, ps)).astype(np.float32)
    psf[rng.random((ps, ps)) < 0.2] = 0.0   # zero taps are skipped
    print(f"matched_filter {h}x{w} psf {ps}", flush=True)
    got = _core().matched_filter(img, psf, var)
    want = oracle.matched_filter(img, psf, var)
    np.testing.assert_array_equal(got, want)


def test_matched_filter_device_entry(gpu):
    import torch

    rng = np.random.default_rng(2)
    img = rng.standard_normal((300, 200)).astype(np.float32)
    var = rng.uniform(0.5, 2.0, img.shape).astype(np.float32)
    psf = _psf(2.0, 0)
    dev = torch.device("cuda:0")
    got = _core().matched_filter_device(torch.from_numpy(img).to(dev), torch.from_numpy(psf).to(dev),
                                        torch.from_numpy(var).to(dev))
    np.testing.assert_array_equal(got.cpu().numpy(), oracle.matched_filter(img, psf, var))


def test_matched_filter_errors(gpu):
    img = np.zeros((10, 10), np.float32)
    with pytest.raises(RuntimeError, match="psf must be square"):
        _core().matched_filter(img, np.zeros((3, 5), np.float32), img)
    with pytest.raises(RuntimeError, match="variance must match image shape"):
        _core().matched_filter(img, np.zeros((3, 3), np.float32), np.zeros((9, 10), np.float32))


# ------------------------------------------------------------------ labels
@pytest.mark.parametrize("h,w,thr,density", [(61, 83, 1.0, 0.0), (400, 300, 0.0, 0.0),
                                             (1000, 1500, 0.4, 0.05), (1, 1, 0.0, 0.0),
                                             (2, 4096, -1.0, 0.0), (3000, 7, 0.2, 0.0)])
def test_labels_bit_exact(gpu, h, w, thr, density):
    rng = np.random.default_rng(h + 7 * w)
    s = rng.standard_normal((h, w)).astype(np.float32)
    s[rng.random((h, w)) < 0.01] = np.nan
    m = (rng.random((h, w)) < density).astype(np.uint8) if density > 0 else None
    for pol in (1, -1):
        print(f"labels {h}x{w} thr {thr} polarity {pol}", flush=True)
        got, n = _core().label_components(s, thr, pol, m)
        want, nw = oracle.label(s, thr, pol, m)
        assert n == nw
        np.testing.assert_array_equal(got, want)


def test_labels_edge_cases(gpu):
    z = np.zeros((33, 47), np.float32)
    lab, n = _core().label_components(z, 0.5)
    assert n == 0 and not lab.any()
    lab, n = _core().label_components(z + 1.0, 0.5)
    assert n == 1 and (lab == 1).all()
    # a snake: one component whose arms meet only at the far end
    s = np.zeros((9, 9), np.float32)
    s[0, :] = s[:, 8] = s[8, :] = s[2:, 0] = s[2, :7] = 1.0
    got, n = _core().label_components(s, 0.5)
    want, nw = oracle.label(s, 0.5)
    assert n == nw == 1
    np.testing.assert_array_equal(got, want)


def test_labels_device_entry(gpu):
    import torch

    rng = np.random.default_rng(9)
    s = rng.standard_normal((512, 768)).astype(np.float32)
    lab, n = _core().label_components_device(torch.from_numpy(s).to("cuda:0"), 0.8)
    want, nw = oracle.label(s, 0.8)
    assert n == nw
    np.testing.assert_array_equal(lab.cpu().numpy(), want)


# ------------------------------------------------------------------ decorrelation
def _correlated(h, w, kern, rng, vs=1.0, vr=1.0):
    ns = np.sqrt(vs) * rng.standard_normal((h, w))
    nr = np.sqrt(vr) * rng.standard_normal((h, w))
    ks = kern.shape[0]
    r = ks // 2
    pad = np.pad(nr, r, mode="wrap")
    conv = np.zeros((h, w))
    kf = kern[::-1, ::-1].astype(np.float64)
    for i in range(ks):
        for j in range(ks):
            conv += kf[i, j] * pad[i:i + h, j:j + w]
    return (ns + conv).astype(np.float32)


def _theta(nc, k, rng, strength):
    theta = np.zeros((nc + 1) * k)
    theta[k - 3] = 1.0
    theta += strength * rng.standard_normal(theta.size) * 0.05
    return theta


@pytest.mark.parametrize("h,w,block,n_max,kcb", [(192, 192, 64, 1, 0), (320, 384, 64, 2, 0),
                                                 (200, 136, 32, 1, 1), (96, 80, 16, 1, 0),
                                                 (512, 512, 256, 2, 0), (300, 260, 128, 1, 2),
                                                 (40, 150, 64, 1, 0), (20, 100, 64, 2, 1),
                                                 (150, 30, 32, 1, 0), (300, 700, 256, 2, 0),
                                                 (200, 180, 256, 1, 0), (130, 1000, 256, 2, 1),
                                                 (140, 260, 512, 1, 0)])
def test_decorrelate_matches_oracle(gpu, h, w, block, n_max, kcb):
    rng = np.random.default_rng(h + w + block)
    beta = 2.0
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    phi = _psf(beta, n_max)
    nc = (n_max + 1) * (n_max + 2) // 2
    theta = _theta(nc, knots.shape[0], rng, 1.0)
    diff = _correlated(h, w, phi, rng)
    vs = rng.uniform(0.8, 1.2, (h, w)).astype(np.float32)
    vr = rng.uniform(0.8, 1.2, (h, w)).astype(np.float32)
    vs[rng.random((h, w)) < 0.02] = 1e30      # sentinel fill
    vr[rng.random((h, w)) < 0.02] = np.nan
    print(f"decorrelate {h}x{w} block {block} n_max {n_max} cells {kcb}", flush=True)
    got = _core().decorrelate(diff, knots, theta, beta, n_max, vs, vr, block=block,
                              kernel_cell_blocks=kcb)
    want = noise.decorrelate(diff, knots, theta, beta, n_max, vs, vr, block=block,
                             kernel_cell_blocks=kcb)
    scale = float(np.std(want["difference"]))
    err = np.abs(got["difference"] - want["difference"]).max()
    print(f"  max |error| {err:.3e} (scale {scale:.3f})")
    assert err <= 2e-5 * scale
    np.testing.assert_allclose(got["variance"], want["variance"], rtol=1e-6)


def test_decorrelate_errors(gpu):
    z = np.zeros((32, 32), np.float32)
    knots = oracle.grid_knots(0.0, 0.0, 31.0, 31.0, 3, 3)
    theta = np.zeros(9 * 2)
    with pytest.raises(RuntimeError, match="power of two"):
        _core().decorrelate(z, knots, theta, 1.0, 0, z, z, block=48)
    with pytest.raises(RuntimeError, match="power of two"):
        _core().decorrelate(z, knots, theta, 1.0, 0, z, z, block=4)


@pytest.mark.parametrize("ps,block,h,w", [(17, 64, 128, 128), (13, 256, 300, 260), (9, 16, 50, 40)])
def test_whiten_score_psf_matches_oracle(gpu, ps, block, h, w):
    rng = np.random.default_rng(ps + block)
    beta, n_max = 2.0, 1
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    theta = _theta(3, knots.shape[0], rng, 1.0)
    psf = rng.random((ps, ps)).astype(np.float32)
    psf /= psf.sum()
    vs = rng.uniform(0.5, 2.0, (h, w)).astype(np.float32)
    vr = rng.uniform(0.5, 2.0, (h, w)).astype(np.float32)
    print(f"whiten_score_psf psf {ps} block {block}", flush=True)
    got = _core().whiten_score_psf(psf, knots, theta, beta, n_max, vs, vr, block)
    want = noise.whiten_score_psf(psf, knots, theta, beta, n_max, vs, vr, block)
    np.testing.assert_allclose(got, want, atol=1e-6)
    np.testing.assert_allclose(got.sum(), 1.0, atol=1e-5)
