"""Pins the oracle decorrelation where the interior checks cannot: frames
smaller than a block (reflect padding, src/noise.cpp:393-406), the flush last
block, and the kernel-cell lattice (kgroup rule and representative block,
:296-343).

An independent numpy construction whitens every block with scipy's FFT and
blends with per-pixel Hann weight sums in float64; it meets the oracle within
2e-5 of the noise scale on frames shorter than a block in one axis, shorter
than half a block (the reflection clamps), and with kgroup 1-3.  Wrap or
zero padding, Python's round-half-even for kgroup, and the cell's first block
instead of its middle one each miss by far more.
"""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest
import scipy.fft as sfft

import oracle
from oracle import noise


def _lround_libm():
    lib = ctypes.CDLL(ctypes.util.find_library("m"))
    lib.lround.restype = ctypes.c_long
    lib.lround.argtypes = [ctypes.c_double]
    return lib.lround


def test_lround_equals_the_c_library():
    lr = _lround_libm()
    vals = [0.0, 0.49999999999999994, 0.5, 1.5, 2.5, 2.4999999999999996, 3.5000000000000004,
            7.0, 1e6 + 0.5, 12.25]
    rng = np.random.default_rng(0)
    vals += list(rng.uniform(0, 50, 2000)) + [k + 0.5 for k in range(40)]
    for v in vals:
        assert noise.lround(v) == lr(v), v
    assert round(2.5) == 2          # half-even: the slip this guards against


def _starts(n, block):
    if n <= block:
        return [0]
    out = list(range(0, n - block + 1, block // 2))
    if out[-1] != n - block:
        out.append(n - block)
    return out


def _gather(img, y0, x0, block, pad):
    h, w = img.shape
    if pad == "zero":
        p = np.zeros((h + 2 * block, w + 2 * block), np.float32)
        p[block:block + h, block:block + w] = img
        return p[block + y0:block + y0 + block, block + x0:block + x0 + block]
    if pad == "wrap":
        return np.take(np.take(img, np.arange(y0, y0 + block), 0, mode="wrap"),
                       np.arange(x0, x0 + block), 1, mode="wrap")
    # reflect about the edge (x = -1 -> 0, x = w -> w - 1), then clamp what is
    # still outside when the block exceeds twice the frame
    p = np.pad(img, ((0, max(0, y0 + block - h)), (0, max(0, x0 + block - w))), mode="symmetric")
    ys = np.minimum(np.arange(y0, y0 + block), np.where(np.arange(y0, y0 + block) < 2 * h,
                                                          np.arange(y0, y0 + block), 0))
    xs = np.minimum(np.arange(x0, x0 + block), np.where(np.arange(x0, x0 + block) < 2 * w,
                                                          np.arange(x0, x0 + block), 0))
    return p[np.ix_(ys, xs)]


def _median_level(v, x0, y0, block):
    h, w = v.shape
    step = max(1, block // 64)
    sub = v[y0:min(y0 + block, h):step, x0:min(x0 + block, w):step].ravel()
    sub = np.sort(sub[np.isfinite(sub) & (sub > 0)])
    return float(sub[sub.size // 2]) if sub.size else 0.0


def decorrelate_numpy(diff, knots, theta, beta, n_max, vs_img, vr_img, block, kcb=0,
                      pad="reflect", rounding=None, rep="middle"):
    h, w = diff.shape
    half = block // 2
    radius = oracle.default_radius(beta, n_max)
    g = oracle.gh_basis1d(beta, n_max, radius)
    phi = [np.outer(g[ny], g[nx]).astype(np.float32) for nx, ny in oracle.gh_orders(n_max)]
    xs0, ys0 = _starts(w, block), _starts(h, block)
    if kcb > 0:
        kg = kcb
    else:
        q = oracle.tps_min_spacing(knots) / 4.0 / half
        kg = max(1, (rounding or noise.lround)(q))
    nc = len(phi)
    coef = np.asarray(theta).reshape(nc + 1, -1)

    def cell_kernel(cix, ciy):
        pick = (lambda c, n: min(c * kg + kg // 2, n - 1)) if rep == "middle" else \
            (lambda c, n: min(c * kg, n - 1))
        cx = min(max(xs0[pick(cix, len(xs0))] + block / 2.0, 0.0), w - 1.0)
        cy = min(max(ys0[pick(ciy, len(ys0))] + block / 2.0, 0.0), h - 1.0)
        a = oracle.tps_design(knots, np.array([[cx, cy]]))[0] @ coef.T
        kern = sum(np.float64(np.float32(a[n])) * phi[n].astype(np.float64) for n in range(nc))
        emb = np.zeros((block, block))
        ks = kern.shape[0]
        idx = (np.arange(ks) - ks // 2) % block
        emb[np.ix_(idx, idx)] = kern
        return np.abs(sfft.rfft2(emb)) ** 2, float((kern ** 2).sum())

    hann = 0.5 * (1 - np.cos(2 * math.pi * (np.arange(block) + 0.5) / block))
    num = np.zeros((h, w))
    vnum = np.zeros((h, w))
    den = np.zeros((h, w))
    cache = {}
    for iy, y0 in enumerate(ys0):
        for ix, x0 in enumerate(xs0):
            key = (ix // kg, iy // kg)
            if key not in cache:
                cache[key] = cell_kernel(*key)
            power, sk2 = cache[key]
            vs = _median_level(vs_img, x0, y0, block)
            vr = _median_level(vr_img, x0, y0, block)
            d = vs + vr * power
            filt = np.where(d > 0, np.sqrt(max(vs + vr * sk2, 0.0)) / np.sqrt(np.where(d > 0, d, 1)),
                            0.0)
            buf = _gather(diff, y0, x0, block, pad).astype(np.float64)
            out = sfft.irfft2(sfft.rfft2(buf) * filt, s=(block, block))
            ny, nx = min(block, h - y0), min(block, w - x0)
            wt = np.outer(hann[:ny], hann[:nx])
            num[y0:y0 + ny, x0:x0 + nx] += wt * out[:ny, :nx]
            vnum[y0:y0 + ny, x0:x0 + nx] += wt * (vs + vr * sk2)
            den[y0:y0 + ny, x0:x0 + nx] += wt
    return np.where(den > 0, num / np.where(den > 0, den, 1), 0.0), \
        np.where(den > 0, vnum / np.where(den > 0, den, 1), 0.0), kg


def _case(h, w, nk, seed):
    rng = np.random.default_rng(seed)
    beta, n_max = 1.6, 2
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, nk, nk) if nk > 1 else np.zeros((1, 2))
    nc = (n_max + 1) * (n_max + 2) // 2
    k = knots.shape[0]
    theta = 0.3 * rng.standard_normal((nc + 1) * k)
    theta[(k - 3) if k >= 3 else 0] += 1.0
    diff = (5 * rng.standard_normal((h, w))).astype(np.float32)
    diff += np.linspace(0, 20, w, dtype=np.float32)[None, :]    # a ramp: edges matter
    vs = rng.uniform(20, 30, (h, w)).astype(np.float32)
    vr = rng.uniform(10, 40, (h, w)).astype(np.float32)
    return diff, knots, theta, beta, n_max, vs, vr


@pytest.mark.parametrize("h,w,block,nk,kcb", [
    (40, 150, 64, 3, 0),      # shorter than a block in y: reflect padding, flush block in x
    (20, 100, 64, 2, 1),      # shorter than half a block: the reflection clamps
    (150, 30, 32, 3, 0),      # tall and narrow
    (300, 300, 32, 3, 0),     # kgroup from the knot spacing (lround(150/4/16) = 2)
    (256, 264, 32, 3, 3),     # explicit kgroup 3, flush block
])
def test_decorrelate_edges_and_cells_match_numpy(h, w, block, nk, kcb):
    diff, knots, theta, beta, n_max, vs, vr = _case(h, w, nk, h + w)
    got = noise.decorrelate(diff, knots, theta, beta, n_max, vs, vr, block=block,
                            kernel_cell_blocks=kcb)
    want, wvar, kg = decorrelate_numpy(diff, knots, theta, beta, n_max, vs, vr, block, kcb)
    scale = float(np.std(want))
    err = np.abs(got["difference"] - want).max()
    assert err <= 2e-5 * scale, (err, scale)
    np.testing.assert_allclose(got["variance"], wvar, rtol=1e-6)
    if h < block or w < block:
        for pad in ("zero", "wrap"):
            alt = decorrelate_numpy(diff, knots, theta, beta, n_max, vs, vr, block, kcb, pad=pad)[0]
            assert np.abs(alt - want).max() > 1e-2 * scale, pad
    if kg > 1:
        alt = decorrelate_numpy(diff, knots, theta, beta, n_max, vs, vr, block, kcb,
                                rep="first")[0]
        assert np.abs(alt - want).max() > 1e-3 * scale


def test_kgroup_rule_rounds_halves_away_from_zero():
    """A knot spacing that makes (spacing / 4) / stride exactly 2.5: lround
    gives 3 cells, round-half-even would give 2, and the outputs differ."""
    h = w = 321                        # 3x3 knots 160 px apart: 160 / 4 / 16 = 2.5
    diff, knots, theta, beta, n_max, vs, vr = _case(h, w, 3, 9)
    assert oracle.tps_min_spacing(knots) / 4.0 / 16 == 2.5
    assert noise.kernel_cell_group(knots, 16, 0) == 3
    got = noise.decorrelate(diff, knots, theta, beta, n_max, vs, vr, block=32)
    want = decorrelate_numpy(diff, knots, theta, beta, n_max, vs, vr, 32)[0]
    even = decorrelate_numpy(diff, knots, theta, beta, n_max, vs, vr, 32, rounding=round)[0]
    scale = float(np.std(want))
    assert np.abs(got["difference"] - want).max() <= 2e-5 * scale
    assert np.abs(even - want).max() > 1e-3 * scale
