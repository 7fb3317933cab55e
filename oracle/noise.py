"""Oracle of delta's noise decorrelation -- TEST INFRASTRUCTURE ONLY.

A restatement of src/noise.cpp in numpy, with scipy.fft for the FFTs.
scipy.fft is PocketFFT, the library the reference vendors
(extern/pocketfft, noise.cpp:20-62), run in single precision as the
reference does (float32 in, complex64 half spectrum, unnormalised both ways).

  origins / kernel_cell_group     noise.cpp:256-269, 310-324
  local_kernel                    noise.cpp:167-189
  kernel_power                    noise.cpp:65-84
  block_variance                  noise.cpp:200-224
  decorrelation_filter            noise.cpp:98-117
  apply_filter_fft                noise.cpp:119-137
  decorrelate_block               bindings.cpp:725-745
  decorrelation_kernel            noise.cpp:139-161, bindings.cpp:748-763
  decorrelate                     noise.cpp:228-483
  whiten_psf / whiten_score_psf   noise.cpp:485-533, bindings.cpp:796-874

Arithmetic that the reference does in double (weights, Phi, the noise
levels) is done in double here; image arithmetic is float32 as in the
reference.  The blend keeps the reference's block order (9-colour classes
by (ix % 3, iy % 3), blocks in raster order within a class), so the float
accumulation order is the reference's.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.fft as sfft

from . import default_radius, gh_basis1d, gh_orders, tps_design, tps_min_spacing


def hann1d(n: int) -> np.ndarray:
    i = np.arange(n, dtype=np.float64)
    return 0.5 * (1.0 - np.cos(2.0 * math.pi * (i + 0.5) / n))


def origins(extent: int, blk: int, strd: int) -> list[int]:
    if extent <= blk:
        return [0]
    o = list(range(0, extent - blk + 1, strd))
    if o[-1] + blk < extent:
        o.append(extent - blk)
    return o


def kernel_cell_group(knots, stride: int, kernel_cell_blocks: int) -> int:
    if kernel_cell_blocks > 0:
        return kernel_cell_blocks
    spacing = tps_min_spacing(np.asarray(knots, dtype=np.float64))
    if spacing > 0.0 and stride > 0:
        cell_px = spacing / 4.0
        return max(1, lround(cell_px / stride))
    return 1


def lround(x: float) -> int:
    """std::lround of a non-negative double: halves away from zero, exactly
    (floor(x + 0.5) rounds 0.49999999999999994 up, because x + 0.5 rounds)."""
    f = math.floor(x)
    return int(f) + (1 if x - f >= 0.5 else 0)


def phi_stack(beta: float, n_max: int, radius: int) -> np.ndarray:
    """kernel2d(c) for every component (basis.cpp:66-79), float32 (nc, ks, ks)."""
    b = gh_basis1d(beta, n_max, radius)
    return np.stack([(b[ny][:, None] * b[nx][None, :]).astype(np.float32)
                     for nx, ny in gh_orders(n_max)])


def local_kernel(knots, theta, phi: np.ndarray, cx: float, cy: float) -> np.ndarray:
    """K = sum_n a_n(cx, cy) phi_n, a_n rounded to float; FMA accumulation."""
    knots = np.asarray(knots, dtype=np.float64)
    k = knots.shape[0]
    nc = phi.shape[0]
    c = np.asarray(theta, dtype=np.float64).reshape(nc + 1, k).T   # column-major k x (nc+1)
    a = tps_design(knots, np.array([[cx, cy]], dtype=np.float64)) @ c
    kern = np.zeros(phi.shape[1:], dtype=np.float32)
    for n in range(nc):
        an = np.float32(a[0, n])
        kern = (np.float64(an) * phi[n].astype(np.float64) + kern.astype(np.float64)).astype(np.float32)
    return kern


def _embed(kern: np.ndarray, n: int) -> np.ndarray:
    ks = kern.shape[0]
    r = ks // 2
    g = np.zeros((n, n), dtype=np.float32)
    idx = (np.arange(ks) - r) % n
    g[np.ix_(idx, idx)] = kern
    return g


def kernel_power(kern: np.ndarray, n: int) -> np.ndarray:
    f = sfft.rfft2(_embed(kern, n))
    re = f.real.astype(np.float64)
    im = f.imag.astype(np.float64)
    return re * re + im * im


def block_variance(var: np.ndarray, bx: int, by: int, block: int) -> float:
    h, w = var.shape
    step = max(1, block // 64)
    ys = [y for y in range(by, by + block, step) if 0 <= y < h]
    xs = [x for x in range(bx, bx + block, step) if 0 <= x < w]
    if not ys or not xs:
        return 0.0
    v = var[np.ix_(ys, xs)].ravel()
    v = v[np.isfinite(v) & (v > 0.0)]
    if v.size == 0:
        return 0.0
    mid = v.size // 2
    return float(np.partition(v, mid)[mid])


def filter_from_power(power: np.ndarray, sumk2: float, vs: float, vr: float) -> np.ndarray:
    cnum = math.sqrt(max(vs + vr * sumk2, 0.0))
    denom = vs + vr * power
    out = np.zeros(power.shape, dtype=np.float32)
    pos = denom > 0.0
    out[pos] = (cnum / np.sqrt(denom[pos])).astype(np.float32)
    return out


def decorrelation_filter(kern, vs, vr, n):
    kern = np.asarray(kern, dtype=np.float32)
    sumk2 = float((kern.astype(np.float64) ** 2).sum())
    return filter_from_power(kernel_power(kern, n), sumk2, vs, vr)


def apply_filter_fft(block: np.ndarray, filt: np.ndarray) -> np.ndarray:
    n = block.shape[0]
    spec = sfft.rfft2(np.asarray(block, dtype=np.float32))
    spec = spec * filt
    return sfft.irfft2(spec, s=(n, n), norm="forward") * np.float32(1.0 / (n * n))


def decorrelate_block(image, kernel, vs, vr) -> np.ndarray:
    """One square block whitened with a constant kernel (bindings.cpp:725-745)."""
    img = np.asarray(image, dtype=np.float32)
    return apply_filter_fft(img, decorrelation_filter(kernel, vs, vr, img.shape[0]))


def decorrelation_kernel(kernel, vs, vr, n) -> np.ndarray:
    """ifft(Phi) / n^2, fft-shifted so the zero-phase kernel is centred
    (decorrelation_kernel_image, noise.cpp:139-161)."""
    filt = decorrelation_filter(kernel, vs, vr, n)
    raw = sfft.irfft2(filt.astype(np.complex64), s=(n, n), norm="forward")
    raw = raw.astype(np.float32) * np.float32(1.0 / (n * n))
    return np.roll(raw, (-(n // 2), -(n // 2)), axis=(0, 1))


def _reflect(idx: np.ndarray, extent: int) -> np.ndarray:
    s = np.where(idx < 0, -idx - 1, idx)
    s = np.where(s >= extent, 2 * extent - s - 1, s)
    return np.clip(s, 0, extent - 1)


def decorrelate(difference, knots, theta, beta, n_max, var_science, var_reference,
                block=256, radius=0, kernel_cell_blocks=0):
    diff = np.asarray(difference, dtype=np.float32)
    var_s = np.asarray(var_science, dtype=np.float32)
    var_r = np.asarray(var_reference, dtype=np.float32)
    h, w = diff.shape
    if block < 8 or (block & (block - 1)) != 0:
        raise RuntimeError("decorrelate: block must be a power of two >= 8")
    radius = radius if radius > 0 else default_radius(beta, n_max)
    phi = phi_stack(beta, n_max, radius)
    stride = block // 2
    oxs, oys = origins(w, block, stride), origins(h, block, stride)
    nox, noy = len(oxs), len(oys)
    kgroup = kernel_cell_group(knots, stride, kernel_cell_blocks)
    ncellx = (nox + kgroup - 1) // kgroup
    ncelly = (noy + kgroup - 1) // kgroup
    cell_power, cell_sumk2 = [], []
    for ci in range(ncellx * ncelly):
        cix, ciy = ci % ncellx, ci // ncellx
        repx = min(cix * kgroup + kgroup // 2, nox - 1)
        repy = min(ciy * kgroup + kgroup // 2, noy - 1)
        cx = min(max(oxs[repx] + block / 2.0, 0.0), w - 1.0)
        cy = min(max(oys[repy] + block / 2.0, 0.0), h - 1.0)
        kern = local_kernel(knots, theta, phi, cx, cy)
        cell_power.append(kernel_power(kern, block))
        cell_sumk2.append(float((kern.astype(np.float64) ** 2).sum()))

    hwin = hann1d(block)
    wx_total = np.zeros(w)
    wy_total = np.zeros(h)
    for bx in oxs:
        sl = slice(bx, min(bx + block, w))
        wx_total[sl] += hwin[: sl.stop - bx]
    for by in oys:
        sl = slice(by, min(by + block, h))
        wy_total[sl] += hwin[: sl.stop - by]

    acc = np.zeros((h, w), dtype=np.float32)
    var_acc = np.zeros((h, w), dtype=np.float32)
    inv_n2 = np.float32(1.0) / np.float32(block * block)
    colours = [[] for _ in range(9)]
    for iy in range(noy):
        for ix in range(nox):
            cell = (iy // kgroup) * ncellx + (ix // kgroup)
            colours[(ix % 3) * 3 + (iy % 3)].append((oxs[ix], oys[iy], cell))
    ar = np.arange(block)
    for blocks in colours:
        for bx, by, cell in blocks:
            sy = _reflect(by + ar, h)
            sx = _reflect(bx + ar, w)
            buf = diff[np.ix_(sy, sx)]
            vs = block_variance(var_s, bx, by, block)
            vr = block_variance(var_r, bx, by, block)
            filt = filter_from_power(cell_power[cell], cell_sumk2[cell], vs, vr)
            spec = sfft.rfft2(buf) * filt
            real = sfft.irfft2(spec, s=(block, block), norm="forward")
            cnum = math.sqrt(max(vs + vr * cell_sumk2[cell], 0.0))
            num = cnum * cnum
            y0, y1 = by, min(by + block, h)
            x0, x1 = bx, min(bx + block, w)
            wy_sig = hwin[: y1 - y0] * np.float64(inv_n2)
            wxs = hwin[: x1 - x0]
            fr = real[: y1 - y0, : x1 - x0].astype(np.float64)
            a = acc[y0:y1, x0:x1].astype(np.float64)
            acc[y0:y1, x0:x1] = (a + (wy_sig[:, None] * wxs[None, :]) * fr).astype(np.float32)
            v = var_acc[y0:y1, x0:x1].astype(np.float64)
            var_acc[y0:y1, x0:x1] = (v + (hwin[: y1 - y0][:, None] * wxs[None, :]) * num).astype(np.float32)
    wsum = wy_total[:, None] * wx_total[None, :]
    pos = wsum > 0.0
    out = np.zeros((h, w), dtype=np.float32)
    vout = np.zeros((h, w), dtype=np.float32)
    out[pos] = (acc.astype(np.float64)[pos] / wsum[pos]).astype(np.float32)
    vout[pos] = (var_acc.astype(np.float64)[pos] / wsum[pos]).astype(np.float32)
    return {"difference": out, "variance": vout}


def whiten_psf(psf, kern, vs, vr, n):
    psf = np.asarray(psf, dtype=np.float32)
    ps = psf.shape[0]
    if ps > n:
        raise RuntimeError("whiten_psf: psf_size must be <= FFT size n")
    filt = decorrelation_filter(kern, vs, vr, n)
    spec = sfft.rfft2(_embed(psf, n)) * filt
    real = sfft.irfft2(spec, s=(n, n), norm="forward")
    inv = np.float32(1.0) / np.float32(n) / np.float32(n)
    r = ps // 2
    idx = (np.arange(ps) - r) % n
    out = (real[np.ix_(idx, idx)] * inv).astype(np.float32)
    s = float(out.astype(np.float64).sum())
    if s > 0.0:
        out = (out * np.float32(1.0 / s)).astype(np.float32)
    return out


def whiten_score_psf(psf, knots, theta, beta, n_max, var_science, var_reference, block,
                     radius=0):
    psf = np.asarray(psf, dtype=np.float32)
    if psf.ndim != 2 or psf.shape[0] != psf.shape[1]:
        raise RuntimeError("whiten_score_psf: psf must be square")
    if block < psf.shape[0]:
        raise RuntimeError("whiten_score_psf: block must be >= psf side")
    var_s = np.asarray(var_science, dtype=np.float32)
    var_r = np.asarray(var_reference, dtype=np.float32)
    if var_s.shape != var_r.shape:
        raise RuntimeError("whiten_score_psf: variance shapes disagree")
    h, w = var_s.shape
    radius = radius if radius > 0 else default_radius(beta, n_max)
    phi = phi_stack(beta, n_max, radius)
    kern = local_kernel(knots, theta, phi, 0.5 * (w - 1), 0.5 * (h - 1))
    bx = max(0, w // 2 - block // 2)
    by = max(0, h // 2 - block // 2)
    vs = block_variance(var_s, bx, by, block)
    vr = block_variance(var_r, bx, by, block)
    return whiten_psf(psf, kern, vs, vr, block)
