"""Deterministic synthetic image pairs for the benchmark configurations.

Star fields as in the reference's validation helpers (python/delta/validation.py:
gaussian_psf :17-24, render_stars :34-55), extended to elliptical, spatially
varying science PSFs.  All inputs come from one numpy PCG64 stream.

Noise model: electrons = gain * ADU; Poisson in electrons plus Gaussian read
noise.  The reference's variance synthesis takes read noise in ADU
(python/delta/_inputs.py:75-86), so 4 e- at gain 1.5 is 2.667 ADU; the
variance maps here are the true expectation (sky + signal)/gain + rn_adu^2.
"""
from __future__ import annotations

import dataclasses

import numpy as np

FWHM_PER_SIGMA = 2.354820045030949


@dataclasses.dataclass
class PairConfig:
    size: int = 4096
    n_stars: int = 8000
    flux_min: float = 1e2
    flux_max: float = 1e6
    ref_fwhm: float = 2.5
    sci_fwhm0: float = 3.0          # science FWHM at x = 0
    sci_fwhm1: float = 4.5          # science FWHM at x = W-1
    ellip_max: float = 0.1          # ellipticity at y = H-1 (0 at y = 0)
    position_angle_deg: float = 30.0
    sky: float = 1000.0
    gain: float = 1.5
    read_noise_e: float = 4.0
    seed: int = 1
    bad_fraction: float = 0.0       # raw bad-column fraction of the science mask
    n_transients: int = 0
    transient_snr: tuple = (5.0, 50.0)


def _stamp(sig_x, sig_y, pa, dx, dy, rad):
    yy, xx = np.mgrid[-rad:rad + 1, -rad:rad + 1].astype(np.float64)
    xx = xx - dx
    yy = yy - dy
    c, s = np.cos(pa), np.sin(pa)
    u = c * xx + s * yy
    v = -s * xx + c * yy
    g = np.exp(-0.5 * ((u / sig_x) ** 2 + (v / sig_y) ** 2))
    return g / g.sum()


def _render(img, x, y, flux, sig_x, sig_y, pa):
    h, w = img.shape
    rad = int(np.ceil(5.0 * max(sig_x, sig_y)))
    ix, iy = int(np.floor(x + 0.5)), int(np.floor(y + 0.5))
    st = flux * _stamp(sig_x, sig_y, pa, x - ix, y - iy, rad)
    x0, x1 = max(ix - rad, 0), min(ix + rad + 1, w)
    y0, y1 = max(iy - rad, 0), min(iy + rad + 1, h)
    if x0 >= x1 or y0 >= y1:
        return
    img[y0:y1, x0:x1] += st[y0 - (iy - rad):y1 - (iy - rad), x0 - (ix - rad):x1 - (ix - rad)]


def make_pair(cfg: PairConfig):
    """Return dict with science, reference, science_var, reference_var (float32),
    the star list and, when requested, masks and transient positions."""
    rng = np.random.default_rng(cfg.seed)
    n = cfg.size
    xs = rng.uniform(0, n - 1, cfg.n_stars)
    ys = rng.uniform(0, n - 1, cfg.n_stars)
    flux = np.exp(rng.uniform(np.log(cfg.flux_min), np.log(cfg.flux_max), cfg.n_stars))
    ref_sig = cfg.ref_fwhm / FWHM_PER_SIGMA
    pa = np.deg2rad(cfg.position_angle_deg)
    ref_model = np.zeros((n, n))
    sci_model = np.zeros((n, n))
    for x, y, f in zip(xs, ys, flux):
        _render(ref_model, x, y, f, ref_sig, ref_sig, 0.0)
        fw = cfg.sci_fwhm0 + (cfg.sci_fwhm1 - cfg.sci_fwhm0) * x / (n - 1)
        e = cfg.ellip_max * y / (n - 1)
        sig = fw / FWHM_PER_SIGMA
        _render(sci_model, x, y, f, sig / np.sqrt(1.0 - e), sig * np.sqrt(1.0 - e), pa)
    trans = None
    if cfg.n_transients:
        tx = rng.uniform(32, n - 33, cfg.n_transients)
        ty = rng.uniform(32, n - 33, cfg.n_transients)
        snr = rng.uniform(*cfg.transient_snr, cfg.n_transients)
        sig_t = cfg.sci_fwhm1 / FWHM_PER_SIGMA
        # matched-filter S/N of a Gaussian of width s on flat variance V is
        # F / sqrt(4 pi s^2 V); invert for the flux
        var_sky = cfg.sky / cfg.gain + (cfg.read_noise_e / cfg.gain) ** 2
        tflux = snr * np.sqrt(4.0 * np.pi * sig_t ** 2 * 2.0 * var_sky)
        for x, y, f in zip(tx, ty, tflux):
            _render(sci_model, x, y, f, sig_t, sig_t, 0.0)
        trans = np.column_stack([tx, ty, tflux, snr])
    rn_adu = cfg.read_noise_e / cfg.gain
    out = {}
    for name, model in (("reference", ref_model), ("science", sci_model)):
        mean = cfg.sky + model
        e = rng.poisson(np.maximum(mean, 0.0) * cfg.gain).astype(np.float64)
        img = e / cfg.gain + rng.normal(0.0, rn_adu, mean.shape)
        out[name] = img.astype(np.float32)
        out[name + "_var"] = (mean / cfg.gain + rn_adu ** 2).astype(np.float32)
    if cfg.bad_fraction > 0:
        ncols = max(1, int(round(cfg.bad_fraction * n)))
        cols = rng.choice(n, ncols, replace=False)
        m = np.zeros((n, n), np.uint8)
        m[:, cols] = 1
        out["science_mask"] = m
    out["stars"] = np.column_stack([xs, ys, flux])
    out["transients"] = trans
    return out


def pick_stamps(stars, size, n_max_stamps=200, stamp_radius=15, edge=None, flux_max=None):
    """Stamp centres given to the fit: the brightest stars that sit a full
    stamp + kernel away from the edge, are isolated (no brighter star within
    two stamp widths) and are below `flux_max`.  Stands in for the reference's
    catalogue-driven selection (select_stamps_from_catalog, src/detect.cpp:289-314)."""
    edge = 2 * stamp_radius + 1 if edge is None else edge
    xs, ys, fl = stars[:, 0], stars[:, 1], stars[:, 2]
    order = np.argsort(-fl)
    chosen = []
    for i in order:
        x, y = xs[i], ys[i]
        if x < edge or y < edge or x > size - 1 - edge or y > size - 1 - edge:
            continue
        if flux_max is not None and fl[i] > flux_max:
            continue
        near = (np.abs(xs - x) < 2 * stamp_radius) & (np.abs(ys - y) < 2 * stamp_radius)
        near[i] = False
        if np.any(fl[near] > 0.05 * fl[i]):
            continue
        chosen.append(i)
        if len(chosen) >= n_max_stamps:
            break
    idx = np.array(chosen, dtype=int)
    return (np.round(xs[idx]).astype(np.int32), np.round(ys[idx]).astype(np.int32))


# the reference scales a FWHM by this rounded reciprocal (python/delta/pipeline.py:23);
# dividing by FWHM_PER_SIGMA instead differs in the last bit for about a third of inputs
SIGMA_PER_FWHM = 1.0 / FWHM_PER_SIGMA


def default_beta(fwhm_a: float, fwhm_b: float) -> float:
    """Kernel scale heuristic of the reference (python/delta/pipeline.py:267-282):
    0.8 x the width of the Gaussian that broadens the narrower PSF into the
    broader one, floored at 0.5 px; 2.0 without a usable FWHM.  The operations
    are the reference's, in its order, so beta is bit-equal to its value."""
    fw = [f for f in (fwhm_a, fwhm_b) if np.isfinite(f) and f > 0]
    if not fw:
        return 2.0
    broad, narrow = max(fw), min(fw)
    d = (broad * SIGMA_PER_FWHM) ** 2 - (narrow * SIGMA_PER_FWHM) ** 2
    ks = np.sqrt(d) if d > 0 else 0.0
    return float(max(0.8 * ks, 0.5))


def smooth_theta(knots, nc, rng, scale=0.05):
    """A smooth, plausible coefficient vector: dominant Gaussian term near 1,
    small higher-order and background terms (for subtraction-only tests)."""
    k = knots.shape[0]
    th = scale * rng.standard_normal((nc + 1) * k) / np.arange(1, nc + 2).repeat(k)
    if k >= 3:
        th[k - 3] += 1.0  # affine intercept of component 0
    else:
        th[0] += 1.0
    return th


# Generating kernel of the differentiation workload: Gauss-Hermite order 4
# (15 components, ordered by total order then ny) at beta = 1.3 px, shifted by
# (0.3, -0.2) px; scaled to unit sum by make_kernel_pair.
KERNEL_TRUE_COEFFS = (1.0, 0.0, 0.0, -0.10, 0.03, -0.08, 0.02, 0.0, 0.0, -0.02,
                      0.05, 0.0, 0.04, 0.0, 0.05)
KERNEL_TRUE_BETA = 1.3
KERNEL_TRUE_SHIFT = (0.3, -0.2)


def kernel_true_params(n_max=6, radius=15):
    """params = [a_0..a_{nc-1}, b, beta, dx, dy] of the generating kernel in an
    order-n_max basis (a_n = 0 beyond order 4), normalised to unit sum."""
    from .adjoint import shifted_kernel

    nc = (n_max + 1) * (n_max + 2) // 2
    p = np.zeros(nc + 4)
    m = min(nc, len(KERNEL_TRUE_COEFFS))
    p[:m] = KERNEL_TRUE_COEFFS[:m]
    p[nc + 1] = KERNEL_TRUE_BETA
    p[nc + 2], p[nc + 3] = KERNEL_TRUE_SHIFT
    p[:nc] /= shifted_kernel(p, n_max, radius).sum()
    return p


def make_kernel_pair(size=2048, n_stars=None, seed=3, n_max=6, radius=15, params=None,
                     flux_min=1e2, flux_max=1e6, ref_fwhm=2.5, sky=1000.0, gain=1.5,
                     read_noise_e=4.0):
    """Image pair whose science PSF is the reference PSF convolved with a known
    shifted Gauss-Hermite kernel.  Stars are rendered on a canvas padded by
    the kernel radius, convolved there (FFT), then cropped, so the science
    frame carries no edge loss.  Returns science, reference, their variance
    maps, the true params and the star list."""
    from scipy.signal import fftconvolve

    from .adjoint import shifted_kernel

    rng = np.random.default_rng(seed)
    if n_stars is None:
        n_stars = int(round(8000 * (size / 4096) ** 2))
    p = kernel_true_params(n_max, radius) if params is None else np.asarray(params, np.float64)
    k = shifted_kernel(p, n_max, radius)
    pad = radius + 16
    m = size + 2 * pad
    xs = rng.uniform(0, m - 1, n_stars)
    ys = rng.uniform(0, m - 1, n_stars)
    flux = np.exp(rng.uniform(np.log(flux_min), np.log(flux_max), n_stars))
    sig = ref_fwhm / FWHM_PER_SIGMA
    ref_model = np.zeros((m, m))
    for x, y, f in zip(xs, ys, flux):
        _render(ref_model, x, y, f, sig, sig, 0.0)
    sci_model = fftconvolve(ref_model, k, mode="same")
    crop = np.s_[pad:pad + size, pad:pad + size]
    rn_adu = read_noise_e / gain
    out = {}
    for name, model in (("reference", ref_model[crop]), ("science", sci_model[crop])):
        mean = sky + model
        e = rng.poisson(np.maximum(mean, 0.0) * gain).astype(np.float64)
        out[name] = (e / gain + rng.normal(0.0, rn_adu, mean.shape)).astype(np.float32)
        out[name + "_var"] = (mean / gain + rn_adu ** 2).astype(np.float32)
    inside = (xs >= pad) & (xs < pad + size) & (ys >= pad) & (ys < pad + size)
    out["stars"] = np.column_stack([xs[inside] - pad, ys[inside] - pad, flux[inside]])
    out["params"] = p
    return out
