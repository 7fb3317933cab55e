"""Synthetic star fields for the stamp-selection tests (CPU and GPU).

``add_star`` and ``peak_flux`` restate the helpers of the reference's
tests/test_detect.py: a unit-sum 2-D Gaussian over the whole frame scaled to
the total flux, and the flux that gives a chosen peak.
"""
import numpy as np

FWHM_PER_SIGMA = 2.354820045030949


def add_star(img, x0, y0, sigma, flux):
    h, w = img.shape
    yy, xx = np.mgrid[0:h, 0:w]
    g = np.exp(-((xx - x0) ** 2 + (yy - y0) ** 2) / (2.0 * sigma ** 2))
    img += (flux * g / g.sum()).astype(img.dtype)


def peak_flux(peak, sigma):
    return peak * 2.0 * np.pi * sigma ** 2


def add_star_local(img, x0, y0, sigma, peak):
    """A Gaussian of the given peak rendered in a +-6 sigma box (fast for big frames)."""
    h, w = img.shape
    r = int(np.ceil(6 * sigma))
    xi, yi = int(round(x0)), int(round(y0))
    x0b, x1b = max(xi - r, 0), min(xi + r + 1, w)
    y0b, y1b = max(yi - r, 0), min(yi + r + 1, h)
    if x0b >= x1b or y0b >= y1b:
        return
    yy, xx = np.mgrid[y0b:y1b, x0b:x1b]
    g = peak * np.exp(-((xx - x0) ** 2 + (yy - y0) ** 2) / (2.0 * sigma ** 2))
    img[y0b:y1b, x0b:x1b] += g.astype(img.dtype)


def star_field(seed, h, w, n_stars, sky=100.0, noise=1.0, sigma=(1.2, 3.0), peak=(5.0, 3000.0),
               nan_frac=0.0, mask_frac=0.0, flat_tops=0, shift=None):
    """(image, mask or None, stars) -- stars uniform over the frame, log-uniform peaks.

    ``flat_tops`` clips that many stars' cores to a plateau (equal-valued
    local maxima); ``shift`` moves every star by (dx, dy) pixels.
    """
    rng = np.random.default_rng(seed)
    img = (sky + rng.normal(0.0, noise, (h, w))).astype(np.float32)
    xs = rng.uniform(0, w - 1, n_stars)
    ys = rng.uniform(0, h - 1, n_stars)
    sg = rng.uniform(sigma[0], sigma[1], n_stars)
    pk = np.exp(rng.uniform(np.log(peak[0]), np.log(peak[1]), n_stars))
    dx, dy = (0.0, 0.0) if shift is None else shift
    unit = noise if noise > 0 else 1.0  # peaks in units of the noise (of 1 without noise)
    for i in range(n_stars):
        add_star_local(img, xs[i] + dx, ys[i] + dy, sg[i], pk[i] * unit)
    for i in range(min(flat_tops, n_stars)):
        xi, yi = int(round(xs[i] + dx)), int(round(ys[i] + dy))
        y0b, y1b, x0b, x1b = max(yi - 1, 0), min(yi + 2, h), max(xi - 1, 0), min(xi + 2, w)
        img[y0b:y1b, x0b:x1b] = np.float32(sky + 0.9 * pk[i] * unit)
    if nan_frac > 0:
        img[rng.random((h, w)) < nan_frac] = np.nan
    mask = None
    if mask_frac > 0:
        mask = (rng.random((h, w)) < mask_frac).astype(np.uint8)
    return img, mask, np.stack([xs, ys, sg, pk], axis=1)


def threshold_straddle_case():
    """(image, detect kwargs, detected with the unfused threshold): a 64x64
    frame whose one peak lies between median + round(ts * sigma) and
    fma(ts, sigma, median), so the two roundings of the threshold disagree on
    it.  The peak's exact value leaves the median and MAD unchanged."""
    from fractions import Fraction as F

    import oracle

    rng = np.random.default_rng(9)
    img = (100.0 + rng.normal(0.0, 2.0, (64, 64))).astype(np.float32)
    img[30:33, 30:33] = np.float32(160.0)
    img[31, 31] = np.float32(171.3)
    m, s = oracle.estimate_background(img)
    for i in range(200):
        peak = float(np.float32(171.3) + np.float32(1.5e-5) * i)  # compared in double
        ts = (peak - m) / s
        for _ in range(4):
            ts = np.nextafter(ts, -np.inf)
        for _ in range(9):
            fused = float(F(ts) * F(s) + F(m))
            plain = float(F(m) + F(ts * s))
            if (peak > fused) != (peak > plain):
                img[31, 31] = np.float32(peak)
                assert oracle.estimate_background(img) == (m, s)
                return img, {"stamp_radius": 4, "threshold_sigma": float(ts)}, peak > plain
            ts = np.nextafter(ts, np.inf)
    raise AssertionError("no straddling threshold found")
