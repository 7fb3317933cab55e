"""Synthetic score images for the catalog tests (CPU and GPU)."""
import numpy as np

_BAD, _COSMIC = 1, 4


def blob(shape, x, y, peak, sigma):
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    return (peak * np.exp(-((xx - x) ** 2 + (yy - y) ** 2) / (2 * sigma ** 2))).astype(np.float32)


def stamp(size, sigma):
    r = size // 2
    yy, xx = np.mgrid[-r:r + 1, -r:r + 1]
    g = np.exp(-(xx ** 2 + yy ** 2) / (2.0 * sigma ** 2))
    return (g / g.sum()).astype(np.float32)


def score_field(seed, h, w, n_pos, n_neg, noise=1.0, plateaus=0, mask_frac=0.0):
    """(score, difference, variance, mask, blob positions) of a synthetic frame."""
    rng = np.random.default_rng(seed)
    s = rng.normal(0.0, noise, (h, w)).astype(np.float32)
    pos = []
    for i in range(n_pos + n_neg):
        x, y = rng.uniform(-3, w + 2), rng.uniform(-3, h + 2)
        amp = rng.uniform(6, 60) * (1 if i < n_pos else -1)
        sig = rng.uniform(0.8, 3.0)
        r = int(np.ceil(7 * sig))
        x0, x1 = max(int(x) - r, 0), min(int(x) + r + 2, w)
        y0, y1 = max(int(y) - r, 0), min(int(y) + r + 2, h)
        if x0 < x1 and y0 < y1:
            yy, xx = np.mgrid[y0:y1, x0:x1]
            s[y0:y1, x0:x1] += (amp * np.exp(-((xx - x) ** 2 + (yy - y) ** 2) /
                                             (2 * sig ** 2))).astype(np.float32)
        pos.append((x, y))
    for i in range(plateaus):
        x, y = int(rng.integers(2, w - 5)), int(rng.integers(2, h - 5))
        s[y:y + 3, x:x + 3] = np.float32(9.0)
    d = (s * np.float32(3.0) + rng.normal(0, 0.5, (h, w)).astype(np.float32)).astype(np.float32)
    d[rng.random((h, w)) < 0.002] = np.nan
    v = rng.uniform(0.5, 2.0, (h, w)).astype(np.float32)
    v[rng.random((h, w)) < 0.01] = 0.0
    m = None
    if mask_frac:
        m = np.where(rng.random((h, w)) < mask_frac,
                     rng.choice([_BAD, _COSMIC, 32], (h, w)), 0).astype(np.uint8)
    return s, d, v, m, np.array(pos)


CASES = [
    # (field kwargs, build_catalog kwargs)
    ({}, {}),
    ({"plateaus": 6}, {"threshold_sigma": 4.0, "expected_fwhm": 2.5}),
    ({"mask_frac": 0.02}, {}),
    ({"mask_frac": 0.02}, {"exclude_bad_pixels": False, "aperture_radius": 6}),
    ({"mask_frac": 0.01}, {"return_negative": True, "threshold_sigma_dipole": 4.0}),
    ({}, {"fit_psf_shape": True, "shape_chi2_tolerance": 2.0, "expected_fwhm": 4.0}),
    ({"mask_frac": 0.01}, {"fit_psf_shape": True, "bright": True, "bright_star_radius": 12.0}),
    ({}, {"bright": True, "return_negative": True}),
    ({}, {"bright": True, "fit_psf_shape": True, "max_shape_fit": 3}),
    ({"noise": 0.0}, {"threshold_sigma": 0.5, "threshold_sigma_dipole": 0.5}),
]


def case_args(case, seed=100):
    fk, ck = CASES[case]
    s, d, v, m, pos = score_field(seed + case, 150, 170, 25, 12, **fk)
    ck = dict(ck)
    if ck.pop("bright", False):
        ck["bright_x"], ck["bright_y"] = pos[::3, 0], pos[::3, 1]
    if ck.get("fit_psf_shape"):
        ck["psf"], ck["variance"] = stamp(9, 1.6), v
    return s, d, dict(mask=m, **ck)
