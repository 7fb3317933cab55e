"""Write the committed golden fixtures.

Run in the build container, where /root/reference exists:
    make -C oracle ref && python tests/golden/make_golden.py
basis_convolve_ref.npz comes from oracle/_ref/libdeltaref.so -- the
reference's basis.cpp, convolve.cpp and catalog.cpp compiled where they lie.
caller_numpy.npz holds the outputs of the caller's numpy arithmetic
(python/delta/pipeline.py:66-86,285-302), restated below, on the inputs of
tests/_caller_cases.py.  The GPU tests read only the .npz files; this script
does not travel with them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def bn_fixture():
    rng = np.random.default_rng(20240611)
    cases = []
    for (h, w, beta, n_max, radius) in ((41, 53, 1.3, 4, 7), (50, 60, 1.7, 3, 6),
                                        (40, 52, 1.0, 6, 15), (33, 35, 2.2, 2, 10)):
        img = (1000 + 50 * rng.standard_normal((h, w))).astype(np.float32)
        img[rng.integers(0, h), :] += 4e4  # a bright streak
        bn = oracle.ref_basis_convolve(img, beta, n_max, radius)
        b1d = np.zeros((n_max + 1) * (2 * radius + 1))
        oracle.ref().ref_basis1d(beta, n_max, radius, b1d)
        cases.append(dict(img=img, beta=beta, n_max=n_max, radius=radius, bn=bn,
                          b1d=b1d.reshape(n_max + 1, -1)))
    out = {}
    for i, c in enumerate(cases):
        for k, v in c.items():
            out[f"c{i}_{k}"] = np.asarray(v)
    out["n_cases"] = np.asarray(len(cases))
    np.savez_compressed(os.path.join(HERE, "basis_convolve_ref.npz"), **out)


MEDIAN_CHI2_1 = 0.4549364231195724   # median of chi-square with one dof


def score_psf_numpy(image, xs, ys, radius):
    """The caller's score-PSF estimate (python/delta/pipeline.py:285-302) in
    numpy: every cutout of side 2*radius+1 that lies inside the frame, in
    float64, less its median, is divided by its sum when that sum is positive;
    the per-pixel median over those unit-sum cutouts, floored at zero and
    renormalised, is the PSF (float32).  The numpy calls are the ones whose
    rounding the reference fixes (median, the contiguous pairwise sum, the
    float64 division), so the result is bit-comparable."""
    img = np.asarray(image, dtype=np.float32)
    h, w = img.shape
    side = 2 * radius + 1
    unit = []
    for cx, cy in zip(np.asarray(xs).tolist(), np.asarray(ys).tolist()):
        x0, y0 = int(cx) - radius, int(cy) - radius
        if x0 < 0 or y0 < 0 or x0 + side > w or y0 + side > h:
            continue
        patch = np.array(img[y0:y0 + side, x0:x0 + side], dtype=np.float64)
        patch -= np.median(patch)
        norm = patch.sum()
        if norm > 0:
            unit.append(patch / norm)
    if not unit:
        raise ValueError("no usable sources for PSF estimation")
    med = np.median(np.asarray(unit), axis=0)
    med = np.maximum(med, 0.0)
    return (med / med.sum()).astype(np.float32)


def rescale_numpy(d, v, m=None):
    """The caller's variance rescale factor (python/delta/pipeline.py:66-86):
    median of d^2/v in float64 over finite pixels with v > 0 (and mask 0),
    over the median of chi-square with one dof; 1.0 when no pixel qualifies."""
    ok = np.isfinite(d) & np.isfinite(v) & (v > 0)
    if m is not None:
        ok &= m == 0
    if not ok.any():
        return 1.0
    dd = d[ok].astype(np.float64)
    return float(np.median(dd * dd / v[ok].astype(np.float64)) / MEDIAN_CHI2_1)


def caller_fixture():
    from tests import _caller_cases as cc

    out = {}
    for n_src, radius in cc.PSF_CASES:
        img, xs, ys = cc.psf_case(n_src, radius)
        key = f"psf_{n_src}_{radius}"
        out[key] = score_psf_numpy(img, xs, ys, radius)
        out[key + "_digest"] = np.asarray(cc.digest(img, xs, ys))
    for shape, masked in cc.RESCALE_CASES:
        d, v, m = cc.rescale_case(shape, masked)
        key = f"rescale_{shape[0]}_{shape[1]}_{int(masked)}"
        out[key] = np.asarray(rescale_numpy(d, v, m))
        out[key + "_digest"] = np.asarray(cc.digest(d, v, m))
        for keep in (11, 12):
            dk, vk = cc.rescale_count_case(shape, keep)
            kk = f"{key}_keep{keep}"
            out[kk] = np.asarray(rescale_numpy(dk, vk))
            out[kk + "_digest"] = np.asarray(cc.digest(dk, vk))
    np.savez_compressed(os.path.join(HERE, "caller_numpy.npz"), **out)


def main():
    caller_fixture()
    if oracle.ref() is None:
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    bn_fixture()
    if hasattr(oracle, "ccl_fixture"):
        oracle.ccl_fixture(HERE)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
