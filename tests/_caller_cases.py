"""Deterministic inputs for the caller-arithmetic parity cases.

The score-PSF estimate and the variance rescale factor are the reference
caller's numpy expressions (python/delta/pipeline.py:66-86,285-302).  Their
expected outputs are committed under tests/golden/caller_numpy.npz, written in
the build container by tests/golden/make_golden.py; the GPU tests rebuild the
same inputs here and compare the device result with the stored one.  Each case
carries a digest of its inputs, so a drift in the generator fails loudly
instead of comparing against vectors of other inputs.
"""
import hashlib

import numpy as np

PSF_CASES = ((60, 8), (61, 8), (7, 3), (25, 15))
RESCALE_CASES = (((301, 257), False), ((300, 256), True), ((64, 33), True))


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        if a is None:
            h.update(b"none")
            continue
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode() + str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()[:16]


def psf_case(n_src, radius):
    """A 256^2 star field and the brightest n_src positions, with one source
    off the frame edge, a NaN inside a cutout and a blank-sky cutout whose
    background-subtracted sum is negative."""
    from delta_amd import synth

    p = synth.make_pair(synth.PairConfig(size=256, n_stars=300, seed=n_src))
    img = p["science"].copy()
    rng = np.random.default_rng(n_src)
    st = p["stars"]
    order = np.argsort(-st[:, 2])[:n_src]
    xs = np.round(st[order, 0]).astype(np.int32)
    ys = np.round(st[order, 1]).astype(np.int32)
    xs[0], ys[0] = 2, 100                       # leaves the frame for radius >= 3
    img[ys[1], xs[1]] = np.nan                  # NaN inside a cutout
    xs[2], ys[2] = rng.integers(40, 200, 2)     # a blank-sky cutout
    img[ys[2] - radius:ys[2] + radius + 1, xs[2] - radius:xs[2] + radius + 1] = -5.0
    img[ys[2], xs[2]] = -50.0
    return img, xs, ys


def rescale_case(shape, masked):
    """Difference/variance frames with zero and non-finite variances, a NaN
    difference, a zero row and (optionally) a 10% mask."""
    rng = np.random.default_rng(shape[0])
    d = (rng.standard_normal(shape) * 3).astype(np.float32)
    v = rng.uniform(0.5, 4.0, shape).astype(np.float32)
    v[::7, ::5] = 0.0
    d[3, 3] = np.nan
    v[4, 4] = np.inf
    d[5, :] = 0.0
    m = (rng.random(shape) < 0.1).astype(np.uint8) if masked else None
    return d, v, m


def rescale_count_case(shape, keep):
    """Only the first ``keep`` pixels have a positive variance (odd and even
    good counts, so the median averages two middle elements or not)."""
    d, _, _ = rescale_case(shape, False)
    vv = np.zeros(shape, np.float32)
    vv.flat[:keep] = 1.0 + np.arange(keep, dtype=np.float32)
    return d, vv
