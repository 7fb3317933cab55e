"""Pins of the seed-mask + 8-connected labelling oracle (src/catalog.cpp:61-261).

The labels must be bit-exact: components numbered in raster order of their
first pixel.  Pinned against the reference's own compiled catalog.cpp
(oracle/_ref, built in this container only) and against
scipy.ndimage.label with a 3x3 structure, which numbers components the same
way.  Hand cases restate tests/test_catalog.py (spike, block, mask severing
8-connectivity, frame corners).
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

import oracle

_BAD = 1


def _random_case(rng, t):
    h, w = (int(v) for v in rng.integers(1, 90, 2))
    s = rng.standard_normal((h, w)).astype(np.float32)
    s[rng.random((h, w)) < 0.02] = np.nan
    s[rng.random((h, w)) < 0.01] = np.inf
    m = (rng.random((h, w)) < 0.1).astype(np.uint8) if t % 2 else None
    thr = float(rng.uniform(-0.5, 1.0))
    pol = 1 if t % 3 else -1
    return s, m, thr, pol


def test_seed_mask_predicate():
    s = np.array([[1.0, 0.99, np.nan, np.inf, -2.0, -np.inf]], np.float32)
    np.testing.assert_array_equal(oracle.seed_mask(s, 1.0, 1), [[1, 0, 0, 0, 0, 0]])
    np.testing.assert_array_equal(oracle.seed_mask(s, 1.0, -1), [[0, 0, 0, 0, 1, 0]])
    m = np.array([[1, 0, 0, 0, 0, 0]], np.uint8)
    np.testing.assert_array_equal(oracle.seed_mask(s, 1.0, 1, m), [[0, 0, 0, 0, 0, 0]])
    # the threshold is compared as float32 (catalog.cpp:71)
    thr = 1.0000000001
    np.testing.assert_array_equal(oracle.seed_mask(s, thr, 1), [[1, 0, 0, 0, 0, 0]])


def test_labels_match_scipy_raster_first_numbering():
    rng = np.random.default_rng(0)
    for t in range(60):
        s, m, thr, pol = _random_case(rng, t)
        got, n = oracle.label(s, thr, pol, m)
        want, nw = ndi.label(oracle.seed_mask(s, thr, pol, m), structure=np.ones((3, 3)))
        assert n == nw
        np.testing.assert_array_equal(got, want)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference sources not mounted")
def test_labels_match_reference_build():
    rng = np.random.default_rng(1)
    for t in range(40):
        s, m, thr, pol = _random_case(rng, t)
        got, n = oracle.label(s, thr, pol, m)
        want, nw = oracle.ref_label(s, thr, pol, m)
        assert n == nw
        np.testing.assert_array_equal(got, want)
    # large enough for the reference's multi-block phase 1 + seam stitch
    s = rng.standard_normal((700, 900)).astype(np.float32)
    for thr in (0.3, 1.5, -3.0):
        got, n = oracle.label(s, thr)
        want, nw = oracle.ref_label(s, thr)
        assert n == nw
        np.testing.assert_array_equal(got, want)


def test_hand_cases():
    z = np.zeros((60, 60), np.float32)
    s = z.copy()
    s[30, 30] = 8.0
    lab, n = oracle.label(s, 5.0)
    assert n == 1 and lab[30, 30] == 1 and lab.sum() == 1
    s = z.copy()
    s[20:35, 20:35] = 6.0
    lab, n = oracle.label(s, 5.0)
    assert n == 1 and (lab == 1).sum() == 225
    # a masked column severs 8-connectivity (test_catalog.py:142)
    s = np.zeros((40, 40), np.float32)
    s[18:23, 18:23] = 8.0
    m = np.zeros((40, 40), np.uint8)
    m[18:23, 20] = _BAD
    lab, n = oracle.label(s, 5.0, 1, m)
    assert n == 2 and (lab == 1).sum() == 10 and (lab == 2).sum() == 10
    # diagonal touch joins; corners of the frame
    s = np.zeros((5, 5), np.float32)
    s[0, 0] = s[1, 1] = s[4, 4] = s[0, 4] = 1.0
    lab, n = oracle.label(s, 0.5)
    assert n == 3
    np.testing.assert_array_equal(lab[[0, 1, 0, 4], [0, 1, 4, 4]], [1, 1, 2, 3])
    # U shape: the later arm merges into the earlier first pixel's label
    s = np.zeros((4, 5), np.float32)
    s[0, 0] = s[1, 0] = s[2, 0] = s[0, 4] = s[1, 4] = s[2, 4] = 1.0
    s[3, :] = 1.0
    lab, n = oracle.label(s, 0.5)
    assert n == 1 and (lab[s > 0] == 1).all()


def test_empty_and_full_frames():
    lab, n = oracle.label(np.zeros((7, 9), np.float32), 0.5)
    assert n == 0 and not lab.any()
    lab, n = oracle.label(np.ones((7, 9), np.float32), 0.5)
    assert n == 1 and (lab == 1).all()
    lab, n = oracle.label(np.ones((1, 1), np.float32), 0.5)
    assert n == 1
