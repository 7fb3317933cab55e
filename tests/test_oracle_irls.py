"""Pins the oracle's IRLS reweight and stamp clip (src/fit.cpp:25-59,411-483)
with independent numpy constructions, each chosen so a plausible slip fails:

  * reweight: 1 / (var_t + sum K^2 var_c) with K = sum_n a_n phi_n frozen at the
    stamp centre.  The fields vary strongly across the frame, so evaluating K
    at the pixel instead of the centre moves the weights by far more than the
    tolerance; masked reference pixels carry a huge variance, so including them
    fails; zero, negative and non-finite var_c are skipped; pixels near the
    frame edge see a zero-padded footprint; rows of inactive stamps keep their
    weights;
  * clip: thr = med + clip_sigma * 1.4826 * MAD with medians taken as element
    n/2 of the sorted values (the upper median for an even count), rejections
    worst first while more than the floor remain; inactive stamps take no part.
    Hand-built chi^2 vectors put a stamp between the thresholds that the upper
    median and the mean of the two middle values would give.
"""
import numpy as np

import oracle

BETA, N_MAX, R = 1.2, 2, 5


def _frame(rng, h=90, w=110):
    var = rng.uniform(2.0, 30.0, (h, w)).astype(np.float32)
    var[rng.random((h, w)) < 0.03] = 0.0
    var[rng.random((h, w)) < 0.02] = -3.0
    var[5, 7] = np.nan
    var[60, 20] = np.inf
    mask = np.zeros((h, w), np.uint8)
    bad = rng.random((h, w)) < 0.03
    mask[bad] = 1
    var[bad] = 1e6            # counted only if the mask were ignored
    return var, mask


def _theta(knots, nc, rng):
    k = knots.shape[0]
    th = 0.3 * rng.standard_normal((nc + 1) * k)
    th[k - 3] += 1.0          # the Gaussian term's intercept
    return th


def _footprints():
    g = oracle.gh_basis1d(BETA, N_MAX, R)
    return np.stack([np.outer(g[ny], g[nx]).astype(np.float32)
                     for nx, ny in oracle.gh_orders(N_MAX)])


def _numpy_weights(var, mask, knots, theta, cx, cy, active, px, py, stamp, vt, w0,
                   centre=True):
    nc = (N_MAX + 1) * (N_MAX + 2) // 2
    phi = _footprints()
    h, w = var.shape
    coef = np.asarray(theta).reshape(nc + 1, -1)
    good = (mask == 0) & np.isfinite(var) & (var > 0)
    vpad = np.zeros((h + 2 * R, w + 2 * R))
    vpad[R:R + h, R:R + w] = np.where(good, var.astype(np.float64), 0.0)
    out = np.array(w0, dtype=np.float64)
    for j in range(len(px)):
        s = stamp[j]
        if not active[s]:
            continue
        at = (cx[s], cy[s]) if centre else (px[j], py[j])
        a = oracle.tps_design(knots, np.array([at], np.float64))[0] @ coef.T
        kern = np.zeros(phi.shape[1:], np.float32)
        for n in range(nc):
            kern = kern + np.float32(a[n]) * phi[n]
        k2 = (kern * kern).astype(np.float64)
        # sum_{du,dv} K^2[dv+R, du+R] var(y-dv, x-du): the window flipped
        win = vpad[py[j]:py[j] + 2 * R + 1, px[j]:px[j] + 2 * R + 1][::-1, ::-1]
        tot = vt[j] + float((k2 * win).sum())
        if tot > 0:
            out[j] = 1.0 / tot
    return out


def test_irls_weights_match_numpy_and_discriminate():
    rng = np.random.default_rng(21)
    var, mask = _frame(rng)
    h, w = var.shape
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    nc = (N_MAX + 1) * (N_MAX + 2) // 2
    theta = _theta(knots, nc, rng)
    cx = np.array([3, 55, 100, 30, 106], np.int32)
    cy = np.array([4, 45, 85, 70, 2], np.int32)
    active = np.array([1, 1, 0, 1, 1], np.uint8)
    px, py, st = [], [], []
    for s in range(len(cx)):
        for dy in range(-4, 5):
            for dx in range(-4, 5):
                x, y = cx[s] + dx, cy[s] + dy
                if 0 <= x < w and 0 <= y < h:
                    px.append(x), py.append(y), st.append(s)
    px, py, st = (np.array(v, np.int32) for v in (px, py, st))
    vt = rng.uniform(1.0, 5.0, px.size)
    vt[::17] = -1e3                                # a non-positive total keeps its weight
    w0 = rng.uniform(0.01, 0.02, px.size)
    got = oracle.irls_weights(var, knots, theta, BETA, N_MAX, R, cx, cy, active, px, py, st, vt,
                              w0, reference_mask=mask)
    want = _numpy_weights(var, mask, knots, theta, cx, cy, active, px, py, st, vt, w0)
    np.testing.assert_allclose(got, want, rtol=2e-6)
    np.testing.assert_array_equal(got[st == 2], w0[st == 2])   # inactive stamp untouched
    assert np.array_equal(got[vt < 0], w0[vt < 0])
    # the discriminating variants miss by far more than the tolerance
    at_pixel = _numpy_weights(var, mask, knots, theta, cx, cy, active, px, py, st, vt, w0,
                              centre=False)
    assert np.abs(at_pixel / want - 1).max() > 1e-3
    no_mask = oracle.irls_weights(var, knots, theta, BETA, N_MAX, R, cx, cy, active, px, py, st,
                                  vt, w0)
    assert np.abs(no_mask / want - 1).max() > 0.5


def test_clip_hand_built_vectors():
    chi2 = np.array([1.00, 1.02, 0.98, 1.01, 0.99, 1.03, 0.97, 1.00, 1.5, 3.0, 2.0, 100.0])
    active = np.ones(12, np.uint8)
    active[11] = 0                        # inactive: no part in the statistics or the clip
    # 11 active: median = sorted[5] = 1.01; |v - med| sorted[5] = 0.02;
    # thr = 1.01 + 4 * 1.4826 * 0.02 = 1.1286 -> 3.0 (9), 2.0 (10), 1.5 (8), worst first
    assert oracle.clip_stamps(chi2, active, 5) == [9, 10, 8]
    assert oracle.clip_stamps(chi2, active, 9) == [9, 10]     # stops at the floor
    assert oracle.clip_stamps(chi2, active, 11) == []         # n_active <= floor: no clip
    assert oracle.clip_stamps(chi2, active, 5, clip_sigma=20.0) == [9, 10]   # thr = 1.6030


def test_clip_even_count_uses_the_upper_median():
    # 10 active values; sorted: 0.9 0.95 1.0 1.0 1.0 | 1.2 1.25 1.3 1.3 1.44
    chi2 = np.array([1.0, 0.95, 1.2, 1.0, 1.3, 0.9, 1.25, 1.0, 1.3, 1.44])
    active = np.ones(10, np.uint8)
    # upper median (element 5) = 1.2; |v - 1.2| sorted element 5 = 0.2;
    # thr = 1.2 + 1.4826 * 0.2 = 1.49652 with clip_sigma 1 -> nothing above
    assert oracle.clip_stamps(chi2, active, 2, clip_sigma=1.0) == []
    # with the mean of the two middle values (1.1) and its MAD (0.15 -> thr 1.3224)
    # 1.44 would go; the reference's element-n/2 rule keeps it.  clip_sigma 0.8:
    # thr = 1.2 + 0.8 * 1.4826 * 0.2 = 1.43722 < 1.44 -> only index 9
    assert oracle.clip_stamps(chi2, active, 2, clip_sigma=0.8) == [9]
