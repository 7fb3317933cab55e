"""Pins the oracle's coarse-field ramp and its spatially varying variance.

  * ramp (src/subtract.cpp:119-205,287-313): from the oracle's FP64 lattice
    nodes, a numpy float32 restatement -- the stride rule, the node
    coordinates 0, S, 2S, ..., len-1, the cell and weight tables, and the
    three fused multiply-adds of base + slope * wx -- reproduces every field
    value bit for bit, while the unfused variant misses on a large fraction of
    pixels.  Fused multiply-adds are emulated exactly (two-sum error and a
    midpoint correction), so the check does not depend on numpy having one;
  * variance (src/subtract.cpp:641-824): an independent per-tile numpy
    construction -- K frozen at each 64^2 tile's centre from the point
    fields, squared and mirrored, Q = sum K^2, the smallest centred box
    holding (1 - 1e-4) Q, the zero-padded window, the 1e-3 flat shortcut
    Q * Var(R) and otherwise the dense correlation, plus Var(S) -- matches
    the oracle to 3e-6 on a frame whose fields vary strongly, with tiles of
    both kinds.  A centre one pixel off, the full footprint without the box
    shrink, and the dense sum in flat tiles all miss by more than 3x that.
"""
import numpy as np

import oracle

F32 = np.float32


def fmaf(a, b, c):
    """Correctly rounded float32 fused multiply-add of float32 arrays."""
    a64, b64, c64 = (np.asarray(v, F32).astype(np.float64) for v in (a, b, c))
    p = a64 * b64                      # exact: 24 + 24 significant bits
    s = p + c64
    bb = s - p
    err = (p - (s - bb)) + (c64 - bb)  # s + err is the exact value
    r = s.astype(F32)
    d = s - r.astype(np.float64)
    away = np.nextafter(r, np.where(d > 0, np.inf, -np.inf).astype(F32))
    mid = (d != 0) & (np.abs(d) * 2 == np.abs(away.astype(np.float64) - r.astype(np.float64)))
    fix = (err != 0) & mid             # the double sum sits on a float32 tie
    if fix.any():
        s2 = np.nextafter(s, np.where(err > 0, np.inf, -np.inf))
        r = np.where(fix, s2.astype(F32), r)
    return r


def test_fmaf_emulation_on_known_ties():
    # a*b = 1 + 2^-11 + 2^-24 exactly: a float32 tie.  Adding 2^-60 puts the
    # exact value just above it (fma rounds up), but the double sum rounds back
    # onto the tie, which float32 rounding then sends to even (down)
    a, b = F32(1 + 2 ** -12), F32(1 + 2 ** -12)
    c = F32(2.0 ** -60)
    up = F32(1 + 2 ** -11 + 2 ** -23)
    assert fmaf(a, b, c) == up
    assert F32(a.astype(np.float64) * b + c) != up    # the naive double route misses
    assert F32(a * b) + c == F32(1 + 2 ** -11)        # unfused: the tie goes to even
    assert fmaf(a, b, F32(0)) == F32(1 + 2 ** -11)    # an exact tie stays at even
    rng = np.random.default_rng(0)
    x = rng.standard_normal(100000).astype(F32)
    y = rng.standard_normal(100000).astype(F32)
    z = rng.standard_normal(100000).astype(F32)
    exact = [float(np.float64(p) * q + r) for p, q, r in zip(x[:2000], y[:2000], z[:2000])]
    assert np.all(np.abs(fmaf(x[:2000], y[:2000], z[:2000]) - np.array(exact)) <=
                  np.abs(np.spacing(np.array(exact, F32))) / 2)


def _axis(n, stride, coords):
    cell = np.minimum(np.arange(n) // stride, coords.size - 2)
    x0, x1 = coords[cell], coords[cell + 1]
    wt = np.where(x1 > x0, (np.arange(n) - x0).astype(F32) / (x1 - x0).astype(F32), F32(0))
    return cell, wt.astype(F32)


def _ramp(lat, w, h, fused=True):
    stride, nodes = lat["stride"], lat["nodes"].astype(F32)
    cx, wx = _axis(w, stride, lat["gx"])
    cy, fy = _axis(h, stride, lat["gy"])
    a00 = nodes[:, cy][:, :, cx]
    a01 = nodes[:, cy][:, :, cx + 1]
    a10 = nodes[:, cy + 1][:, :, cx]
    a11 = nodes[:, cy + 1][:, :, cx + 1]
    fyb, wxb = fy[None, :, None], wx[None, None, :]
    top = a01 - a00
    if fused:
        base = fmaf(a10 - a00, fyb, a00)
        slope = fmaf((a11 - a10) - top, fyb, top)
        return fmaf(slope, wxb, base)
    base = (a10 - a00) * fyb + a00
    slope = ((a11 - a10) - top) * fyb + top
    return slope * wxb + base


def test_coarse_ramp_bit_exact_vs_float32_restatement():
    h, w, n_max = 400, 430, 3
    nc = (n_max + 1) * (n_max + 2) // 2
    rng = np.random.default_rng(17)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    theta = rng.standard_normal((nc + 1) * knots.shape[0])
    lat = oracle.field_lattice(w, h, knots, theta, nc)
    assert not lat["exact"]
    # choose_field_stride: clamp(int(min knot spacing / 16), 1, 64)
    spacing = oracle.tps_min_spacing(knots)
    assert lat["stride"] == min(max(int(spacing / 16), 1), 64)
    s = lat["stride"]
    assert list(lat["gx"]) == list(range(0, w - 1, s)) + [w - 1]
    assert list(lat["gy"]) == list(range(0, h - 1, s)) + [h - 1]
    # the nodes are the FP64 fields at the node coordinates
    pts = np.array([(x, y) for y in lat["gy"] for x in lat["gx"]], np.float64)
    want_nodes = (oracle.tps_design(knots, pts) @ theta.reshape(nc + 1, -1).T).T
    np.testing.assert_allclose(lat["nodes"].reshape(nc + 1, -1), want_nodes, rtol=1e-12,
                               atol=1e-12)
    got, exact = oracle.fields(w, h, knots, theta, nc, bits=32)
    assert not exact
    want = _ramp(lat, w, h)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    unfused = _ramp(lat, w, h, fused=False)
    assert (unfused != got).mean() > 0.05


# ------------------------------------------------------------------ variance
BETA, N_MAX, RK = 1.4, 3, 12


def _footprints():
    g = oracle.gh_basis1d(BETA, N_MAX, RK)
    return np.stack([np.outer(g[ny], g[nx]).astype(F32) for nx, ny in oracle.gh_orders(N_MAX)])


def variance_numpy(vr, vs, fields32, phi, rk, dx=0, shrink=True, flat=True):
    h, w = vr.shape
    ks = 2 * rk + 1
    out = np.zeros((h, w))
    vpad = np.zeros((h + 2 * rk, w + 2 * rk), F32)
    vpad[rk:rk + h, rk:rk + w] = vr
    kinds = []
    for by in range(0, h, 64):
        for bx in range(0, w, 64):
            bw, bh = min(64, w - bx), min(64, h - by)
            cpx, cpy = min(bx + bw // 2 + dx, w - 1), min(by + bh // 2, h - 1)
            kern = np.zeros((ks, ks), F32)
            for n in range(phi.shape[0]):
                kern = fmaf(fields32[n, cpy, cpx], phi[n], kern)
            kf = (kern * kern)[::-1, ::-1]
            q = float((kern.astype(np.float64) ** 2).sum())
            hw = rk
            if shrink:
                target = (1.0 - float(F32(1e-4))) * q
                for cand in range(rk):
                    box = kf[rk - cand:rk + cand + 1, rk - cand:rk + cand + 1].astype(np.float64)
                    if box.sum() >= target:
                        hw = cand
                        break
            win = vpad[by:by + bh + 2 * rk, bx:bx + bw + 2 * rk]
            wmin, wmax = win.min(), win.max()
            if flat and wmax > 0 and (wmax - wmin) <= F32(1e-3) * wmax:
                kinds.append("flat")
                out[by:by + bh, bx:bx + bw] = F32(q) * win[rk:rk + bh, rk:rk + bw].astype(np.float64)
                continue
            kinds.append("dense")
            acc = np.zeros((bh, bw))
            for ly in range(rk - hw, rk + hw + 1):
                for lx in range(rk - hw, rk + hw + 1):
                    acc += float(kf[ly, lx]) * win[ly:ly + bh, lx:lx + bw]
            out[by:by + bh, bx:bx + bw] = acc
    return out + vs.astype(np.float64), kinds


def test_varying_kernel_variance_matches_per_tile_construction():
    h, w = 200, 230
    nc = (N_MAX + 1) * (N_MAX + 2) // 2
    rng = np.random.default_rng(8)
    knots = oracle.grid_knots(0.0, 0.0, w - 1.0, h - 1.0, 3, 3)
    k = knots.shape[0]
    theta = 0.4 * rng.standard_normal((nc + 1) * k)
    theta[k - 3] += 1.0
    vr = rng.uniform(20.0, 80.0, (h, w)).astype(F32)
    # a flat patch (relative spread 4e-4) big enough for whole tiles with halos
    vr[40:200, 40:230] = (50.0 * (1 + 4e-4 * rng.random((160, 190)))).astype(F32)
    vs = rng.uniform(1.0, 3.0, (h, w)).astype(F32)
    sci = np.zeros((h, w), F32)
    ref = np.zeros((h, w), F32)
    got = oracle.subtract(sci, ref, knots, theta, BETA, N_MAX, radius=RK, science_var=vs,
                          reference_var=vr, bits=32)["variance"]
    f32, exact = oracle.fields(w, h, knots, theta, nc, bits=32)
    phi = _footprints()
    want, kinds = variance_numpy(vr, vs, f32, phi, RK)
    assert "flat" in kinds and "dense" in kinds
    rel = np.abs(got / want - 1)
    assert rel.max() < 3e-6, rel.max()      # measured 8.2e-7 (FP32 vs FP64 sums)
    for variant in (dict(dx=1), dict(shrink=False), dict(flat=False)):
        alt, _ = variance_numpy(vr, vs, f32, phi, RK, **variant)
        assert np.abs(alt / want - 1).max() > 1e-5, variant
