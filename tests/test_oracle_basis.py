"""Pins of the whole-image basis convolution, the basis tables, the TPS
penalty and the single-block decorrelation QA functions.

* oracle.basis_convolve (src/convolve.cpp:90-113) equals the reference's own
  compiled basis_convolve (the frozen fixture, and the live build where it
  exists) bit for bit, and a brute-force FP64 2-D convolution with the
  separable kernel to 1e-10;
* the library's host basis tables (gauss_hermite_basis1d / _kernels,
  src/bindings.cpp:171-206) equal the reference's samples bit for bit, the
  kernels are float(g_ny g_nx) and the orders run by total order then ny;
* tps_penalty equals the oracle's penalty and is symmetric with the affine
  functions in its null space (test_spatial.py:27-93);
* decorrelate_block / decorrelation_kernel (noise.cpp:98-161) in the oracle:
  a unit kernel gives Phi = 1 (the block unchanged, the kernel image a centred
  impulse), and the block apply equals the circular convolution of the block
  with the unshifted kernel image.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import noise as onoise

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _conv2d_same(img, kernel):
    kh, kw = kernel.shape
    ry, rx = kh // 2, kw // 2
    h, w = img.shape
    pad = np.zeros((h + 2 * ry, w + 2 * rx))
    pad[ry:ry + h, rx:rx + w] = img
    kf = kernel[::-1, ::-1]
    out = np.zeros((h, w))
    for i in range(kh):
        for j in range(kw):
            out += kf[i, j] * pad[i:i + h, j:j + w]
    return out


def test_basis_convolve_equals_reference_fixture():
    z = np.load(os.path.join(GOLDEN, "basis_convolve_ref.npz"))
    for i in range(int(z["n_cases"])):
        beta, n_max, radius = float(z[f"c{i}_beta"]), int(z[f"c{i}_n_max"]), int(z[f"c{i}_radius"])
        got = oracle.basis_convolve(z[f"c{i}_img"], beta, n_max, radius)
        np.testing.assert_array_equal(got, z[f"c{i}_bn"], err_msg=f"case {i}")


@pytest.mark.parametrize("shape,beta,n_max,radius", [((70, 90), 1.3, 4, 7), ((33, 20), 2.0, 2, 12)])
def test_basis_convolve_equals_live_reference(shape, beta, n_max, radius):
    if oracle.ref() is None:
        pytest.skip("reference build not present")
    rng = np.random.default_rng(3)
    img = (100 + 30 * rng.standard_normal(shape)).astype(np.float32)
    img[5, 7] = np.nan
    img[-2, 3] = np.inf
    got = oracle.basis_convolve(img, beta, n_max, radius)
    want = oracle.ref_basis_convolve(img, beta, n_max, radius)
    np.testing.assert_array_equal(got, want)


def test_basis_convolve_matches_direct_2d_convolution():
    rng = np.random.default_rng(9)
    img = rng.standard_normal((37, 51)).astype(np.float32)
    beta, n_max, radius = 1.1, 3, 5
    got = oracle.basis_convolve(img, beta, n_max, radius, bits=64)
    g = oracle.gh_basis1d(beta, n_max, radius)
    for c, (nx, ny) in enumerate(oracle.gh_orders(n_max)):
        want = _conv2d_same(img.astype(np.float64), np.outer(g[ny], g[nx]))
        np.testing.assert_allclose(got[c], want, rtol=1e-10, atol=1e-12)


def test_host_basis_tables_equal_reference_samples():
    from delta_amd import _core

    z = np.load(os.path.join(GOLDEN, "basis_convolve_ref.npz"))
    for i in range(int(z["n_cases"])):
        beta, n_max, radius = float(z[f"c{i}_beta"]), int(z[f"c{i}_n_max"]), int(z[f"c{i}_radius"])
        g = _core.gauss_hermite_basis1d(beta, n_max, radius)
        np.testing.assert_array_equal(g, z[f"c{i}_b1d"])
        orders, kern = _core.gauss_hermite_kernels(beta, n_max, radius)
        np.testing.assert_array_equal(orders, np.array(oracle.gh_orders(n_max), np.int32))
        for c, (nx, ny) in enumerate(orders):
            np.testing.assert_array_equal(kern[c], np.outer(g[ny], g[nx]).astype(np.float32))
    # radius 0 resolves to the default half-width
    assert _core.gauss_hermite_basis1d(1.7, 5).shape == (6, 2 * _core.default_radius(1.7, 5) + 1)


def test_host_tps_penalty():
    from delta_amd import _core

    knots = oracle.grid_knots(0.0, 0.0, 300.0, 200.0, 4, 3)
    pen = _core.tps_penalty(knots)
    np.testing.assert_allclose(pen, oracle.tps_setup(knots)[2], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pen, pen.T, atol=1e-12)
    # the trend block [1 u v] carries no bending energy
    np.testing.assert_array_equal(pen[-3:], 0.0)
    np.testing.assert_array_equal(pen[:, -3:], 0.0)
    assert np.linalg.eigvalsh(pen[:-3, :-3]).min() > 0


def test_oracle_decorrelation_unit_kernel_closed_form():
    rng = np.random.default_rng(1)
    n = 32
    img = rng.standard_normal((n, n)).astype(np.float32)
    unit = np.zeros((5, 5), np.float32)
    unit[2, 2] = 1.0
    np.testing.assert_allclose(onoise.decorrelate_block(img, unit, 3.0, 2.0), img, atol=1e-5)
    k = onoise.decorrelation_kernel(unit, 3.0, 2.0, n)
    want = np.zeros((n, n), np.float32)
    want[n // 2, n // 2] = 1.0
    np.testing.assert_allclose(k, want, atol=1e-6)


def test_oracle_block_apply_is_circular_convolution_with_kernel_image():
    rng = np.random.default_rng(2)
    n = 16
    img = rng.standard_normal((n, n)).astype(np.float32)
    kern = rng.random((5, 5)).astype(np.float32)
    kern /= kern.sum()
    got = onoise.decorrelate_block(img, kern, 1.0, 0.7).astype(np.float64)
    kimg = onoise.decorrelation_kernel(kern, 1.0, 0.7, n).astype(np.float64)
    k0 = np.roll(kimg, (n // 2, n // 2), axis=(0, 1))   # zero-phase (unshifted) kernel
    want = np.zeros((n, n))
    for dy in range(n):
        for dx in range(n):
            want += k0[dy, dx] * np.roll(img.astype(np.float64), (dy, dx), axis=(0, 1))
    np.testing.assert_allclose(got, want, atol=1e-5)


def test_host_basis_samples_equal_live_reference_over_random_parameters():
    """The FP64 samples behind every table (taps, footprints, component sums)
    equal the reference's compiled basis.cpp over 300 random (beta, n_max,
    radius); a sample differing in its last bit (a fused multiply-add placed
    on the other product) fails."""
    from delta_amd import _core

    lib = oracle.ref()
    if lib is None:
        pytest.skip("reference build not present")
    rng = np.random.default_rng(11)
    for beta, n_max, radius in zip(rng.uniform(0.3, 4.0, 300), rng.integers(0, 11, 300),
                                   rng.integers(1, 25, 300)):
        beta, n_max, radius = float(beta), int(n_max), int(radius)
        want = np.zeros((n_max + 1) * (2 * radius + 1))
        lib.ref_basis1d(beta, n_max, radius, want)
        np.testing.assert_array_equal(_core.gauss_hermite_basis1d(beta, n_max, radius).ravel(),
                                      want, err_msg=f"{beta} {n_max} {radius}")


def _boundary_betas():
    """beta values that put 3 beta sqrt(n_max+1) within an ulp of an integer,
    where the association of the product decides the ceiling."""
    out = []
    for n_max in range(1, 12):
        s = np.sqrt(n_max + 1.0)
        for k in range(3, 60):
            b = k / (3 * s)
            out += [(b, n_max), (np.nextafter(b, 0.0), n_max), (np.nextafter(b, 9.0), n_max)]
    return out


def test_default_radius_and_component_sums_equal_live_reference():
    """GaussHermiteBasis::default_radius (src/basis.cpp:9-12) on the betas that
    put 3 beta sqrt(n+1) on an integer -- (3 beta) sqrt(n+1) rounds to the
    other side for ~1 in 10 of them -- and component_sums (:81-93, the 1-D
    sums in index order) over random parameters: the library and the oracle
    equal the compiled reference bit for bit."""
    import ctypes as C

    from delta_amd import _core

    lib = oracle.ref()
    if lib is None:
        pytest.skip("reference build not present")
    lib.ref_default_radius.restype = C.c_int
    lib.ref_default_radius.argtypes = [C.c_double, C.c_int]
    lib.ref_component_sums.argtypes = [C.c_double, C.c_int, C.c_int,
                                       np.ctypeslib.ndpointer(np.float64, flags="C")]
    for beta, n_max in _boundary_betas():
        want = lib.ref_default_radius(float(beta), n_max)
        assert _core.default_radius(beta, n_max) == want, (beta, n_max)
        assert oracle.default_radius(float(beta), n_max) == want, (beta, n_max)
    rng = np.random.default_rng(12)
    for beta, n_max, radius in zip(rng.uniform(0.3, 4.0, 300), rng.integers(0, 11, 300),
                                   rng.integers(0, 25, 300)):
        beta, n_max, radius = float(beta), int(n_max), int(radius)
        r = radius if radius > 0 else lib.ref_default_radius(beta, n_max)
        want = np.zeros((n_max + 1) * (n_max + 2) // 2)
        lib.ref_component_sums(beta, n_max, r, want)
        np.testing.assert_array_equal(_core.component_sums(beta, n_max, radius), want,
                                      err_msg=f"{beta} {n_max} {radius}")


def test_default_radius_association_is_discriminated():
    """The boundary set above holds betas where the two associations of
    3 * beta * sqrt(n+1) give different ceilings, so the live comparison
    would catch the slip."""
    import math

    diff = sum(math.ceil(3.0 * b * math.sqrt(n + 1)) != math.ceil(3.0 * (b * math.sqrt(n + 1)))
               for b, n in _boundary_betas())
    assert diff > 50
