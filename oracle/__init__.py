"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle is a C restatement of delta's PSF-matching algorithm, compiled in
two precisions (``liboracle32.so``: float image arithmetic as delta's own CPU
path does it; ``liboracle64.so``: the same algorithm in double, the accuracy
reference).  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` import this module; the product package ``delta_amd``
never does.

``ref()`` loads ``oracle/_ref/libdeltaref.so`` -- the reference's own
Eigen-free translation units (basis.cpp, convolve.cpp, catalog.cpp) compiled
in this container by ``make -C oracle ref`` -- or returns None where that
library is absent (on a machine without the reference sources).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict = {}

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile the oracle (and the reference wrapper where sources exist)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


class _Opt:
    """ctypes converter that maps None to NULL for an optional ndarray."""

    def __init__(self, dtype):
        self.dtype = dtype

    def from_param(self, v):
        if v is None:
            return None
        a = np.ascontiguousarray(v, dtype=self.dtype)
        return a.ctypes.data_as(C.c_void_p)


def _lib(bits: int):
    if bits in _LIBS:
        return _LIBS[bits]
    path = os.path.join(_HERE, f"liboracle{bits}.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    real = _f32p if bits == 32 else _f64p
    lib.or_subtract.restype = C.c_int
    lib.or_subtract.argtypes = [C.c_int, C.c_int, _f32p, _f32p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p, _f64p, C.c_int, _f64p, C.c_double,
                                C.c_int, C.c_int, C.c_double, real, C.c_void_p, C.c_void_p]
    lib.or_gh_basis1d.argtypes = [C.c_double, C.c_int, C.c_int, _f64p]
    lib.or_gh_default_radius.argtypes = [C.c_double, C.c_int]
    lib.or_gh_default_radius.restype = C.c_int
    lib.or_tps_setup.argtypes = [_f64p, C.c_int, _f64p, _f64p, _f64p]
    lib.or_tps_design.argtypes = [_f64p, C.c_int, _f64p, C.c_int, _f64p]
    lib.or_tps_min_spacing.argtypes = [_f64p, C.c_int]
    lib.or_tps_min_spacing.restype = C.c_double
    lib.or_fields.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, _f64p, C.c_int, real]
    lib.or_fill_value.argtypes = [_f32p, C.c_void_p, C.c_long]
    lib.or_fill_value.restype = C.c_float
    _extra_signatures(lib, real)
    _LIBS[bits] = lib
    return lib


def _extra_signatures(lib, real):
    """Signatures of the detection units (seed mask, CCL, matched filter)."""
    lib.or_seed_mask.argtypes = [_f32p, C.c_void_p, C.c_long, C.c_double, C.c_int, _u8p]
    lib.or_label_seed.restype = C.c_int
    lib.or_label_seed.argtypes = [_u8p, C.c_int, C.c_int, _i32p]
    lib.or_label.restype = C.c_int
    lib.or_label.argtypes = [_f32p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, _i32p]
    lib.or_matched_filter.argtypes = [_f32p, _f32p, C.c_int, _f32p, C.c_int, C.c_int, _f32p]
    lib.or_gh_shifted.restype = C.c_int
    lib.or_gh_shifted.argtypes = [C.c_double, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                  C.c_void_p, C.c_void_p]
    lib.or_shifted_kernel.restype = C.c_int
    lib.or_shifted_kernel.argtypes = [_f64p, C.c_int, C.c_int, _f64p, C.c_void_p]
    lib.or_kernel_chi2.restype = C.c_int
    lib.or_kernel_chi2.argtypes = [C.c_int, C.c_int, _f32p, _f32p, C.c_void_p, _f64p, C.c_int,
                                   C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]


def set_threads(n):
    """OpenMP threads of both oracle builds' parallel regions; returns the count."""
    got = 0
    for bits in (32, 64):
        lib = _lib(bits)
        lib.or_set_threads.restype = C.c_int
        lib.or_set_threads.argtypes = [C.c_int]
        got = lib.or_set_threads(int(n))
    return got


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _cu8(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint8)


# --------------------------------------------------------------------------- basis / TPS
def gh_basis1d(beta, n_max, radius, bits=64):
    ks = 2 * radius + 1
    out = np.zeros((n_max + 1) * ks)
    _lib(bits).or_gh_basis1d(beta, n_max, radius, out)
    return out.reshape(n_max + 1, ks)


def default_radius(beta, n_max):
    return _lib(64).or_gh_default_radius(beta, n_max)


def gh_orders(n_max):
    return [(tot - ny, ny) for tot in range(n_max + 1) for ny in range(tot + 1)]


def grid_knots(x0, y0, x1, y1, nx, ny):
    """Regular knot grid (src/spatial.cpp:113-129)."""
    out = np.zeros((nx * ny, 2))
    i = 0
    for iy in range(ny):
        fy = 0.5 if ny == 1 else iy / (ny - 1)
        for ix in range(nx):
            fx = 0.5 if nx == 1 else ix / (nx - 1)
            out[i] = (x0 + fx * (x1 - x0), y0 + fy * (y1 - y0))
            i += 1
    return out


def tps_setup(knots):
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    k = knots.shape[0]
    norm = np.zeros(3)
    z = np.zeros(max(k * (k - 3), 1))
    pen = np.zeros(k * k)
    _lib(64).or_tps_setup(knots, k, norm, z, pen)
    return norm, z[: k * (k - 3)].reshape(k, k - 3) if k > 3 else np.zeros((k, 0)), pen.reshape(k, k)


def tps_design(knots, points):
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    pts = np.ascontiguousarray(points, dtype=np.float64)
    out = np.zeros((pts.shape[0], knots.shape[0]))
    _lib(64).or_tps_design(knots, knots.shape[0], pts, pts.shape[0], out)
    return out


def tps_min_spacing(knots):
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    return _lib(64).or_tps_min_spacing(knots, knots.shape[0])


def fields(w, h, knots, theta, nc, bits=32):
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    out = np.zeros((nc + 1, h, w), dtype=np.float32 if bits == 32 else np.float64)
    ex = _lib(bits).or_fields(w, h, knots, knots.shape[0],
                              np.ascontiguousarray(theta, dtype=np.float64), nc, out)
    return out, bool(ex)


def field_lattice(w, h, knots, theta, nc):
    """The coarse lattice of the fields (src/subtract.cpp:113-316):
    {exact, stride, nodes (nf, ncy, ncx) FP64, gx, gy}."""
    lib = _lib(64)
    if not getattr(lib, "_fl_ready", False):
        lib.or_field_lattice.restype = C.c_int
        lib.or_field_lattice.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, _f64p, C.c_int, _i32p,
                                         C.c_void_p, C.c_long, C.c_void_p, C.c_int, C.c_void_p,
                                         C.c_int]
        lib._fl_ready = True
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    info = np.zeros(4, np.int32)
    lib.or_field_lattice(w, h, knots, knots.shape[0], theta, nc, info, None, 0, None, 0, None, 0)
    exact, stride, ncx, ncy = (int(v) for v in info)
    out = {"exact": bool(exact), "stride": stride}
    if exact:
        return out
    lat = np.zeros((nc + 1, ncy, ncx))
    gx, gy = np.zeros(ncx, np.int32), np.zeros(ncy, np.int32)
    rc = lib.or_field_lattice(w, h, knots, knots.shape[0], theta, nc, info,
                              lat.ctypes.data_as(C.c_void_p), lat.size,
                              gx.ctypes.data_as(C.c_void_p), ncx, gy.ctypes.data_as(C.c_void_p),
                              ncy)
    assert rc == 0
    out.update(nodes=lat, gx=gx, gy=gy)
    return out


def photometric_scale(knots, theta, component_sums, h, w, bits=32):
    """sum_c S_c a_c(x, y) from the coarse-lattice fields (src/fit.cpp:513-530)."""
    lib = _lib(bits)
    if not getattr(lib, "_ps_ready", False):
        lib.or_photometric_scale.restype = C.c_int
        lib.or_photometric_scale.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, _f64p, _f64p,
                                             C.c_int, C.c_void_p]
        lib._ps_ready = True
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    sums = np.ascontiguousarray(component_sums, dtype=np.float64)
    out = np.zeros((h, w), dtype=np.float32 if bits == 32 else np.float64)
    lib.or_photometric_scale(w, h, knots, knots.shape[0],
                             np.ascontiguousarray(theta, dtype=np.float64), sums, sums.size,
                             _ptr(out))
    return out


def fill_value(ref, mask=None):
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    m = _cu8(mask)
    return float(_lib(32).or_fill_value(ref, _ptr(m), ref.size))


# --------------------------------------------------------------------------- subtract
def subtract(science, reference, knots, theta, beta, n_max, radius=0, saturation=0.0,
             science_var=None, reference_var=None, science_mask=None,
             reference_mask=None, bits=32):
    """Oracle of ``_core.subtract`` (src/bindings.cpp:603-638)."""
    sci = np.ascontiguousarray(science, dtype=np.float32)
    ref = np.ascontiguousarray(reference, dtype=np.float32)
    h, w = sci.shape
    knots = np.ascontiguousarray(knots, dtype=np.float64)
    radius = radius if radius > 0 else default_radius(beta, n_max)
    dt = np.float32 if bits == 32 else np.float64
    sv, rv = _c32(science_var), _c32(reference_var)
    sm, rm = _cu8(science_mask), _cu8(reference_mask)
    diff = np.zeros((h, w), dtype=dt)
    var = np.zeros((h, w), dtype=dt) if (sv is not None or rv is not None) else None
    mask = (np.zeros((h, w), dtype=np.uint8)
            if (sm is not None or rm is not None or saturation > 0) else None)
    rc = _lib(bits).or_subtract(w, h, sci, ref, _ptr(sv), _ptr(rv), _ptr(sm), _ptr(rm), knots,
                                knots.shape[0], np.ascontiguousarray(theta, dtype=np.float64),
                                beta, n_max, radius, saturation, diff, _ptr(var), _ptr(mask))
    if rc != 0:
        raise ValueError(f"oracle subtract failed ({rc})")
    return {"difference": diff, "variance": var, "mask": mask}


def basis_convolve(img, beta, n_max, radius=0, bits=32):
    """Oracle of ``_core.basis_convolve`` (src/convolve.cpp:90-113): (nc, H, W)."""
    lib = _lib(bits)
    lib.or_basis_convolve.restype = C.c_int
    lib.or_basis_convolve.argtypes = [_f32p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                      C.c_void_p]
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape
    radius = radius if radius > 0 else default_radius(beta, n_max)
    nc = (n_max + 1) * (n_max + 2) // 2
    out = np.zeros((nc, h, w), dtype=np.float32 if bits == 32 else np.float64)
    rc = lib.or_basis_convolve(img, w, h, beta, n_max, radius, out.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise ValueError(f"oracle basis_convolve failed ({rc})")
    return out


# --------------------------------------------------------------------------- reference build
_REF = None
_REF_PATH = os.path.join(_HERE, "_ref", "libdeltaref.so")


def ref_available():
    """Whether the reference build exists, without loading it (test
    collection calls this; only a test that runs loads the library)."""
    return os.path.exists(_REF_PATH)


def ref():
    """The reference's own compiled basis/convolve/CCL, or None when absent."""
    global _REF
    if _REF is not None:
        return _REF or None
    if not ref_available():
        _REF = False
        return None
    lib = C.CDLL(_REF_PATH)
    lib.ref_basis1d.argtypes = [C.c_double, C.c_int, C.c_int, _f64p]
    lib.ref_basis_convolve.argtypes = [_f32p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _f32p]
    lib.ref_label.restype = C.c_int
    lib.ref_label.argtypes = [_f32p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, _i32p]
    _REF = lib
    return lib


def ref_basis_convolve(img, beta, n_max, radius):
    lib = ref()
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape
    nc = (n_max + 1) * (n_max + 2) // 2
    out = np.zeros((nc, h, w), dtype=np.float32)
    lib.ref_basis_convolve(img, w, h, beta, n_max, radius, out)
    return out


# --------------------------------------------------------------------------- fit
def _fit_sigs(lib):
    if getattr(lib, "_fit_ready", False):
        return
    vp = C.c_void_p
    lib.or_solve_gls_gcv.restype = C.c_int
    lib.or_solve_gls_gcv.argtypes = [_f64p, C.c_int, _f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int,
                                     _f64p, C.c_int, _f64p, _f64p, _f64p]
    lib.or_assemble.restype = C.c_int
    lib.or_assemble.argtypes = [_f64p, C.c_int, _f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int,
                                _f64p, _f64p, _f64p]
    lib.or_gather.restype = C.c_int
    lib.or_gather.argtypes = [C.c_int, C.c_int, _f32p, _f32p, vp, vp, vp, vp, _i32p, _i32p,
                              C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _i32p, _i32p,
                              _i32p, _f64p, _f64p]
    lib.or_fit_kernel.restype = C.c_int
    lib.or_fit_kernel.argtypes = [C.c_int, C.c_int, _f32p, _f32p, vp, vp, vp, vp, _f64p, C.c_int,
                                  _i32p, _i32p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                  _f64p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                  _f64p, _f64p, _f64p, _i32p, _i32p, _i32p, _f64p, _u8p]
    lib.or_solve_gls_cv.restype = C.c_int
    lib.or_solve_gls_cv.argtypes = [_f64p, C.c_int, _f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int,
                                    _f64p, C.c_int, _i32p, C.c_int, C.c_int, _f64p, _f64p, _f64p]
    lib.or_irls_weights.restype = C.c_int
    lib.or_irls_weights.argtypes = [C.c_int, C.c_int, _f32p, vp, _f64p, C.c_int, _f64p,
                                    C.c_double, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _u8p,
                                    C.c_int, _i32p, _i32p, _i32p, _f64p, _f64p]
    lib.or_clip_stamps.restype = C.c_int
    lib.or_clip_stamps.argtypes = [_f64p, _u8p, C.c_int, C.c_int, C.c_double, _i32p]
    lib._fit_ready = True


def irls_weights(reference_var, knots, theta, beta, n_max, radius, stamp_cx, stamp_cy, active,
                 px, py, stamp, var_t, weights, reference_mask=None, bits=64):
    """One IRLS reweight (src/fit.cpp:38-59,411-446) on caller rows: K frozen at
    each active stamp's centre; returns the updated weights."""
    lib = _lib(bits)
    _fit_sigs(lib)
    rv = _c32(reference_var)
    h, w = rv.shape
    rm = _cu8(reference_mask)
    knots = np.ascontiguousarray(knots, np.float64)
    out = np.array(weights, dtype=np.float64)
    i32 = lambda a: np.ascontiguousarray(a, np.int32).ravel()  # noqa: E731
    rc = lib.or_irls_weights(w, h, rv, _ptr(rm), knots, knots.shape[0],
                             np.ascontiguousarray(theta, np.float64), beta, n_max, radius,
                             len(stamp_cx), i32(stamp_cx), i32(stamp_cy),
                             np.ascontiguousarray(active, np.uint8), len(px), i32(px), i32(py),
                             i32(stamp), np.ascontiguousarray(var_t, np.float64), out)
    if rc != 0:
        raise ValueError("oracle irls_weights: bad arguments")
    return out


def clip_stamps(chi2, active, floor_stamps, clip_sigma=4.0):
    """One pass of the robust stamp clip (src/fit.cpp:456-483): indices of the
    rejected stamps, worst first."""
    lib = _lib(64)
    _fit_sigs(lib)
    chi2 = np.ascontiguousarray(chi2, np.float64)
    rej = np.zeros(max(chi2.size, 1), np.int32)
    n = lib.or_clip_stamps(chi2, np.ascontiguousarray(active, np.uint8), chi2.size,
                           int(floor_stamps), float(clip_sigma), rej)
    return rej[:n].tolist()


def solve_gls_gcv(knots, points, target, weights, bn, lambda_grid, bits=64):
    """Oracle of ``_core.solve_gls_gcv`` (src/bindings.cpp:540-551)."""
    lib = _lib(bits)
    _fit_sigs(lib)
    knots = np.ascontiguousarray(knots, np.float64)
    pts = np.ascontiguousarray(points, np.float64)
    bn = np.ascontiguousarray(bn, np.float64)
    n, nc = bn.shape
    grid = np.ascontiguousarray(lambda_grid, np.float64)
    p = (nc + 1) * knots.shape[0]
    theta, scal, curve = np.zeros(p), np.zeros(4), np.zeros(grid.size)
    rc = lib.or_solve_gls_gcv(knots, knots.shape[0], pts, np.ascontiguousarray(target, np.float64),
                              np.ascontiguousarray(weights, np.float64), bn, n, nc, grid, grid.size,
                              theta, scal, curve)
    if rc != 0:
        raise RuntimeError(f"solve_gls_gcv failed ({rc})")
    return {"theta": theta, "n_components": nc, "n_spatial": knots.shape[0], "lambda": scal[0],
            "gcv": scal[1], "effective_dof": scal[2], "rss": scal[3], "lambda_grid": grid,
            "gcv_curve": curve}


def solve_gls_cv(knots, points, target, weights, bn, lambda_grid, group, n_groups, warm_start=-1,
                 bits=64):
    """Oracle of ``_core.solve_gls_cv`` (src/bindings.cpp:553-566): leave-group-out
    CV over the grid (coarse-to-fine, or a descent from ``warm_start``)."""
    lib = _lib(bits)
    _fit_sigs(lib)
    knots = np.ascontiguousarray(knots, np.float64)
    pts = np.ascontiguousarray(points, np.float64)
    bn = np.ascontiguousarray(bn, np.float64)
    n, nc = bn.shape
    grid = np.ascontiguousarray(lambda_grid, np.float64)
    grp = np.ascontiguousarray(group, np.int32)
    if grp.size != n:
        raise RuntimeError("solve_gls_cv: group.size() must match points.rows()")
    p = (nc + 1) * knots.shape[0]
    theta, scal, curve = np.zeros(p), np.zeros(4), np.zeros(grid.size)
    rc = lib.or_solve_gls_cv(knots, knots.shape[0], pts, np.ascontiguousarray(target, np.float64),
                             np.ascontiguousarray(weights, np.float64), bn, n, nc, grid, grid.size,
                             grp, int(n_groups), int(warm_start), theta, scal, curve)
    if rc == -3:
        raise RuntimeError("solve_gls_cv: group id out of range")
    if rc != 0:
        raise RuntimeError(f"solve_gls_cv failed ({rc})")
    return {"theta": theta, "n_components": nc, "n_spatial": knots.shape[0], "lambda": scal[0],
            "gcv": scal[1], "effective_dof": scal[2], "rss": scal[3], "lambda_grid": grid,
            "gcv_curve": curve}


def assemble(knots, points, target, weights, bn, bits=64):
    lib = _lib(bits)
    _fit_sigs(lib)
    knots = np.ascontiguousarray(knots, np.float64)
    bn = np.ascontiguousarray(bn, np.float64)
    n, nc = bn.shape
    p = (nc + 1) * knots.shape[0]
    m, rhs, ywy = np.zeros(p * p), np.zeros(p), np.zeros(1)
    lib.or_assemble(knots, knots.shape[0], np.ascontiguousarray(points, np.float64),
                    np.ascontiguousarray(target, np.float64),
                    np.ascontiguousarray(weights, np.float64), bn, n, nc, m, rhs, ywy)
    return m.reshape(p, p), rhs, float(ywy[0])


def gather(science, reference, stamp_x, stamp_y, stamp_radius, beta, n_max, radius=0,
           science_var=None, reference_var=None, science_mask=None, reference_mask=None,
           bits=32):
    lib = _lib(bits)
    _fit_sigs(lib)
    sci = np.ascontiguousarray(science, np.float32)
    ref = np.ascontiguousarray(reference, np.float32)
    h, w = sci.shape
    radius = radius if radius > 0 else default_radius(beta, n_max)
    sx = np.ascontiguousarray(stamp_x, np.int32)
    sy = np.ascontiguousarray(stamp_y, np.int32)
    cap = max(1, sx.size * (2 * stamp_radius + 1) ** 2)
    nc = (n_max + 1) * (n_max + 2) // 2
    px, py, st = (np.zeros(cap, np.int32) for _ in range(3))
    tg, bn = np.zeros(cap), np.zeros(cap * nc)
    sv, rv, sm, rm = (_c32(science_var), _c32(reference_var), _cu8(science_mask),
                      _cu8(reference_mask))
    n = lib.or_gather(w, h, sci, ref, _ptr(sv), _ptr(rv), _ptr(sm), _ptr(rm), sx, sy, sx.size,
                      stamp_radius, beta, n_max, radius, px, py, st, tg, bn)
    return {"x": px[:n], "y": py[:n], "stamp": st[:n], "target": tg[:n],
            "bn": bn[:n * nc].reshape(n, nc)}


def fit_kernel(science, reference, knots, stamp_x, stamp_y, stamp_radius, beta, n_max,
               lambda_grid, radius=0, clip_sigma=4.0, clip_iterations=5, min_stamps=5,
               science_var=None, reference_var=None, science_mask=None, reference_mask=None,
               cv_folds=0, stamp_approx=-1, bits=32):
    """Oracle of ``_core.fit_kernel`` (src/bindings.cpp:643-693); ``stamp_approx``
    forces the per-stamp (1) or exact (0) CV build, -1 the knot-spacing gate."""
    lib = _lib(bits)
    _fit_sigs(lib)
    sci = np.ascontiguousarray(science, np.float32)
    ref = np.ascontiguousarray(reference, np.float32)
    h, w = sci.shape
    radius = radius if radius > 0 else default_radius(beta, n_max)
    knots = np.ascontiguousarray(knots, np.float64)
    k = knots.shape[0]
    nc = (n_max + 1) * (n_max + 2) // 2
    sx = np.ascontiguousarray(stamp_x, np.int32)
    sy = np.ascontiguousarray(stamp_y, np.int32)
    ns = sx.size
    grid = np.ascontiguousarray(lambda_grid, np.float64)
    theta, scal, curve = np.zeros((nc + 1) * k), np.zeros(6), np.zeros(grid.size)
    ints = np.zeros(5, np.int32)
    stx, sty = np.zeros(max(ns, 1), np.int32), np.zeros(max(ns, 1), np.int32)
    chi2, acc = np.zeros(max(ns, 1)), np.zeros(max(ns, 1), np.uint8)
    sv, rv, sm, rm = (_c32(science_var), _c32(reference_var), _cu8(science_mask),
                      _cu8(reference_mask))
    rc = lib.or_fit_kernel(w, h, sci, ref, _ptr(sv), _ptr(rv), _ptr(sm), _ptr(rm), knots, k, sx,
                           sy, ns, stamp_radius, beta, n_max, radius, grid, grid.size, clip_sigma,
                           clip_iterations, min_stamps, int(cv_folds), int(stamp_approx), theta,
                           scal, curve, ints, stx, sty, chi2, acc)
    if rc == -3:
        raise RuntimeError("fit_kernel: no usable stamp pixels")
    if rc == -4:
        raise RuntimeError("fit_kernel: all stamp pixels clipped")
    if rc != 0:
        raise ValueError(f"fit_kernel: invalid arguments ({rc})")
    s = ints[2]
    return {"theta": theta, "n_components": nc, "n_spatial": k, "lambda": scal[0],
            "gcv": scal[1], "effective_dof": scal[2], "rss": scal[3], "lambda_grid": grid,
            "gcv_curve": curve, "reduced_chi2": scal[4], "n_pixels": int(ints[0]),
            "n_stamps_used": int(ints[1]), "n_stamps_total": int(s),
            "n_stamps_rejected": int(ints[3]), "passes": int(ints[4]),
            "stamp_x": stx[:s], "stamp_y": sty[:s], "stamp_chi2": chi2[:s],
            "stamp_accepted": acc[:s], "cv_exact_design_bytes": int(scal[5])}


# --------------------------------------------------------------------------- detection
def seed_mask(score, threshold, polarity=1, mask=None):
    """Candidate bytes: finite, over threshold in the polarity's direction, unmasked."""
    s = np.ascontiguousarray(score, dtype=np.float32)
    m = _cu8(mask)
    out = np.zeros(s.shape, dtype=np.uint8)
    _lib(32).or_seed_mask(s, _ptr(m), s.size, threshold, polarity, out)
    return out


def label_seed(seed):
    """8-connected labels of a 0/1 image, numbered in raster order of first pixels."""
    s = np.ascontiguousarray(seed, dtype=np.uint8)
    h, w = s.shape
    out = np.zeros((h, w), dtype=np.int32)
    n = _lib(32).or_label_seed(s, w, h, out)
    return out, n


def label(score, threshold, polarity=1, mask=None):
    s = np.ascontiguousarray(score, dtype=np.float32)
    m = _cu8(mask)
    h, w = s.shape
    out = np.zeros((h, w), dtype=np.int32)
    n = _lib(32).or_label(s, _ptr(m), w, h, threshold, polarity, out)
    return out, n


def ref_label(score, threshold, polarity=1, mask=None):
    """The reference's own seed_mask + label_components (compiled catalog.cpp)."""
    lib = ref()
    s = np.ascontiguousarray(score, dtype=np.float32)
    m = _cu8(mask)
    h, w = s.shape
    out = np.zeros((h, w), dtype=np.int32)
    n = lib.ref_label(s, _ptr(m), w, h, threshold, polarity, out)
    return out, n


def matched_filter(image, psf, variance):
    img = np.ascontiguousarray(image, dtype=np.float32)
    p = np.ascontiguousarray(psf, dtype=np.float32)
    v = np.ascontiguousarray(variance, dtype=np.float32)
    h, w = img.shape
    out = np.zeros((h, w), dtype=np.float32)
    _lib(32).or_matched_filter(img, p, p.shape[0], v, w, h, out)
    return out


# --------------------------------------------------------------------------- adjoint
def gh_shifted(beta, n_max, radius, shift):
    """(g, dg/dx, dg/dbeta), each (n_max+1, ks), sampled at i - radius - shift."""
    ks = 2 * radius + 1
    g, dx, db = (np.zeros((n_max + 1) * ks) for _ in range(3))
    if _lib(64).or_gh_shifted(beta, n_max, radius, shift, _ptr(g), _ptr(dx), _ptr(db)):
        raise ValueError("gh_shifted: invalid arguments")
    return tuple(a.reshape(n_max + 1, ks) for a in (g, dx, db))


def shifted_kernel(params, n_max, radius, derivatives=False):
    """K (ks, ks) of params = [a_0..a_{nc-1}, b, beta, dx, dy]; with
    derivatives=True also (dK/dbeta, dK/ddx, dK/ddy) as (3, ks, ks)."""
    ks = 2 * radius + 1
    p = np.ascontiguousarray(params, np.float64)
    k2 = np.zeros(ks * ks)
    dk = np.zeros(3 * ks * ks) if derivatives else None
    if _lib(64).or_shifted_kernel(p, n_max, radius, k2, _ptr(dk)):
        raise ValueError("shifted_kernel: invalid arguments")
    k2 = k2.reshape(ks, ks)
    return (k2, dk.reshape(3, ks, ks)) if derivatives else k2


def kernel_chi2(science, reference, weight, params, n_max, radius, grad=True, diff=False,
                bits=64):
    """chi^2 = sum w (S - b - K (x) R)^2 of a constant shifted Gauss-Hermite
    kernel, its gradient with respect to params (nc + 4) and optionally D."""
    sci = np.ascontiguousarray(science, np.float32)
    ref = np.ascontiguousarray(reference, np.float32)
    h, w = sci.shape
    wt = _c32(weight)
    p = np.ascontiguousarray(params, np.float64)
    chi2 = np.zeros(1)
    g = np.zeros(p.size) if grad else None
    d = np.zeros((h, w), np.float32 if bits == 32 else np.float64) if diff else None
    rc = _lib(bits).or_kernel_chi2(w, h, sci, ref, _ptr(wt), p, n_max, radius, _ptr(chi2),
                                   _ptr(g), _ptr(d))
    if rc != 0:
        raise ValueError("kernel_chi2: invalid arguments")
    out = {"chi2": float(chi2[0])}
    if grad:
        out["grad"] = g
    if diff:
        out["difference"] = d
    return out


# --------------------------------------------------------------------------- stamp selection
class DetectParams(C.Structure):
    """Field-for-field ``delta::DetectParams`` (include/delta/detect.hpp:35-44)."""

    _fields_ = [("stamp_radius", C.c_int), ("threshold_sigma", C.c_double),
                ("max_stamps", C.c_int), ("saturation", C.c_double),
                ("isolation_radius", C.c_int), ("border", C.c_int),
                ("fwhm_radius", C.c_int), ("match_radius", C.c_int)]


def detect_params(stamp_radius=15, threshold_sigma=5.0, max_stamps=200, saturation=0.0,
                  isolation_radius=0, border=0, fwhm_radius=0, match_radius=2):
    return DetectParams(int(stamp_radius), float(threshold_sigma), int(max_stamps),
                        float(saturation), int(isolation_radius), int(border),
                        int(fwhm_radius), int(match_radius))


def _stamp_sigs(lib):
    if getattr(lib, "_stamp_ready", False):
        return
    lib.or_estimate_background.argtypes = [_f32p, C.c_void_p, C.c_int, C.c_int,
                                           C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.or_estimate_fwhm.restype = C.c_double
    lib.or_estimate_fwhm.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_double]
    lib.or_detect_stamps.restype = C.c_int
    lib.or_detect_stamps.argtypes = [_f32p, C.c_void_p, C.c_int, C.c_int,
                                     C.POINTER(DetectParams)] + [C.c_void_p] * 5
    lib.or_select_stamps.restype = C.c_int
    lib.or_select_stamps.argtypes = ([_f32p, C.c_void_p, C.c_int, C.c_int] * 2 +
                                     [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(DetectParams)] +
                                     [C.c_void_p] * 4 + [C.POINTER(C.c_double)] * 2 +
                                     [C.POINTER(C.c_int)])
    lib._stamp_ready = True


def estimate_background(data, mask=None):
    """(median, 1.4826 * MAD) of the decimated good sample (src/detect.cpp:84-113)."""
    lib = _lib(64)
    _stamp_sigs(lib)
    d = np.ascontiguousarray(data, np.float32)
    m = _cu8(mask)
    h, w = d.shape
    med, sig = C.c_double(), C.c_double()
    lib.or_estimate_background(d, _ptr(m), w, h, C.byref(med), C.byref(sig))
    return med.value, sig.value


def estimate_fwhm(data, x, y, radius, background):
    lib = _lib(64)
    _stamp_sigs(lib)
    d = np.ascontiguousarray(data, np.float32)
    h, w = d.shape
    return lib.or_estimate_fwhm(d, w, h, int(x), int(y), int(radius), float(background))


def _stamp_dict(n, cols):
    return {k: v[:n].copy() for k, v in cols.items()}


def detect_stamps(data, mask=None, **kw):
    """Stamps in selection order as the reference's dict (src/detect.cpp:159-241)."""
    lib = _lib(64)
    _stamp_sigs(lib)
    p = detect_params(**kw)
    d = np.ascontiguousarray(data, np.float32)
    m = _cu8(mask)
    h, w = d.shape
    cap = max(p.max_stamps, 0) + 1
    cols = {"x": np.zeros(cap, np.int32), "y": np.zeros(cap, np.int32),
            "flux": np.zeros(cap), "snr": np.zeros(cap), "fwhm": np.zeros(cap)}
    n = lib.or_detect_stamps(d, _ptr(m), w, h, C.byref(p), *[_ptr(cols[k]) for k in
                                                             ("x", "y", "flux", "snr", "fwhm")])
    if n < 0:
        raise ValueError("detect_stamps: border resolves below one pixel")
    return _stamp_dict(n, cols)


def select_stamps(science, reference, science_mask=None, reference_mask=None, catalog_x=None,
                  catalog_y=None, **kw):
    """Matched stamps, FWHMs, medians and direction (src/detect.cpp:243-314)."""
    lib = _lib(64)
    _stamp_sigs(lib)
    p = detect_params(**kw)
    s = np.ascontiguousarray(science, np.float32)
    r = np.ascontiguousarray(reference, np.float32)
    sm, rm = _cu8(science_mask), _cu8(reference_mask)
    cat = catalog_x is not None and catalog_y is not None
    cx = np.ascontiguousarray(catalog_x, np.int32) if cat else None
    cy = np.ascontiguousarray(catalog_y, np.int32) if cat else None
    ncat = min(cx.size, cy.size) if cat else 0
    cap = (ncat if cat else max(p.max_stamps, 0)) + 1
    cols = {"x": np.zeros(cap, np.int32), "y": np.zeros(cap, np.int32),
            "fwhm_science": np.zeros(cap), "fwhm_reference": np.zeros(cap)}
    ms, mr, dr = C.c_double(), C.c_double(), C.c_int()
    n = lib.or_select_stamps(s, _ptr(sm), s.shape[1], s.shape[0], r, _ptr(rm), r.shape[1],
                             r.shape[0], _ptr(cx), _ptr(cy), ncat, C.byref(p),
                             *[_ptr(cols[k]) for k in ("x", "y", "fwhm_science",
                                                       "fwhm_reference")],
                             C.byref(ms), C.byref(mr), C.byref(dr))
    if n < 0:
        raise ValueError("select_stamps: border resolves below one pixel")
    out = _stamp_dict(n, cols)
    out["median_fwhm_science"] = ms.value
    out["median_fwhm_reference"] = mr.value
    out["direction"] = "reference" if dr.value == 0 else "science"
    return out


def _ref_params(p):
    ip = np.array([p.stamp_radius, p.max_stamps, p.isolation_radius, p.border, p.fwhm_radius,
                   p.match_radius], np.int32)
    dp = np.array([p.threshold_sigma, p.saturation], np.float64)
    return ip, dp


def _ref_stamp_sigs(lib):
    if getattr(lib, "_stamp_ready", False):
        return
    lib.ref_estimate_background.argtypes = [_f32p, C.c_void_p, C.c_int, C.c_int, _f64p]
    lib.ref_estimate_fwhm.restype = C.c_double
    lib.ref_estimate_fwhm.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double]
    lib.ref_detect_stamps.restype = C.c_int
    lib.ref_detect_stamps.argtypes = ([_f32p, C.c_void_p, C.c_int, C.c_int, _i32p, _f64p,
                                       C.c_int] + [C.c_void_p] * 5)
    lib.ref_select_stamps.restype = C.c_int
    lib.ref_select_stamps.argtypes = ([_f32p, C.c_void_p, C.c_int, C.c_int] * 2 +
                                      [C.c_void_p, C.c_void_p, C.c_int, _i32p, _f64p, C.c_int] +
                                      [C.c_void_p] * 4 + [_f64p, C.POINTER(C.c_int)])
    lib._stamp_ready = True


def ref_estimate_background(data, mask=None):
    """The reference's own estimate_background (compiled detect.cpp)."""
    lib = ref()
    _ref_stamp_sigs(lib)
    d = np.ascontiguousarray(data, np.float32)
    m = _cu8(mask)
    out = np.zeros(2)
    lib.ref_estimate_background(d, _ptr(m), d.shape[1], d.shape[0], out)
    return float(out[0]), float(out[1])


def ref_estimate_fwhm(data, x, y, radius, background):
    lib = ref()
    _ref_stamp_sigs(lib)
    d = np.ascontiguousarray(data, np.float32)
    return lib.ref_estimate_fwhm(d, d.shape[1], d.shape[0], int(x), int(y), int(radius),
                                 float(background))


def ref_detect_stamps(data, mask=None, **kw):
    """The reference's own detect_stamps (compiled detect.cpp)."""
    lib = ref()
    _ref_stamp_sigs(lib)
    p = detect_params(**kw)
    ip, dp = _ref_params(p)
    d = np.ascontiguousarray(data, np.float32)
    m = _cu8(mask)
    cap = max(p.max_stamps, 0) + 1
    cols = {"x": np.zeros(cap, np.int32), "y": np.zeros(cap, np.int32),
            "flux": np.zeros(cap), "snr": np.zeros(cap), "fwhm": np.zeros(cap)}
    n = lib.ref_detect_stamps(d, _ptr(m), d.shape[1], d.shape[0], ip, dp, cap,
                              *[_ptr(cols[k]) for k in ("x", "y", "flux", "snr", "fwhm")])
    return _stamp_dict(n, cols)


def ref_select_stamps(science, reference, science_mask=None, reference_mask=None,
                      catalog_x=None, catalog_y=None, **kw):
    """The reference's own select_stamps / select_stamps_from_catalog."""
    lib = ref()
    _ref_stamp_sigs(lib)
    p = detect_params(**kw)
    ip, dp = _ref_params(p)
    s = np.ascontiguousarray(science, np.float32)
    r = np.ascontiguousarray(reference, np.float32)
    sm, rm = _cu8(science_mask), _cu8(reference_mask)
    cat = catalog_x is not None and catalog_y is not None
    cx = np.ascontiguousarray(catalog_x, np.int32) if cat else None
    cy = np.ascontiguousarray(catalog_y, np.int32) if cat else None
    ncat = min(cx.size, cy.size) if cat else 0
    cap = (ncat if cat else max(p.max_stamps, 0)) + 1
    cols = {"x": np.zeros(cap, np.int32), "y": np.zeros(cap, np.int32),
            "fwhm_science": np.zeros(cap), "fwhm_reference": np.zeros(cap)}
    med = np.zeros(2)
    dr = C.c_int()
    n = lib.ref_select_stamps(s, _ptr(sm), s.shape[1], s.shape[0], r, _ptr(rm), r.shape[1],
                              r.shape[0], _ptr(cx), _ptr(cy), ncat, ip, dp, cap,
                              *[_ptr(cols[k]) for k in ("x", "y", "fwhm_science",
                                                        "fwhm_reference")],
                              med, C.byref(dr))
    out = _stamp_dict(n, cols)
    out["median_fwhm_science"] = float(med[0])
    out["median_fwhm_reference"] = float(med[1])
    out["direction"] = "reference" if dr.value == 0 else "science"
    return out


# --------------------------------------------------------------------------- catalog
CATALOG_ENTRY = np.dtype([
    ("x", "f8"), ("y", "f8"), ("peak_snr", "f8"), ("flux", "f8"), ("expected_n_pix", "f8"),
    ("fwhm_ratio", "f8"), ("psf_chi2", "f8"), ("psf_amplitude", "f8"),
    ("psf_amplitude_err", "f8"), ("bright_star_dist", "f8"), ("peak_x", "i4"), ("peak_y", "i4"),
    ("n_pix", "i4"), ("psf_chi2_dof", "i4"), ("fwhm_consistent", "u1"), ("mask_flags", "u1"),
    ("is_dipole", "u1"), ("shape_consistent", "u1"), ("near_bright_star", "u1"),
    ("quality", "u1"), ("pad", "u1", (2,))])
CATALOG_COLUMNS = [n for n in CATALOG_ENTRY.names if n != "pad"]


class CatalogParams(C.Structure):
    """Field for field ``delta::CatalogParams`` (include/delta/catalog.hpp)."""

    _fields_ = [("threshold_sigma", C.c_double), ("threshold_sigma_dipole", C.c_double),
                ("expected_fwhm", C.c_double), ("fwhm_tolerance_lo", C.c_double),
                ("fwhm_tolerance_hi", C.c_double), ("aperture_radius", C.c_int),
                ("exclude_bad_pixels", C.c_int), ("fit_psf_shape", C.c_int),
                ("shape_chi2_tolerance", C.c_double), ("bright_star_radius", C.c_double),
                ("max_shape_fit", C.c_int)]


_CAT_DEFAULTS = dict(mask=None, threshold_sigma=5.0, threshold_sigma_dipole=3.0,
                     expected_fwhm=3.5, fwhm_tolerance_lo=0.3, fwhm_tolerance_hi=3.0,
                     aperture_radius=0, exclude_bad_pixels=True, return_negative=False,
                     fit_psf_shape=False, shape_chi2_tolerance=3.0, variance=None, psf=None,
                     bright_star_radius=0.0, max_shape_fit=5000, bright_x=None, bright_y=None)


def _cat_args(score, difference, kw):
    a = dict(_CAT_DEFAULTS)
    a.update(kw)
    s = np.ascontiguousarray(score, np.float32)
    d = np.ascontiguousarray(difference, np.float32)
    m = _cu8(a["mask"])
    v = _c32(a["variance"])
    p = _c32(a["psf"])
    bx = None if a["bright_x"] is None else np.ascontiguousarray(a["bright_x"], np.float64)
    by = None if a["bright_y"] is None else np.ascontiguousarray(a["bright_y"], np.float64)
    return a, s, d, m, v, p, bx, by


def _cat_columns(raw, n):
    return {k: raw[k][:n].copy() for k in CATALOG_COLUMNS}


def build_catalog(score, difference, **kw):
    """Column dict of build_catalog (src/catalog.cpp:600-697, bindings.cpp:348-483)."""
    lib = _lib(64)
    if not getattr(lib, "_cat_ready", False):
        lib.or_build_catalog.restype = C.c_int
        lib.or_build_catalog.argtypes = ([_f32p, _f32p, C.c_void_p, C.c_int, C.c_int,
                                          C.POINTER(CatalogParams), C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_int])
        lib._cat_ready = True
    a, s, d, m, v, p, bx, by = _cat_args(score, difference, kw)
    h, w = s.shape
    cp = CatalogParams(a["threshold_sigma"], a["threshold_sigma_dipole"], a["expected_fwhm"],
                       a["fwhm_tolerance_lo"], a["fwhm_tolerance_hi"], int(a["aperture_radius"]),
                       int(bool(a["exclude_bad_pixels"])), int(bool(a["fit_psf_shape"])),
                       a["shape_chi2_tolerance"], a["bright_star_radius"], int(a["max_shape_fit"]))
    cap = 4096
    while True:
        out = np.zeros(cap, CATALOG_ENTRY)
        n = lib.or_build_catalog(s, d, _ptr(m), w, h, C.byref(cp), int(bool(a["return_negative"])),
                                 _ptr(v), _ptr(p), 0 if p is None else p.shape[0], _ptr(bx),
                                 _ptr(by), 0 if bx is None else bx.size, _ptr(out), cap)
        if n <= cap:
            return _cat_columns(out, n)
        cap = n


def ref_build_catalog(score, difference, **kw):
    """The reference's own build_catalog (compiled catalog.cpp, one thread)."""
    lib = ref()
    if not getattr(lib, "_cat_ready", False):
        lib.ref_build_catalog.restype = C.c_int
        lib.ref_build_catalog.argtypes = ([_f32p, _f32p, C.c_void_p, C.c_int, C.c_int, _f64p,
                                           _i32p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int])
        lib._cat_ready = True
    a, s, d, m, v, p, bx, by = _cat_args(score, difference, kw)
    h, w = s.shape
    dp = np.array([a["threshold_sigma"], a["threshold_sigma_dipole"], a["expected_fwhm"],
                   a["fwhm_tolerance_lo"], a["fwhm_tolerance_hi"], a["shape_chi2_tolerance"],
                   a["bright_star_radius"]], np.float64)
    ip = np.array([a["aperture_radius"], int(bool(a["exclude_bad_pixels"])),
                   int(bool(a["fit_psf_shape"])), a["max_shape_fit"]], np.int32)
    cap = 4096
    while True:
        out = np.zeros(cap, CATALOG_ENTRY)
        n = lib.ref_build_catalog(s, d, _ptr(m), w, h, dp, ip, int(bool(a["return_negative"])),
                                  _ptr(v), _ptr(p), 0 if p is None else p.shape[0], _ptr(bx),
                                  _ptr(by), 0 if bx is None else bx.size, _ptr(out), cap)
        if n <= cap:
            return _cat_columns(out, n)
        cap = n
