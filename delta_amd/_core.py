"""ctypes binding of libdeltagpu with the reference's ``_core`` call signatures.

Each function here mirrors the nanobind entry point of the same name in the
reference (src/bindings.cpp) -- argument names, defaults, dict keys and error
types -- and calls the C ABI declared in include/deltagpu.h.  The library is
loaded from this package directory; a missing library raises ImportError at
import time (there is no CPU fallback).

Two calling forms exist for the image-sized entry points:
  * numpy arrays (host memory), as the reference binding takes them;
  * ``*_device`` helpers taking torch CUDA tensors, for device-resident
    pipelines (the data never leaves HBM).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libdeltagpu.so")
if not os.path.exists(_LIB_PATH):
    raise ImportError(
        f"libdeltagpu.so not found at {_LIB_PATH}; build it with "
        "`python -m delta_amd.build` (hipcc, gfx950)")
_lib = C.CDLL(_LIB_PATH)

DG_OK, DG_ERR_INVALID, DG_ERR_RUNTIME, DG_ERR_DEVICE = 0, 1, 2, 3

_vp = C.c_void_p
_i, _d = C.c_int, C.c_double

_lib.dg_version.restype = _i
_lib.dg_last_error.restype = C.c_char_p
_lib.dg_default_radius.restype = _i
_lib.dg_default_radius.argtypes = [_d, _i]
_SUB_ARGS = [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _d, _i, _i, _d, _vp, _vp, _vp]
_lib.dg_subtract.restype = _i
_lib.dg_subtract.argtypes = _SUB_ARGS
_lib.dg_subtract_device.restype = _i
_lib.dg_subtract_device.argtypes = _SUB_ARGS + [_vp]
_FIT_ARGS = [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _i, _i, _d, _i, _i, _vp,
             _i, _d, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
_lib.dg_fit_kernel.restype = _i
_lib.dg_fit_kernel.argtypes = _FIT_ARGS
_lib.dg_fit_kernel_device.restype = _i
_lib.dg_fit_kernel_device.argtypes = _FIT_ARGS + [_vp]
_lib.dg_solve_gls_gcv.restype = _i
_lib.dg_solve_gls_gcv.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _i, _i, _vp, _vp, _vp]
_lib.dg_solve_gls_cv.restype = _i
_lib.dg_solve_gls_cv.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _i, _vp, _i, _i, _vp,
                                 _vp, _vp]
_DEC_ARGS = [_i, _i, _vp, _vp, _i, _vp, _i, _d, _i, _vp, _vp, _i, _i, _i, _vp, _vp]
_lib.dg_decorrelate.restype = _i
_lib.dg_decorrelate.argtypes = _DEC_ARGS
_lib.dg_decorrelate_device.restype = _i
_lib.dg_decorrelate_device.argtypes = _DEC_ARGS + [_vp]
_lib.dg_whiten_score_psf.restype = _i
_lib.dg_whiten_score_psf.argtypes = [_vp, _i, _vp, _i, _vp, _i, _d, _i, _vp, _vp, _i, _i, _i, _i,
                                     _vp]
_lib.dg_whiten_score_psf_device.restype = _i
_lib.dg_whiten_score_psf_device.argtypes = _lib.dg_whiten_score_psf.argtypes + [_vp]
_lib.dg_gauss_hermite_basis1d.restype = _i
_lib.dg_gauss_hermite_basis1d.argtypes = [_d, _i, _i, _vp]
_lib.dg_component_sums.restype = _i
_lib.dg_component_sums.argtypes = [_d, _i, _i, _vp]
_lib.dg_gauss_hermite_kernels.restype = _i
_lib.dg_gauss_hermite_kernels.argtypes = [_d, _i, _i, _vp, _vp]
_lib.dg_basis_convolve.restype = _i
_lib.dg_basis_convolve.argtypes = [_i, _i, _vp, _d, _i, _i, _vp]
_lib.dg_basis_convolve_device.restype = _i
_lib.dg_basis_convolve_device.argtypes = [_i, _i, _vp, _d, _i, _i, _vp, _vp]
_lib.dg_tps_design.restype = _i
_lib.dg_tps_design.argtypes = [_vp, _i, _vp, _i, _vp]
_lib.dg_tps_evaluate.restype = _i
_lib.dg_tps_evaluate.argtypes = [_vp, _i, _vp, _i, _vp, _i, _vp]
_lib.dg_tps_penalty.restype = _i
_lib.dg_tps_penalty.argtypes = [_vp, _i, _vp]
_lib.dg_tps_fit.restype = _i
_lib.dg_tps_fit.argtypes = [_vp, _i, _vp, _vp, _i, _d, _vp]
_lib.dg_decorrelate_block.restype = _i
_lib.dg_decorrelate_block.argtypes = [_i, _vp, _vp, _i, _d, _d, _vp]
_lib.dg_decorrelation_kernel.restype = _i
_lib.dg_decorrelation_kernel.argtypes = [_vp, _i, _d, _d, _i, _vp]
_lib.dg_synth_variance.restype = _i
_lib.dg_synth_variance.argtypes = [_i, _i, _vp, _d, _d, _vp]
_lib.dg_synth_variance_device.restype = _i
_lib.dg_synth_variance_device.argtypes = [_i, _i, _vp, _d, _d, _vp, _vp]
_lib.dg_variance_rescale_factor.restype = _i
_lib.dg_variance_rescale_factor.argtypes = [_i, _i, _vp, _vp, _vp, C.POINTER(_d)]
_lib.dg_variance_rescale_factor_device.restype = _i
_lib.dg_variance_rescale_factor_device.argtypes = [_i, _i, _vp, _vp, _vp, C.POINTER(_d), _vp]
_lib.dg_estimate_psf.restype = _i
_lib.dg_estimate_psf.argtypes = [_i, _i, _vp, _vp, _vp, _i, _i, _vp]
_lib.dg_estimate_psf_device.restype = _i
_lib.dg_estimate_psf_device.argtypes = [_i, _i, _vp, _vp, _vp, _i, _i, _vp, _vp]
_lib.dg_weighted_mean.restype = _i
_lib.dg_weighted_mean.argtypes = [_i, _i, _vp, _vp, _vp, C.POINTER(_d)]
_lib.dg_matched_filter.restype = _i
_lib.dg_matched_filter.argtypes = [_i, _i, _vp, _vp, _i, _vp, _vp]
_lib.dg_matched_filter_device.restype = _i
_lib.dg_matched_filter_device.argtypes = [_i, _i, _vp, _vp, _i, _vp, _vp, _vp]
_lib.dg_label.restype = _i
_lib.dg_label.argtypes = [_i, _i, _vp, _vp, _d, _i, _vp, C.POINTER(_i)]
_lib.dg_label_device.restype = _i
_lib.dg_label_device.argtypes = [_i, _i, _vp, _vp, _d, _i, _vp, C.POINTER(_i), _vp]
_KC_ARGS = [_i, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]
_lib.dg_kernel_chi2.restype = _i
_lib.dg_kernel_chi2.argtypes = _KC_ARGS
_lib.dg_kernel_chi2_device.restype = _i
_lib.dg_kernel_chi2_device.argtypes = _KC_ARGS + [_vp]
_lib.dg_kernel_difference_device.restype = _i
_lib.dg_kernel_difference_device.argtypes = [_i, _i, _vp, _vp, _vp, _i, _i, _vp, _vp]
_lib.dg_kernel_difference_vjp_device.restype = _i
_lib.dg_kernel_difference_vjp_device.argtypes = [_i, _i, _vp, _vp, _vp, _i, _i, _vp, _vp]


class DetectParams(C.Structure):
    """``dg_detect_params`` (include/deltagpu.h), field for field
    ``delta::DetectParams`` (include/delta/detect.hpp:35-44)."""

    _fields_ = [("stamp_radius", _i), ("threshold_sigma", _d), ("max_stamps", _i),
                ("saturation", _d), ("isolation_radius", _i), ("border", _i),
                ("fwhm_radius", _i), ("match_radius", _i)]


_DP = C.POINTER(DetectParams)
_pd, _pi = C.POINTER(_d), C.POINTER(_i)
_lib.dg_estimate_background.restype = _i
_lib.dg_estimate_background.argtypes = [_i, _i, _vp, _vp, _pd, _pd]
_lib.dg_estimate_background_device.restype = _i
_lib.dg_estimate_background_device.argtypes = [_i, _i, _vp, _vp, _pd, _pd, _vp]
_DET_ARGS = [_i, _i, _vp, _vp, _DP, _vp, _vp, _vp, _vp, _vp, _i, _pi]
_lib.dg_detect_stamps.restype = _i
_lib.dg_detect_stamps.argtypes = _DET_ARGS
_lib.dg_detect_stamps_device.restype = _i
_lib.dg_detect_stamps_device.argtypes = _DET_ARGS + [_vp]
_SEL_ARGS = [_i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _i, _DP, _vp, _vp, _vp, _vp, _i, _pi,
             _vp, _pi]
_lib.dg_select_stamps.restype = _i
_lib.dg_select_stamps.argtypes = _SEL_ARGS
_lib.dg_select_stamps_device.restype = _i
_lib.dg_select_stamps_device.argtypes = _SEL_ARGS + [_vp]


class CatalogParams(C.Structure):
    """``dg_catalog_params`` (include/deltagpu.h), field for field
    ``delta::CatalogParams`` (include/delta/catalog.hpp:57-84)."""

    _fields_ = [("threshold_sigma", _d), ("threshold_sigma_dipole", _d), ("expected_fwhm", _d),
                ("fwhm_tolerance_lo", _d), ("fwhm_tolerance_hi", _d), ("aperture_radius", _i),
                ("exclude_bad_pixels", _i), ("fit_psf_shape", _i),
                ("shape_chi2_tolerance", _d), ("bright_star_radius", _d),
                ("max_shape_fit", _i)]


# ``dg_catalog_entry``: doubles, int32, bytes -- 104 bytes, no interior padding.
CATALOG_ENTRY = np.dtype([
    ("x", "f8"), ("y", "f8"), ("peak_snr", "f8"), ("flux", "f8"), ("expected_n_pix", "f8"),
    ("fwhm_ratio", "f8"), ("psf_chi2", "f8"), ("psf_amplitude", "f8"),
    ("psf_amplitude_err", "f8"), ("bright_star_dist", "f8"), ("peak_x", "i4"), ("peak_y", "i4"),
    ("n_pix", "i4"), ("psf_chi2_dof", "i4"), ("fwhm_consistent", "u1"), ("mask_flags", "u1"),
    ("is_dipole", "u1"), ("shape_consistent", "u1"), ("near_bright_star", "u1"),
    ("quality", "u1"), ("pad", "u1", (2,))])
assert CATALOG_ENTRY.itemsize == 104
_CAT_ARGS = [_i, _i, _vp, _vp, _vp, C.POINTER(CatalogParams), _i, _vp, _vp, _i, _vp, _vp, _i,
             _vp, _i, C.POINTER(_i)]
_lib.dg_build_catalog.restype = _i
_lib.dg_build_catalog.argtypes = _CAT_ARGS
_lib.dg_build_catalog_device.restype = _i
_lib.dg_build_catalog_device.argtypes = _CAT_ARGS + [_vp]
_lib.dg_build_catalog_labelled_device.restype = _i
_lib.dg_build_catalog_labelled_device.argtypes = _CAT_ARGS[:13] + [_vp, _i] + _CAT_ARGS[13:] + [_vp]
_PS_ARGS = [_vp, _i, _vp, _i, _vp, _i, _i, _i, _vp]
_lib.dg_photometric_scale.restype = _i
_lib.dg_photometric_scale.argtypes = _PS_ARGS
_lib.dg_photometric_scale_device.restype = _i
_lib.dg_photometric_scale_device.argtypes = _PS_ARGS + [_vp]
_lib.dg_photometric_scale_at.restype = _i
_lib.dg_photometric_scale_at.argtypes = [_vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp]
_lib.dg_profile_enable.argtypes = [_i]
_lib.dg_profile_read.argtypes = [C.c_char_p, C.POINTER(_i), C.POINTER(_d)]
_lib.dg_profile_read.restype = _i


def profile_enable(on: bool = True) -> None:
    """Record a HIP event pair around each instrumented kernel launch."""
    _lib.dg_profile_enable(1 if on else 0)


def profile_reset() -> None:
    _lib.dg_profile_reset()


def profile_read(name: str):
    """(launch count, total milliseconds) of the named kernel since reset."""
    n, ms = _i(0), _d(0.0)
    _check(_lib.dg_profile_read(name.encode(), C.byref(n), C.byref(ms)))
    return n.value, ms.value


# --------------------------------------------------------------------------- stage timers
# With DELTA_TIMING set (non-empty and not "0", read once, as
# src/timing.cpp:19-25) every call clears the timers and the result dicts
# carry {label: seconds} of the device stages that ran (src/bindings.cpp:
# 591-599); drain_timing() returns the same for calls that return bare
# arrays.  Stages with a counterpart carry the reference's DELTA_TIME labels;
# the rest are named in the same "stage: step" form.  Times are the HIP event
# intervals of the instrumented launches (enclosing scopes are left out, so
# nothing is counted twice).
_TIMING_ON = None
_STAGE_LABELS = {
    "select": "subtract: sanitise reference",
    "fields": "subtract: spatial fields",
    "model": "subtract: model convolve",
    "variance": "subtract: variance propagation",
    "mask": "subtract: mask growth",
    "stamp_bn": "fit: stamp B_n convolve",
    "assemble": "fit: normal equations",
    "gcv": "fit: lambda grid solves",
    "ldlt": "fit: LDLT fallback",
    "rss": "fit: residual sums",
    "resid": "fit: residual sums",
    "chi2": "fit: stamp chi2",
    "reweight": "fit: IRLS reweight",
    "cv_folds": "fit: cross-validation folds",
    "cv_solve": "fit: cross-validation solves",
    "decorrelate_cells": "decorrelate: kernel-power cache",
    "decorrelate_noise": "decorrelate: block filter+blend",
    "decorrelate_blocks": "decorrelate: block filter+blend",
    "decorrelate_normalise": "decorrelate: normalise",
    "decorrelate_variance": "decorrelate: normalise",
    "whiten_psf": "score: whitened psf",
    "matched_filter": "score: blocked correlation",
    "label": "catalog: label positive",
    "catalog": "catalog: rows",
    "basis_convolve": "basis: convolve",
    "background": "detect: background",
    "detect": "detect: peaks",
    "select_stamps": "detect: select stamps",
}


def timing_enabled() -> bool:
    global _TIMING_ON
    if _TIMING_ON is None:
        v = os.environ.get("DELTA_TIMING", "")
        _TIMING_ON = v != "" and v != "0"
    return _TIMING_ON


def _timing_clear() -> None:
    if timing_enabled():
        _lib.dg_profile_reset()
        _lib.dg_profile_enable(1)


def drain_timing():
    """{label: seconds} of the device stages since the last call cleared the
    timers, or None when DELTA_TIMING is off (``_core.drain_timing``,
    src/bindings.cpp:1032-1035)."""
    if not timing_enabled():
        return None
    out = {}
    for name, label in _STAGE_LABELS.items():
        c, ms = profile_read(name)
        if c:
            out[label] = out.get(label, 0.0) + ms * 1e-3
    return out


class DeviceError(RuntimeError):
    """A HIP runtime failure inside libdeltagpu."""


def _check(rc: int) -> None:
    if rc == DG_OK:
        return
    msg = (_lib.dg_last_error() or b"").decode()
    if rc == DG_ERR_INVALID:
        raise ValueError(msg)
    if rc == DG_ERR_RUNTIME:
        raise RuntimeError(msg)
    raise DeviceError(msg)


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _in2d(a, dtype, name):
    """C-contiguous 2-D host array (the reference's InArray contract)."""
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.ndim != 2:
        raise ValueError(f"{name} must be 2-D")
    return a


def _match(a, shape, what):
    if a is not None and a.shape != shape:
        raise RuntimeError(f"{what} shape does not match data")


def version() -> int:
    return int(_lib.dg_version())


def default_radius(beta: float, n_max: int) -> int:
    return int(_lib.dg_default_radius(float(beta), int(n_max)))


def library_path() -> str:
    return _LIB_PATH


# --------------------------------------------------------------------------- subtract
def subtract(science, reference, knots, theta, beta, n_max, radius=0, saturation=0.0,
             science_var=None, reference_var=None, science_mask=None, reference_mask=None):
    """Full-frame subtraction -> {'timing', 'difference', 'variance', 'mask'}.

    Mirrors ``_core.subtract`` (src/bindings.cpp:603-638): variance is None
    unless a variance layer is given, mask is None unless a mask layer is
    given or saturation > 0.
    """
    sci = _in2d(science, np.float32, "science")
    ref = _in2d(reference, np.float32, "reference")
    h, w = sci.shape
    if ref.shape != sci.shape:
        raise RuntimeError("subtract: science/reference shape mismatch")
    sv = _in2d(science_var, np.float32, "science_var")
    rv = _in2d(reference_var, np.float32, "reference_var")
    sm = _in2d(science_mask, np.uint8, "science_mask")
    rm = _in2d(reference_mask, np.uint8, "reference_mask")
    for a, what in ((sv, "variance"), (rv, "variance"), (sm, "mask"), (rm, "mask")):
        _match(a, (h, w), what)
    kn = np.ascontiguousarray(knots, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    if kn.ndim != 2 or kn.shape[1] != 2:
        raise ValueError("knots must be (k, 2)")
    diff = np.empty((h, w), dtype=np.float32)
    var = np.empty((h, w), dtype=np.float32) if (sv is not None or rv is not None) else None
    mask = (np.empty((h, w), dtype=np.uint8)
            if (sm is not None or rm is not None or saturation > 0) else None)
    _timing_clear()
    _check(_lib.dg_subtract(w, h, _p(sci), _p(ref), _p(sv), _p(rv), _p(sm), _p(rm), _p(kn),
                            kn.shape[0], _p(th), th.size, float(beta), int(n_max), int(radius),
                            float(saturation), _p(diff), _p(var), _p(mask)))
    return {"timing": drain_timing(), "difference": diff, "variance": var, "mask": mask}


def _tp(t):
    return None if t is None else _vp(t.data_ptr())


def subtract_device(science, reference, knots, theta, beta, n_max, radius=0, saturation=0.0,
                    science_var=None, reference_var=None, science_mask=None,
                    reference_mask=None, out=None, stream=None):
    """Device-resident subtraction on torch CUDA tensors (float32 / uint8).

    ``out`` may supply preallocated {'difference', 'variance', 'mask'} tensors.
    Runs on ``stream`` (default: torch's current stream) and returns once the
    stream has drained.
    """
    import torch

    h, w = science.shape
    for t in (science, reference, science_var, reference_var):
        if t is not None and (t.dtype != torch.float32 or not t.is_contiguous()
                              or t.shape != (h, w) or not t.is_cuda):
            raise ValueError("device images must be contiguous float32 CUDA tensors of one shape")
    for t in (science_mask, reference_mask):
        if t is not None and (t.dtype != torch.uint8 or not t.is_contiguous()
                              or t.shape != (h, w) or not t.is_cuda):
            raise ValueError("device masks must be contiguous uint8 CUDA tensors")
    want_var = science_var is not None or reference_var is not None
    want_mask = science_mask is not None or reference_mask is not None or saturation > 0
    out = dict(out or {})
    dev = science.device
    if out.get("difference") is None:
        out["difference"] = torch.empty((h, w), dtype=torch.float32, device=dev)
    if want_var and out.get("variance") is None:
        out["variance"] = torch.empty((h, w), dtype=torch.float32, device=dev)
    if want_mask and out.get("mask") is None:
        out["mask"] = torch.empty((h, w), dtype=torch.uint8, device=dev)
    if not want_var:
        out["variance"] = None
    if not want_mask:
        out["mask"] = None
    kn = np.ascontiguousarray(knots, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    _timing_clear()
    _check(_lib.dg_subtract_device(
        w, h, _tp(science), _tp(reference), _tp(science_var), _tp(reference_var),
        _tp(science_mask), _tp(reference_mask), _p(kn), kn.shape[0], _p(th), th.size,
        float(beta), int(n_max), int(radius), float(saturation), _tp(out["difference"]),
        _tp(out["variance"]), _tp(out["mask"]), _vp(s.cuda_stream)))
    out["timing"] = drain_timing()
    return out


# --------------------------------------------------------------------------- kernel fit
PRECISIONS = {"fp32": 0, "mixed": 1, "split_bf16": 1, "fp64": 2}


def _precision(p) -> int:
    if isinstance(p, int):
        return p
    try:
        return PRECISIONS[p]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}") from None


def _fit_outputs(n_theta, n_lambda, n_stamps):
    return (np.empty(n_theta), np.empty(n_lambda), np.empty(6), np.zeros(5, np.int32),
            np.zeros(max(n_stamps, 1), np.int32), np.zeros(max(n_stamps, 1), np.int32),
            np.zeros(max(n_stamps, 1)), np.zeros(max(n_stamps, 1), np.uint8))


def _fit_dict(theta, curve, scal, cnt, sx, sy, chi2, acc, grid, nc, k, comp_sums):
    s = int(cnt[2])
    return {"theta": theta, "n_components": nc, "n_spatial": k, "lambda": float(scal[0]),
            "gcv": float(scal[1]), "effective_dof": float(scal[2]), "rss": float(scal[3]),
            "lambda_grid": grid, "gcv_curve": curve, "timing": drain_timing(),
            "component_sums": comp_sums, "n_pixels": int(cnt[0]),
            "n_stamps_used": int(cnt[1]), "n_stamps_total": s,
            "n_stamps_rejected": int(cnt[3]), "reduced_chi2": float(scal[4]),
            "cv_exact_design_bytes": int(scal[5]), "stamp_x": sx[:s].copy(),
            "stamp_y": sy[:s].copy(),
            "stamp_chi2": chi2[:s].copy(), "stamp_accepted": acc[:s].copy(),
            "irls_passes": int(cnt[4])}


def component_sums(beta, n_max, radius=0):
    """Footprint sums S_c = (sum g_nx)(sum g_ny), each 1-D sum in index order
    (src/basis.cpp:81-93), from the library's basis tables."""
    nc = (int(n_max) + 1) * (int(n_max) + 2) // 2
    out = np.empty(nc)
    _check(_lib.dg_component_sums(float(beta), int(n_max), int(radius), _p(out)))
    return out


def fit_kernel(science, reference, knots, stamp_x, stamp_y, stamp_radius, beta, n_max,
               lambda_grid, radius=0, clip_sigma=4.0, clip_iterations=5, min_stamps=5, cv_folds=0,
               science_var=None, reference_var=None, science_mask=None, reference_mask=None,
               precision="fp64"):
    """Kernel fit -> the reference's result dict (src/bindings.cpp:643-693).

    ``science`` is the fit target and ``reference`` the image convolved, as in
    the reference binding.  ``cv_folds`` > 1 selects lambda by leave-stamp-out
    CV (``gcv_curve`` is then the CV error), otherwise GCV.  ``precision``
    selects the M = X^T W X build: "fp64" (FP64 MFMA; reproduces the
    reference's FP64 normal equations), "fp32" (FP32 MFMA) or "mixed"
    (split-BF16 MFMA); the latter two are faster approximations of an
    ill-conditioned system.
    """
    sci = _in2d(science, np.float32, "science")
    ref = _in2d(reference, np.float32, "reference")
    h, w = sci.shape
    if ref.shape != sci.shape:
        raise RuntimeError("fit_kernel: science/reference shape mismatch")
    sv = _in2d(science_var, np.float32, "science_var")
    rv = _in2d(reference_var, np.float32, "reference_var")
    sm = _in2d(science_mask, np.uint8, "science_mask")
    rm = _in2d(reference_mask, np.uint8, "reference_mask")
    for a, what in ((sv, "variance"), (rv, "variance"), (sm, "mask"), (rm, "mask")):
        _match(a, (h, w), what)
    sx = np.ascontiguousarray(stamp_x, np.int32).ravel()
    sy = np.ascontiguousarray(stamp_y, np.int32).ravel()
    if sx.size != sy.size:
        raise RuntimeError("fit_kernel: stamp_x/stamp_y length mismatch")
    kn = np.ascontiguousarray(knots, np.float64)
    grid = np.ascontiguousarray(lambda_grid, np.float64).ravel()
    nc = (n_max + 1) * (n_max + 2) // 2
    outs = _fit_outputs((nc + 1) * kn.shape[0], grid.size, sx.size)
    _timing_clear()
    _check(_lib.dg_fit_kernel(w, h, _p(sci), _p(ref), _p(sv), _p(rv), _p(sm), _p(rm), _p(kn),
                              kn.shape[0], _p(sx), _p(sy), sx.size, int(stamp_radius),
                              float(beta), int(n_max), int(radius), _p(grid), grid.size,
                              float(clip_sigma), int(clip_iterations), int(min_stamps),
                              int(cv_folds), _precision(precision), *[_p(o) for o in outs]))
    return _fit_dict(*outs, grid, nc, kn.shape[0], component_sums(beta, n_max, radius))


def fit_kernel_device(science, reference, knots, stamp_x, stamp_y, stamp_radius, beta, n_max,
                      lambda_grid, radius=0, clip_sigma=4.0, clip_iterations=5, min_stamps=5,
                      science_var=None, reference_var=None, science_mask=None,
                      reference_mask=None, precision="fp64", stream=None, cv_folds=0):
    """fit_kernel on torch CUDA tensors (images resident in HBM).  Every frame
    is checked (contiguous float32 / uint8 CUDA tensors of the science
    frame's shape) before anything is launched."""
    import torch

    _dev_any_image(science, "science")
    h, w = science.shape
    _dev_image(reference, h, w, "reference")
    for t, name in ((science_var, "science_var"), (reference_var, "reference_var")):
        if t is not None:
            _dev_image(t, h, w, name)
    _dev_mask(science_mask, (h, w), "science_mask")
    _dev_mask(reference_mask, (h, w), "reference_mask")
    sx = np.ascontiguousarray(stamp_x, np.int32).ravel()
    sy = np.ascontiguousarray(stamp_y, np.int32).ravel()
    kn = np.ascontiguousarray(knots, np.float64)
    grid = np.ascontiguousarray(lambda_grid, np.float64).ravel()
    nc = (n_max + 1) * (n_max + 2) // 2
    outs = _fit_outputs((nc + 1) * kn.shape[0], grid.size, sx.size)
    s = stream if stream is not None else torch.cuda.current_stream(science.device)
    _timing_clear()
    _check(_lib.dg_fit_kernel_device(
        w, h, _tp(science), _tp(reference), _tp(science_var), _tp(reference_var),
        _tp(science_mask), _tp(reference_mask), _p(kn), kn.shape[0], _p(sx), _p(sy), sx.size,
        int(stamp_radius), float(beta), int(n_max), int(radius), _p(grid), grid.size,
        float(clip_sigma), int(clip_iterations), int(min_stamps), int(cv_folds),
        _precision(precision), *[_p(o) for o in outs], _vp(s.cuda_stream)))
    return _fit_dict(*outs, grid, nc, kn.shape[0], component_sums(beta, n_max, radius))


def _gls_rows(knots, points, target, weights, bn):
    kn = np.ascontiguousarray(knots, np.float64)
    pts = np.ascontiguousarray(points, np.float64)
    tg = np.ascontiguousarray(target, np.float64).ravel()
    wt = np.ascontiguousarray(weights, np.float64).ravel()
    bn = np.ascontiguousarray(bn, np.float64)
    if bn.ndim != 2 or pts.shape != (tg.size, 2) or wt.size != tg.size or bn.shape[0] != tg.size:
        raise RuntimeError("solve_gls: row counts of design/target/weights/bn disagree")
    return kn, pts, tg, wt, bn


def solve_gls_cv(knots, points, target, weights, bn, lambda_grid, group, n_groups,
                 precision="fp64"):
    """Penalised GLS selecting lambda by k-fold group cross-validation (group =
    fold id per row); gcv_curve holds the CV-error curve (src/bindings.cpp:553-566)."""
    kn, pts, tg, wt, bn = _gls_rows(knots, points, target, weights, bn)
    grp = np.ascontiguousarray(group, np.int32).ravel()
    if grp.size != tg.size:
        raise RuntimeError("solve_gls_cv: group.size() must match points.rows()")
    grid = np.ascontiguousarray(lambda_grid, np.float64).ravel()
    nc = bn.shape[1]
    theta, scal, curve = np.empty((nc + 1) * kn.shape[0]), np.empty(4), np.empty(grid.size)
    _check(_lib.dg_solve_gls_cv(_p(kn), kn.shape[0], _p(pts), _p(tg), _p(wt), _p(bn), tg.size, nc,
                                _p(grid), grid.size, _p(grp), int(n_groups),
                                _precision(precision), _p(theta), _p(scal), _p(curve)))
    return {"theta": theta, "n_components": nc, "n_spatial": kn.shape[0],
            "lambda": float(scal[0]), "gcv": float(scal[1]), "effective_dof": float(scal[2]),
            "rss": float(scal[3]), "lambda_grid": grid, "gcv_curve": curve}


def solve_gls(knots, points, target, weights, bn, lam, precision="fp64"):
    """Penalised GLS at a fixed lambda (src/bindings.cpp:528-537): the GCV
    solve over the one-point grid; as in the reference, ``lambda_grid`` and
    ``gcv_curve`` come back empty (src/bindings.cpp:510-511)."""
    out = solve_gls_gcv(knots, points, target, weights, bn, [float(lam)], precision)
    out["lambda_grid"] = np.empty(0)
    out["gcv_curve"] = np.empty(0)
    return out


def solve_gls_gcv(knots, points, target, weights, bn, lambda_grid, precision="fp64"):
    """Penalised GLS + GCV -> dict (src/bindings.cpp:540-551)."""
    kn = np.ascontiguousarray(knots, np.float64)
    pts = np.ascontiguousarray(points, np.float64)
    tg = np.ascontiguousarray(target, np.float64).ravel()
    wt = np.ascontiguousarray(weights, np.float64).ravel()
    bn = np.ascontiguousarray(bn, np.float64)
    if bn.ndim != 2 or pts.shape != (tg.size, 2) or wt.size != tg.size or bn.shape[0] != tg.size:
        raise RuntimeError("solve_gls: row counts of design/target/weights/bn disagree")
    grid = np.ascontiguousarray(lambda_grid, np.float64).ravel()
    nc = bn.shape[1]
    theta, scal, curve = np.empty((nc + 1) * kn.shape[0]), np.empty(4), np.empty(grid.size)
    _check(_lib.dg_solve_gls_gcv(_p(kn), kn.shape[0], _p(pts), _p(tg), _p(wt), _p(bn), tg.size, nc,
                                 _p(grid), grid.size, _precision(precision), _p(theta), _p(scal),
                                 _p(curve)))
    return {"theta": theta, "n_components": nc, "n_spatial": kn.shape[0], "lambda": scal[0],
            "gcv": scal[1], "effective_dof": scal[2], "rss": scal[3], "lambda_grid": grid,
            "gcv_curve": curve}


# --------------------------------------------------------------------------- noise / score
def _knots_theta(knots, theta):
    kn = np.ascontiguousarray(knots, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    if kn.ndim != 2 or kn.shape[1] != 2:
        raise ValueError("knots must be (k, 2)")
    return kn, th


def decorrelate(difference, knots, theta, beta, n_max, var_science, var_reference, block=256,
                radius=0, kernel_cell_blocks=0):
    """Noise decorrelation -> {'difference', 'variance', 'timing'}.

    Mirrors ``_core.decorrelate`` (src/bindings.cpp:766-792): 'variance' is
    the post-whitening noise level vs + vr sum K^2, Hann-blended.
    """
    diff = _in2d(difference, np.float32, "difference")
    h, w = diff.shape
    vs = _in2d(var_science, np.float32, "var_science")
    vr = _in2d(var_reference, np.float32, "var_reference")
    _match(vs, (h, w), "variance")
    _match(vr, (h, w), "variance")
    kn, th = _knots_theta(knots, theta)
    out = np.empty((h, w), dtype=np.float32)
    var = np.empty((h, w), dtype=np.float32)
    _timing_clear()
    _check(_lib.dg_decorrelate(w, h, _p(diff), _p(kn), kn.shape[0], _p(th), th.size, float(beta),
                               int(n_max), _p(vs), _p(vr), int(block), int(radius),
                               int(kernel_cell_blocks), _p(out), _p(var)))
    return {"difference": out, "variance": var, "timing": drain_timing()}


def decorrelate_device(difference, knots, theta, beta, n_max, var_science, var_reference,
                       block=256, radius=0, kernel_cell_blocks=0, out=None, stream=None):
    """Device-resident decorrelation on torch CUDA float32 tensors."""
    import torch

    h, w = difference.shape
    for t in (difference, var_science, var_reference):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.shape != (h, w) or not t.is_cuda:
            raise ValueError("device images must be contiguous float32 CUDA tensors of one shape")
    out = dict(out or {})
    dev = difference.device
    for key in ("difference", "variance"):
        if out.get(key) is None:
            out[key] = torch.empty((h, w), dtype=torch.float32, device=dev)
    kn, th = _knots_theta(knots, theta)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    _timing_clear()
    _check(_lib.dg_decorrelate_device(
        w, h, _tp(difference), _p(kn), kn.shape[0], _p(th), th.size, float(beta), int(n_max),
        _tp(var_science), _tp(var_reference), int(block), int(radius), int(kernel_cell_blocks),
        _tp(out["difference"]), _tp(out["variance"]), _vp(s.cuda_stream)))
    out["timing"] = drain_timing()
    return out


def whiten_score_psf(psf, knots, theta, beta, n_max, var_science, var_reference, block, radius=0):
    """PSF whitened by the frame-centre decorrelation filter, unit sum.

    Mirrors ``_core.whiten_score_psf`` (src/bindings.cpp:796-874).
    """
    p = _in2d(psf, np.float32, "psf")
    if p.shape[0] != p.shape[1]:
        raise RuntimeError("whiten_score_psf: psf must be square")
    vs = _in2d(var_science, np.float32, "var_science")
    vr = _in2d(var_reference, np.float32, "var_reference")
    if vs.shape != vr.shape:
        raise RuntimeError("whiten_score_psf: variance shapes disagree")
    h, w = vs.shape
    kn, th = _knots_theta(knots, theta)
    out = np.empty(p.shape, dtype=np.float32)
    _check(_lib.dg_whiten_score_psf(_p(p), p.shape[0], _p(kn), kn.shape[0], _p(th), th.size,
                                    float(beta), int(n_max), _p(vs), _p(vr), w, h, int(block),
                                    int(radius), _p(out)))
    return out


def whiten_score_psf_device(psf, knots, theta, beta, n_max, var_science, var_reference, block,
                            radius=0, stream=None):
    """whiten_score_psf with the variance maps as torch CUDA tensors; the PSF
    (host array or CUDA tensor) comes back as a CUDA tensor."""
    import torch

    vs, vr = var_science, var_reference
    if vs.shape != vr.shape:
        raise RuntimeError("whiten_score_psf: variance shapes disagree")
    for t in (vs, vr):
        if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("variance maps must be contiguous float32 CUDA tensors")
    p = torch.as_tensor(np.asarray(psf, dtype=np.float32) if not torch.is_tensor(psf) else psf,
                        dtype=torch.float32, device=vs.device).contiguous()
    if p.ndim != 2 or p.shape[0] != p.shape[1]:
        raise RuntimeError("whiten_score_psf: psf must be square")
    h, w = vs.shape
    kn, th = _knots_theta(knots, theta)
    out = torch.empty_like(p)
    s = stream if stream is not None else torch.cuda.current_stream(vs.device)
    _check(_lib.dg_whiten_score_psf_device(_tp(p), p.shape[0], _p(kn), kn.shape[0], _p(th),
                                           th.size, float(beta), int(n_max), _tp(vs), _tp(vr), w,
                                           h, int(block), int(radius), _tp(out),
                                           _vp(s.cuda_stream)))
    return out


def matched_filter(image, psf, variance):
    """Per-pixel S/N map. Mirrors ``_core.matched_filter`` (src/bindings.cpp:878-898)."""
    img = _in2d(image, np.float32, "image")
    h, w = img.shape
    p = _in2d(psf, np.float32, "psf")
    if p.shape[0] != p.shape[1]:
        raise RuntimeError("matched_filter: psf must be square")
    var = _in2d(variance, np.float32, "variance")
    if var.shape != (h, w):
        raise RuntimeError("matched_filter: variance must match image shape")
    out = np.empty((h, w), dtype=np.float32)
    _timing_clear()
    _check(_lib.dg_matched_filter(w, h, _p(img), _p(p), p.shape[0], _p(var), _p(out)))
    return out


def matched_filter_device(image, psf, variance, out=None, stream=None):
    """Device-resident matched filter on torch CUDA float32 tensors."""
    import torch

    h, w = image.shape
    for t in (image, variance):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.shape != (h, w) or not t.is_cuda:
            raise ValueError("device images must be contiguous float32 CUDA tensors of one shape")
    if psf.dtype != torch.float32 or not psf.is_cuda or psf.shape[0] != psf.shape[1]:
        raise ValueError("psf must be a square float32 CUDA tensor")
    psf = psf.contiguous()
    if out is None:
        out = torch.empty((h, w), dtype=torch.float32, device=image.device)
    s = stream if stream is not None else torch.cuda.current_stream(image.device)
    _timing_clear()
    _check(_lib.dg_matched_filter_device(w, h, _tp(image), _tp(psf), psf.shape[0], _tp(variance),
                                         _tp(out), _vp(s.cuda_stream)))
    return out


# --------------------------------------------------------------------------- basis / TPS
def _ksize(beta, n_max, radius):
    return 2 * (radius if radius > 0 else default_radius(beta, n_max)) + 1


def gauss_hermite_basis1d(beta, n_max, radius=0):
    """1-D sampled Gauss-Hermite functions, (n_max+1, ksize) float64
    (src/bindings.cpp:171-185)."""
    out = np.empty((n_max + 1, _ksize(beta, n_max, radius)), dtype=np.float64)
    _check(_lib.dg_gauss_hermite_basis1d(float(beta), int(n_max), int(radius), _p(out)))
    return out


def gauss_hermite_kernels(beta, n_max, radius=0):
    """(orders int32 (nc, 2), kernels float32 (nc, ks, ks)) of the 2-D
    components (src/bindings.cpp:187-206)."""
    nc, ks = n_components(n_max), _ksize(beta, n_max, radius)
    orders = np.empty((nc, 2), dtype=np.int32)
    kernels = np.empty((nc, ks, ks), dtype=np.float32)
    _check(_lib.dg_gauss_hermite_kernels(float(beta), int(n_max), int(radius), _p(orders),
                                         _p(kernels)))
    return orders, kernels


def basis_convolve(image, beta, n_max, radius=0):
    """``image`` convolved with every basis component -> (nc, H, W) float32
    (src/bindings.cpp:209-227), computed on the GPU."""
    img = _in2d(image, np.float32, "image")
    h, w = img.shape
    out = np.empty((n_components(n_max), h, w), dtype=np.float32)
    _check(_lib.dg_basis_convolve(w, h, _p(img), float(beta), int(n_max), int(radius), _p(out)))
    return out


def basis_convolve_device(image, beta, n_max, radius=0, out=None, stream=None):
    """basis_convolve of a float32 CUDA tensor into an (nc, H, W) CUDA tensor."""
    import torch

    if image.dtype != torch.float32 or not image.is_cuda or image.dim() != 2:
        raise ValueError("image must be a 2-D float32 CUDA tensor")
    image = image.contiguous()
    h, w = image.shape
    shape = (n_components(n_max), h, w)
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=image.device)
    elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {shape}")
    s = stream if stream is not None else torch.cuda.current_stream(image.device)
    _check(_lib.dg_basis_convolve_device(w, h, _tp(image), float(beta), int(n_max), int(radius),
                                         _tp(out), _vp(s.cuda_stream)))
    return out


def grid_knots(x0, y0, x1, y1, nx, ny):
    """Regular nx*ny knot grid over [x0,x1]x[y0,y1], shape (nx*ny, 2): x
    varies fastest, a single knot along an axis sits at the midpoint
    (src/spatial.cpp:113-129)."""
    if nx < 1 or ny < 1:
        raise ValueError("nx, ny must be >= 1")
    fx = np.full(1, 0.5) if nx == 1 else np.arange(nx) / (nx - 1)
    fy = np.full(1, 0.5) if ny == 1 else np.arange(ny) / (ny - 1)
    xs = x0 + fx * (x1 - x0)
    ys = y0 + fy * (y1 - y0)
    gx, gy = np.meshgrid(xs, ys)
    return np.column_stack([gx.ravel(), gy.ravel()]).astype(np.float64)


def _knots_points(knots, points):
    kn = np.ascontiguousarray(knots, np.float64)
    pts = np.ascontiguousarray(points, np.float64)
    if kn.ndim != 2 or kn.shape[1] != 2 or pts.ndim != 2 or pts.shape[1] != 2:
        raise ValueError("knots and points must be (n, 2) arrays")
    return kn, pts


def tps_design(knots, points):
    """Thin-plate design matrix (m, n_knots) (src/bindings.cpp:487-490)."""
    kn, pts = _knots_points(knots, points)
    out = np.empty((pts.shape[0], kn.shape[0]), dtype=np.float64)
    _check(_lib.dg_tps_design(_p(kn), kn.shape[0], _p(pts), pts.shape[0], _p(out)))
    return out


def tps_penalty(knots):
    """Bending-energy penalty (n_knots, n_knots) (src/bindings.cpp:492-494)."""
    kn = np.ascontiguousarray(knots, np.float64)
    out = np.empty((kn.shape[0], kn.shape[0]), dtype=np.float64)
    _check(_lib.dg_tps_penalty(_p(kn), kn.shape[0], _p(out)))
    return out


def tps_fit(knots, points, values, lam):
    """Unweighted penalised fit (D^T D + lam P) theta = D^T values
    (src/bindings.cpp:496-500)."""
    kn, pts = _knots_points(knots, points)
    v = np.ascontiguousarray(values, np.float64).ravel()
    if v.size != pts.shape[0]:
        raise RuntimeError("tps_fit: values must have one entry per point")
    theta = np.empty(kn.shape[0], dtype=np.float64)
    _check(_lib.dg_tps_fit(_p(kn), kn.shape[0], _p(pts), _p(v), v.size, float(lam), _p(theta)))
    return theta


def tps_evaluate(knots, points, coeffs):
    """design(points) @ coeffs, shape (m,) (src/bindings.cpp:502-506)."""
    kn, pts = _knots_points(knots, points)
    c = np.ascontiguousarray(coeffs, np.float64).ravel()
    out = np.empty(pts.shape[0], dtype=np.float64)
    _check(_lib.dg_tps_evaluate(_p(kn), kn.shape[0], _p(pts), pts.shape[0], _p(c), c.size,
                                _p(out)))
    return out


def decorrelate_block(image, kernel, var_science, var_reference):
    """One square block whitened with a constant kernel and noise levels
    (src/bindings.cpp:725-745)."""
    img = _in2d(image, np.float32, "image")
    if img.shape[0] != img.shape[1]:
        raise RuntimeError("decorrelate_block: image must be square")
    k = _in2d(kernel, np.float32, "kernel")
    if k.shape[0] != k.shape[1]:
        raise RuntimeError("decorrelate_block: kernel must be square")
    n = img.shape[0]
    out = np.empty((n, n), dtype=np.float32)
    _check(_lib.dg_decorrelate_block(n, _p(img), _p(k), k.shape[0], float(var_science),
                                     float(var_reference), _p(out)))
    return out


def decorrelation_kernel(kernel, var_science, var_reference, n):
    """Real-space decorrelation kernel, centred, (n, n) (src/bindings.cpp:748-763)."""
    k = _in2d(kernel, np.float32, "kernel")
    out = np.empty((n, n), dtype=np.float32)
    _check(_lib.dg_decorrelation_kernel(_p(k), k.shape[0], float(var_science),
                                        float(var_reference), int(n), _p(out)))
    return out


# --------------------------------------------------------------------------- caller arithmetic
def synth_variance(data, gain, read_noise=0.0):
    """Var = max(data, 0) / gain + read_noise^2, float32 (python/delta/_inputs.py:75-86)."""
    if gain <= 0.0:
        raise ValueError("gain must be > 0")
    d = _in2d(data, np.float32, "data")
    h, w = d.shape
    out = np.empty((h, w), dtype=np.float32)
    _check(_lib.dg_synth_variance(w, h, _p(d), float(gain), float(read_noise), _p(out)))
    return out


def synth_variance_device(data, gain, read_noise=0.0, out=None, stream=None):
    """synth_variance of a float32 CUDA tensor."""
    import torch

    if gain <= 0.0:
        raise ValueError("gain must be > 0")
    if data.dtype != torch.float32 or not data.is_cuda or data.dim() != 2:
        raise ValueError("data must be a 2-D float32 CUDA tensor")
    data = data.contiguous()
    h, w = data.shape
    if out is None:
        out = torch.empty_like(data)
    s = stream if stream is not None else torch.cuda.current_stream(data.device)
    _check(_lib.dg_synth_variance_device(w, h, _tp(data), float(gain), float(read_noise),
                                         _tp(out), _vp(s.cuda_stream)))
    return out


def variance_rescale_factor(difference, variance, mask=None):
    """median(d^2 / v) / median(chi2_1) over good pixels, 1.0 if none
    (python/delta/pipeline.py:66-86)."""
    d = _in2d(difference, np.float32, "difference")
    h, w = d.shape
    v = _in2d(variance, np.float32, "variance")
    m = _in2d(mask, np.uint8, "mask")
    if v.shape != (h, w) or (m is not None and m.shape != (h, w)):
        raise ValueError("difference, variance and mask must share a shape")
    f = C.c_double(1.0)
    _check(_lib.dg_variance_rescale_factor(w, h, _p(d), _p(v), _p(m), C.byref(f)))
    return float(f.value)


def variance_rescale_factor_device(difference, variance, mask=None, stream=None):
    """variance_rescale_factor of CUDA tensors (float32, float32, uint8)."""
    import torch

    h, w = difference.shape
    for t in (difference, variance):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.shape != (h, w) or not t.is_cuda:
            raise ValueError("device images must be contiguous float32 CUDA tensors of one shape")
    if mask is not None and (mask.dtype != torch.uint8 or tuple(mask.shape) != (h, w)
                             or not mask.is_contiguous()):
        raise ValueError("mask must be a contiguous uint8 CUDA tensor of the image shape")
    s = stream if stream is not None else torch.cuda.current_stream(difference.device)
    f = C.c_double(1.0)
    _check(_lib.dg_variance_rescale_factor_device(w, h, _tp(difference), _tp(variance),
                                                  _tp(mask) if mask is not None else None,
                                                  C.byref(f), _vp(s.cuda_stream)))
    return float(f.value)


def _psf_positions(xs, ys):
    x = np.ascontiguousarray(xs, dtype=np.int32).ravel()
    y = np.ascontiguousarray(ys, dtype=np.int32).ravel()
    if x.size != y.size:
        raise ValueError("xs and ys must have the same length")
    return x, y


def estimate_psf(image, xs, ys, radius):
    """Median unit-sum stamp of the sources, float32 (2r+1, 2r+1)
    (python/delta/pipeline.py:285-302), computed on the GPU."""
    img = _in2d(image, np.float32, "image")
    h, w = img.shape
    x, y = _psf_positions(xs, ys)
    ks = 2 * int(radius) + 1
    out = np.empty((ks, ks), dtype=np.float32)
    _check(_lib.dg_estimate_psf(w, h, _p(img), _p(x), _p(y), x.size, int(radius), _p(out)))
    return out


def estimate_psf_device(image, xs, ys, radius, stream=None):
    """estimate_psf of a float32 CUDA image; positions on the host, the PSF
    comes back as a host array."""
    import torch

    if image.dtype != torch.float32 or not image.is_cuda or image.dim() != 2:
        raise ValueError("image must be a 2-D float32 CUDA tensor")
    image = image.contiguous()
    h, w = image.shape
    x, y = _psf_positions(xs, ys)
    ks = 2 * int(radius) + 1
    out = np.empty((ks, ks), dtype=np.float32)
    s = stream if stream is not None else torch.cuda.current_stream(image.device)
    _check(_lib.dg_estimate_psf_device(w, h, _tp(image), _p(x), _p(y), x.size, int(radius),
                                       _p(out), _vp(s.cuda_stream)))
    return out


def weighted_mean(data, variance, mask):
    """Inverse-variance weighted mean over unmasked pixels with positive
    variance (src/bindings.cpp:107-125), reduced on the GPU."""
    d = _in2d(data, np.float32, "data")
    h, w = d.shape
    v = _in2d(variance, np.float32, "variance")
    m = _in2d(mask, np.uint8, "mask")
    if v.shape != d.shape or m.shape != d.shape:
        raise RuntimeError("weighted_mean: data, variance and mask must share a shape")
    out = C.c_double(0.0)
    _check(_lib.dg_weighted_mean(w, h, _p(d), _p(v), _p(m), C.byref(out)))
    return float(out.value)


# --------------------------------------------------------------------------- candidates
def label_components(score, threshold, polarity=1, mask=None):
    """(labels int32 (H, W), count) of the seed pixels of ``score``.

    Seeds are seed_mask (src/catalog.cpp:61-80): finite, score >= threshold
    (polarity > 0) or <= -threshold (polarity < 0), unmasked where a mask is
    given; labels are 8-connected, numbered in raster order of each
    component's first pixel (label_components, src/catalog.cpp:105-261).
    """
    sc = _in2d(score, np.float32, "score")
    h, w = sc.shape
    m = _in2d(mask, np.uint8, "mask")
    _match(m, (h, w), "mask")
    labels = np.empty((h, w), dtype=np.int32)
    n = _i(0)
    _check(_lib.dg_label(w, h, _p(sc), _p(m), float(threshold), int(polarity), _p(labels),
                         C.byref(n)))
    return labels, int(n.value)


def label_components_device(score, threshold, polarity=1, mask=None, out=None, stream=None):
    """Device-resident labelling on a torch CUDA float32 score (uint8 mask)."""
    import torch

    h, w = score.shape
    if score.dtype != torch.float32 or not score.is_contiguous() or not score.is_cuda:
        raise ValueError("score must be a contiguous float32 CUDA tensor")
    if mask is not None and (mask.dtype != torch.uint8 or not mask.is_contiguous()
                             or mask.shape != (h, w)):
        raise ValueError("mask must be a contiguous uint8 CUDA tensor of the score's shape")
    if out is None:
        out = torch.empty((h, w), dtype=torch.int32, device=score.device)
    n = _i(0)
    s = stream if stream is not None else torch.cuda.current_stream(score.device)
    _check(_lib.dg_label_device(w, h, _tp(score), _tp(mask), float(threshold), int(polarity),
                                _tp(out), C.byref(n), _vp(s.cuda_stream)))
    return out, int(n.value)



# --------------------------------------------------------------------------- photometric scale
def _scale_args(knots, theta, component_sums):
    kn, th = _knots_theta(knots, theta)
    sums = np.ascontiguousarray(component_sums, dtype=np.float64).ravel()
    return kn, th, sums


def photometric_scale(knots, theta, component_sums, height, width):
    """Per-pixel photometric scale sum_n a_n(x, y) S_n, float32 (height, width)
    (src/bindings.cpp:695-707, src/fit.cpp:513-530)."""
    kn, th, sums = _scale_args(knots, theta, component_sums)
    out = np.empty((int(height), int(width)), dtype=np.float32)
    _check(_lib.dg_photometric_scale(_p(kn), kn.shape[0], _p(th), th.size, _p(sums), sums.size,
                                     int(width), int(height), _p(out)))
    return out


def photometric_scale_device(knots, theta, component_sums, height, width, device=None,
                             out=None, stream=None):
    """Device form of :func:`photometric_scale`: a float32 CUDA tensor."""
    import torch

    kn, th, sums = _scale_args(knots, theta, component_sums)
    if out is None:
        out = torch.empty((int(height), int(width)), dtype=torch.float32,
                          device=device or torch.device("cuda:0"))
    s = stream if stream is not None else torch.cuda.current_stream(out.device)
    _check(_lib.dg_photometric_scale_device(_p(kn), kn.shape[0], _p(th), th.size, _p(sums),
                                            sums.size, int(width), int(height), _tp(out),
                                            _vp(s.cuda_stream)))
    return out


def photometric_scale_at(knots, theta, component_sums, points):
    """Photometric scale at (m, 2) pixel coordinates, float64 (m,)
    (src/bindings.cpp:710-719, src/fit.cpp:532-545)."""
    kn, th, sums = _scale_args(knots, theta, component_sums)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    out = np.empty(pts.shape[0], dtype=np.float64)
    _check(_lib.dg_photometric_scale_at(_p(kn), kn.shape[0], _p(th), th.size, _p(sums),
                                        sums.size, _p(pts), pts.shape[0], _p(out)))
    return out


# --------------------------------------------------------------------------- catalog
_CATALOG_COLUMNS = ("x", "y", "peak_x", "peak_y", "peak_snr", "n_pix", "flux", "expected_n_pix",
                    "fwhm_ratio", "fwhm_consistent", "mask_flags", "is_dipole", "quality",
                    "psf_chi2", "psf_chi2_dof", "psf_amplitude", "psf_amplitude_err",
                    "shape_consistent", "bright_star_dist", "near_bright_star")


def _catalog_params(threshold_sigma, threshold_sigma_dipole, expected_fwhm, fwhm_tolerance_lo,
                    fwhm_tolerance_hi, aperture_radius, exclude_bad_pixels, fit_psf_shape,
                    shape_chi2_tolerance, bright_star_radius, max_shape_fit):
    """make_catalog_params (src/bindings.cpp:330-345)."""
    return CatalogParams(float(threshold_sigma), float(threshold_sigma_dipole),
                         float(expected_fwhm), float(fwhm_tolerance_lo), float(fwhm_tolerance_hi),
                         int(aperture_radius), int(bool(exclude_bad_pixels)),
                         int(bool(fit_psf_shape)), float(shape_chi2_tolerance),
                         float(bright_star_radius), int(max_shape_fit))


def _catalog_dict(rows):
    return {k: np.ascontiguousarray(rows[k]) for k in _CATALOG_COLUMNS}


def build_catalog(score, difference, mask=None, threshold_sigma=5.0, threshold_sigma_dipole=3.0,
                  expected_fwhm=3.5, fwhm_tolerance_lo=0.3, fwhm_tolerance_hi=3.0,
                  aperture_radius=0, exclude_bad_pixels=True, return_negative=False,
                  fit_psf_shape=False, shape_chi2_tolerance=3.0, variance=None, psf=None,
                  bright_star_radius=0.0, max_shape_fit=5000, bright_x=None, bright_y=None):
    """Connected-component source catalog -> column dict (src/bindings.cpp:348-483,
    src/catalog.cpp:600-697).  ``timing`` as for every call (None unless
    DELTA_TIMING is set)."""
    sc = _in2d(score, np.float32, "score")
    df = _in2d(difference, np.float32, "difference")
    if df.shape != sc.shape:
        raise RuntimeError("build_catalog: difference shape must match score")
    va = _in2d(variance, np.float32, "variance")
    if va is not None and va.shape != sc.shape:
        raise RuntimeError("build_catalog: variance shape must match score")
    ps = _in2d(psf, np.float32, "psf")
    if ps is not None and ps.shape[0] != ps.shape[1]:
        raise RuntimeError("build_catalog: psf must be square")
    if (bright_x is None) != (bright_y is None):
        raise RuntimeError("build_catalog: supply both bright_x and bright_y, or neither")
    bx = by = None
    if bright_x is not None:
        bx = np.ascontiguousarray(bright_x, dtype=np.float64).ravel()
        by = np.ascontiguousarray(bright_y, dtype=np.float64).ravel()
        if bx.size != by.size:
            raise RuntimeError("build_catalog: bright_x and bright_y must be the same length")
    m = _in2d(mask, np.uint8, "mask")
    _match(m, sc.shape, "mask")
    h, w = sc.shape
    p = _catalog_params(threshold_sigma, threshold_sigma_dipole, expected_fwhm,
                        fwhm_tolerance_lo, fwhm_tolerance_hi, aperture_radius,
                        exclude_bad_pixels, fit_psf_shape, shape_chi2_tolerance,
                        bright_star_radius, max_shape_fit)
    cap = 4096
    _timing_clear()
    while True:
        rows = np.zeros(cap, CATALOG_ENTRY)
        n = _i(0)
        _check(_lib.dg_build_catalog(w, h, _p(sc), _p(df), _p(m), C.byref(p),
                                     1 if return_negative else 0, _p(va), _p(ps),
                                     0 if ps is None else ps.shape[0], _p(bx), _p(by),
                                     0 if bx is None else bx.size, _p(rows), cap, C.byref(n)))
        if n.value <= cap:
            out = _catalog_dict(rows[:n.value])
            out["timing"] = drain_timing()
            return out
        cap = n.value


def build_catalog_device(score, difference, mask=None, threshold_sigma=5.0,
                         threshold_sigma_dipole=3.0, expected_fwhm=3.5, fwhm_tolerance_lo=0.3,
                         fwhm_tolerance_hi=3.0, aperture_radius=0, exclude_bad_pixels=True,
                         return_negative=False, fit_psf_shape=False, shape_chi2_tolerance=3.0,
                         variance=None, psf=None, bright_star_radius=0.0, max_shape_fit=5000,
                         bright_x=None, bright_y=None, capacity=4096, stream=None,
                         labels=None, n_labels=None):
    """Device form of :func:`build_catalog` on torch CUDA tensors (psf and
    bright_x/bright_y too).  Returns (rows, n): a uint8 CUDA tensor holding
    ``min(n, capacity)`` rows of :data:`CATALOG_ENTRY` and the full count n
    (call again with ``capacity >= n`` for every row).  ``labels`` /
    ``n_labels``: the caller's labelling of the same seeds (as
    :func:`label_components_device` returns it for this threshold, polarity
    and seed mask), used instead of labelling again."""
    import torch

    _dev_any_image(score, "score")
    _dev_image(difference, score.shape[0], score.shape[1], "difference")
    if variance is not None:
        _dev_image(variance, score.shape[0], score.shape[1], "variance")
    _dev_mask(mask, score.shape, "mask")
    if psf is not None:
        if psf.dim() != 2 or psf.shape[0] != psf.shape[1]:
            raise RuntimeError("build_catalog: psf must be square")
        if psf.dtype != torch.float32 or not psf.is_contiguous() or not psf.is_cuda:
            raise ValueError("build_catalog: psf must be a contiguous float32 CUDA tensor")
    if (bright_x is None) != (bright_y is None):
        raise RuntimeError("build_catalog: supply both bright_x and bright_y, or neither")
    nb = 0
    if bright_x is not None:
        if bright_x.numel() != bright_y.numel():
            raise RuntimeError("build_catalog: bright_x and bright_y must be the same length")
        for t, name in ((bright_x, "bright_x"), (bright_y, "bright_y")):
            if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError(f"build_catalog: {name} must be a contiguous float64 CUDA tensor")
        nb = bright_x.numel()
    h, w = score.shape
    p = _catalog_params(threshold_sigma, threshold_sigma_dipole, expected_fwhm,
                        fwhm_tolerance_lo, fwhm_tolerance_hi, aperture_radius,
                        exclude_bad_pixels, fit_psf_shape, shape_chi2_tolerance,
                        bright_star_radius, max_shape_fit)
    rows = torch.empty((max(capacity, 1), CATALOG_ENTRY.itemsize), dtype=torch.uint8,
                       device=score.device)
    n = _i(0)
    s = stream if stream is not None else torch.cuda.current_stream(score.device)
    if labels is not None:
        if (labels.dtype != torch.int32 or not labels.is_contiguous() or not labels.is_cuda or
                tuple(labels.shape) != (h, w)):
            raise ValueError("build_catalog: labels must be a contiguous int32 CUDA tensor of "
                             "the score's shape")
        if n_labels is None or int(n_labels) < 0:
            raise ValueError("build_catalog: pass the label count with labels")
        _check(_lib.dg_build_catalog_labelled_device(
            w, h, _tp(score), _tp(difference), _tp(mask), C.byref(p),
            1 if return_negative else 0, _tp(variance), _tp(psf),
            0 if psf is None else psf.shape[0], _tp(bright_x), _tp(bright_y), nb, _tp(labels),
            int(n_labels), _tp(rows), capacity, C.byref(n), _vp(s.cuda_stream)))
        return rows, n.value
    _check(_lib.dg_build_catalog_device(
        w, h, _tp(score), _tp(difference), _tp(mask), C.byref(p), 1 if return_negative else 0,
        _tp(variance), _tp(psf), 0 if psf is None else psf.shape[0], _tp(bright_x),
        _tp(bright_y), nb, _tp(rows), capacity, C.byref(n), _vp(s.cuda_stream)))
    return rows, n.value


def catalog_rows_to_dict(rows, n):
    """Column dict of the first n rows of a device catalog buffer."""
    host = rows[:n].cpu().numpy().reshape(-1).view(CATALOG_ENTRY)
    return _catalog_dict(host)


# --------------------------------------------------------------------------- stamp selection
def _detect_params(stamp_radius, threshold_sigma, max_stamps, saturation, isolation_radius,
                   border, fwhm_radius, match_radius):
    """make_params (src/bindings.cpp:70-84)."""
    return DetectParams(int(stamp_radius), float(threshold_sigma), int(max_stamps),
                        float(saturation), int(isolation_radius), int(border), int(fwhm_radius),
                        int(match_radius))


def estimate_background(data, mask=None):
    """Robust (median, MAD-sigma) background over unmasked pixels
    (src/bindings.cpp:231-236, src/detect.cpp:84-113)."""
    d = _in2d(data, np.float32, "data")
    h, w = d.shape
    m = _in2d(mask, np.uint8, "mask")
    _match(m, (h, w), "mask")
    med, sig = _d(0.0), _d(0.0)
    if d.size == 0:
        return 0.0, 0.0
    _check(_lib.dg_estimate_background(w, h, _p(d), _p(m), C.byref(med), C.byref(sig)))
    return med.value, sig.value


def detect_stamps(data, mask=None, stamp_radius=15, threshold_sigma=5.0, max_stamps=200,
                  saturation=0.0, isolation_radius=0, border=0, fwhm_radius=0, match_radius=2):
    """Detect bright, isolated, unsaturated point-source stamps ->
    {'x','y','flux','snr','fwhm'} (src/bindings.cpp:239-273, src/detect.cpp:159-241)."""
    d = _in2d(data, np.float32, "data")
    h, w = d.shape
    m = _in2d(mask, np.uint8, "mask")
    _match(m, (h, w), "mask")
    p = _detect_params(stamp_radius, threshold_sigma, max_stamps, saturation, isolation_radius,
                       border, fwhm_radius, match_radius)
    cap = max(p.max_stamps, 0)
    cols = {"x": np.empty(cap, np.int32), "y": np.empty(cap, np.int32),
            "flux": np.empty(cap), "snr": np.empty(cap), "fwhm": np.empty(cap)}
    n = _i(0)
    if d.size and cap:
        _check(_lib.dg_detect_stamps(w, h, _p(d), _p(m), C.byref(p),
                                     *[_p(cols[k]) for k in ("x", "y", "flux", "snr", "fwhm")],
                                     cap, C.byref(n)))
    return {k: v[:n.value].copy() for k, v in cols.items()}


def _catalog(catalog_x, catalog_y):
    if catalog_x is None or catalog_y is None:
        return None, None, 0
    cx = np.ascontiguousarray(catalog_x, dtype=np.int32)
    cy = np.ascontiguousarray(catalog_y, dtype=np.int32)
    if cx.ndim != 1 or cy.ndim != 1:
        raise ValueError("catalog_x and catalog_y must be 1-D")
    return cx, cy, min(cx.size, cy.size)


def _selection_dict(cols, n, med, direction):
    out = {k: v[:n].copy() for k, v in cols.items()}
    out["median_fwhm_science"] = float(med[0])
    out["median_fwhm_reference"] = float(med[1])
    out["direction"] = "reference" if direction == 0 else "science"
    return out


def select_stamps(science, reference, science_mask=None, reference_mask=None, catalog_x=None,
                  catalog_y=None, stamp_radius=15, threshold_sigma=5.0, max_stamps=200,
                  saturation=0.0, isolation_radius=0, border=0, fwhm_radius=0, match_radius=2):
    """Select matched stamps across both images (detection or a supplied
    catalog) and the convolution direction (src/bindings.cpp:277-320,
    src/detect.cpp:243-314)."""
    sci = _in2d(science, np.float32, "science")
    ref = _in2d(reference, np.float32, "reference")
    sm = _in2d(science_mask, np.uint8, "science_mask")
    rm = _in2d(reference_mask, np.uint8, "reference_mask")
    _match(sm, sci.shape, "mask")
    _match(rm, ref.shape, "mask")
    p = _detect_params(stamp_radius, threshold_sigma, max_stamps, saturation, isolation_radius,
                       border, fwhm_radius, match_radius)
    cx, cy, ncat = _catalog(catalog_x, catalog_y)
    cap = ncat if cx is not None else max(p.max_stamps, 0)
    cols = {"x": np.empty(cap + 1, np.int32), "y": np.empty(cap + 1, np.int32),
            "fwhm_science": np.empty(cap + 1), "fwhm_reference": np.empty(cap + 1)}
    n, direction = _i(0), _i(1)
    med = np.full(2, np.nan)
    if sci.size == 0 or ref.size == 0:  # nothing on-frame: empty selection
        return _selection_dict(cols, 0, med, 1)
    _check(_lib.dg_select_stamps(sci.shape[1], sci.shape[0], _p(sci), _p(sm), ref.shape[1],
                                 ref.shape[0], _p(ref), _p(rm), _p(cx), _p(cy), ncat, C.byref(p),
                                 *[_p(cols[k]) for k in ("x", "y", "fwhm_science",
                                                         "fwhm_reference")],
                                 cap, C.byref(n), _p(med), C.byref(direction)))
    return _selection_dict(cols, n.value, med, direction.value)


def _dev_any_image(t, name):
    import torch

    if t.dtype != torch.float32 or not t.is_contiguous() or t.dim() != 2 or not t.is_cuda:
        raise ValueError(f"{name} must be a contiguous 2-D float32 CUDA tensor")


def _dev_mask(m, shape, name):
    import torch

    if m is not None and (m.dtype != torch.uint8 or not m.is_contiguous() or not m.is_cuda
                          or tuple(m.shape) != tuple(shape)):
        raise ValueError(f"{name} must be a contiguous uint8 CUDA tensor of its image's shape")


def estimate_background_device(data, mask=None, stream=None):
    """Device form of :func:`estimate_background` (torch CUDA tensors)."""
    import torch

    _dev_any_image(data, "data")
    _dev_mask(mask, data.shape, "mask")
    h, w = data.shape
    med, sig = _d(0.0), _d(0.0)
    s = stream if stream is not None else torch.cuda.current_stream(data.device)
    _check(_lib.dg_estimate_background_device(w, h, _tp(data), _tp(mask), C.byref(med),
                                              C.byref(sig), _vp(s.cuda_stream)))
    return med.value, sig.value


def detect_stamps_device(data, mask=None, stamp_radius=15, threshold_sigma=5.0, max_stamps=200,
                         saturation=0.0, isolation_radius=0, border=0, fwhm_radius=0,
                         match_radius=2, stream=None):
    """Device form of :func:`detect_stamps`: columns are CUDA tensors."""
    import torch

    _dev_any_image(data, "data")
    _dev_mask(mask, data.shape, "mask")
    h, w = data.shape
    p = _detect_params(stamp_radius, threshold_sigma, max_stamps, saturation, isolation_radius,
                       border, fwhm_radius, match_radius)
    cap = max(p.max_stamps, 0)
    dev = data.device
    cols = {"x": torch.empty(cap + 1, dtype=torch.int32, device=dev),
            "y": torch.empty(cap + 1, dtype=torch.int32, device=dev),
            "flux": torch.empty(cap + 1, dtype=torch.float64, device=dev),
            "snr": torch.empty(cap + 1, dtype=torch.float64, device=dev),
            "fwhm": torch.empty(cap + 1, dtype=torch.float64, device=dev)}
    n = _i(0)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    _check(_lib.dg_detect_stamps_device(w, h, _tp(data), _tp(mask), C.byref(p),
                                        *[_tp(cols[k]) for k in ("x", "y", "flux", "snr",
                                                                 "fwhm")],
                                        cap, C.byref(n), _vp(s.cuda_stream)))
    return {k: v[:n.value] for k, v in cols.items()}


def select_stamps_device(science, reference, science_mask=None, reference_mask=None,
                         catalog_x=None, catalog_y=None, stamp_radius=15, threshold_sigma=5.0,
                         max_stamps=200, saturation=0.0, isolation_radius=0, border=0,
                         fwhm_radius=0, match_radius=2, stream=None):
    """Device form of :func:`select_stamps`: images (and a catalog, int32)
    are CUDA tensors; x, y and the FWHM columns come back as CUDA tensors."""
    import torch

    _dev_any_image(science, "science")
    _dev_any_image(reference, "reference")
    _dev_mask(science_mask, science.shape, "science_mask")
    _dev_mask(reference_mask, reference.shape, "reference_mask")
    p = _detect_params(stamp_radius, threshold_sigma, max_stamps, saturation, isolation_radius,
                       border, fwhm_radius, match_radius)
    cat = catalog_x is not None and catalog_y is not None
    ncat = 0
    if cat:
        for c in (catalog_x, catalog_y):
            if c.dtype != torch.int32 or not c.is_cuda or not c.is_contiguous() or c.dim() != 1:
                raise ValueError("catalog_x/catalog_y must be contiguous 1-D int32 CUDA tensors")
        ncat = min(catalog_x.numel(), catalog_y.numel())
    cap = ncat if cat else max(p.max_stamps, 0)
    dev = science.device
    cols = {"x": torch.empty(cap + 1, dtype=torch.int32, device=dev),
            "y": torch.empty(cap + 1, dtype=torch.int32, device=dev),
            "fwhm_science": torch.empty(cap + 1, dtype=torch.float64, device=dev),
            "fwhm_reference": torch.empty(cap + 1, dtype=torch.float64, device=dev)}
    n, direction = _i(0), _i(0)
    med = np.zeros(2)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    _check(_lib.dg_select_stamps_device(
        science.shape[1], science.shape[0], _tp(science), _tp(science_mask),
        reference.shape[1], reference.shape[0], _tp(reference), _tp(reference_mask),
        _tp(catalog_x) if cat else None, _tp(catalog_y) if cat else None, ncat, C.byref(p),
        *[_tp(cols[k]) for k in ("x", "y", "fwhm_science", "fwhm_reference")], cap,
        C.byref(n), _p(med), C.byref(direction), _vp(s.cuda_stream)))
    out = {k: v[:n.value] for k, v in cols.items()}
    out["median_fwhm_science"] = float(med[0])
    out["median_fwhm_reference"] = float(med[1])
    out["direction"] = "reference" if direction.value == 0 else "science"
    return out


# --------------------------------------------------------------------------- constant-kernel model
def n_components(n_max: int) -> int:
    return (int(n_max) + 1) * (int(n_max) + 2) // 2


def kernel_chi2(science, reference, weight, params, n_max, radius, grad=True):
    """chi^2 = sum w (S - b - K (x) R)^2 of a spatially constant Gauss-Hermite
    kernel with free scale and sub-pixel shift, and its gradient.

    params = [a_0 .. a_{nc-1}, b, beta, dx, dy]; returns {'chi2', 'grad'}
    (grad None when not requested).  Host arrays; the work runs on the GPU.
    No counterpart in the reference (whose basis is fixed at dx = dy = 0,
    src/basis.cpp:14-59, and which has no gradient code).
    """
    sci = _in2d(science, np.float32, "science")
    ref = _in2d(reference, np.float32, "reference")
    h, w = sci.shape
    _match(ref, (h, w), "reference")
    wt = _in2d(weight, np.float32, "weight")
    _match(wt, (h, w), "weight")
    nc = n_components(n_max)
    p = np.ascontiguousarray(params, np.float64).ravel()
    if p.size != nc + 4:
        raise ValueError(f"params must have n_components + 4 = {nc + 4} entries")
    out = np.zeros(nc + 5)
    _check(_lib.dg_kernel_chi2(w, h, _p(sci), _p(ref), _p(wt), _p(p), int(n_max), int(radius),
                               1 if grad else 0, _p(out)))
    return {"chi2": float(out[0]), "grad": out[1:].copy() if grad else None}


def _dev_image(t, h, w, name):
    import torch

    if t.dtype != torch.float32 or not t.is_contiguous() or t.shape != (h, w) or not t.is_cuda:
        raise ValueError(f"{name} must be a contiguous float32 CUDA tensor of the frame's shape")


def _dev_params(params, nc):
    import torch

    if (params.dtype != torch.float64 or not params.is_cuda or not params.is_contiguous()
            or params.numel() != nc + 4):
        raise ValueError(f"params must be a contiguous float64 CUDA tensor of {nc + 4} entries")


def kernel_chi2_device(science, reference, weight, params, n_max, radius, grad=True, out=None,
                       stream=None):
    """Device form of :func:`kernel_chi2` on torch CUDA tensors.  ``params``
    is a float64 CUDA tensor; returns a float64 CUDA tensor
    [chi2, grad...] (nc + 5 entries; only entry 0 valid when grad=False)."""
    import torch

    h, w = science.shape
    _dev_image(science, h, w, "science")
    _dev_image(reference, h, w, "reference")
    if weight is not None:
        _dev_image(weight, h, w, "weight")
    nc = n_components(n_max)
    _dev_params(params, nc)
    if out is None:
        out = torch.zeros(nc + 5, dtype=torch.float64, device=science.device)
    s = stream if stream is not None else torch.cuda.current_stream(science.device)
    _check(_lib.dg_kernel_chi2_device(w, h, _tp(science), _tp(reference), _tp(weight),
                                      _tp(params), int(n_max), int(radius), 1 if grad else 0,
                                      _tp(out), _vp(s.cuda_stream)))
    return out


def kernel_difference_device(science, reference, params, n_max, radius, out=None, stream=None):
    """D = S - b - K (x) R of the constant-kernel model (torch CUDA tensors)."""
    import torch

    h, w = science.shape
    _dev_image(science, h, w, "science")
    _dev_image(reference, h, w, "reference")
    _dev_params(params, n_components(n_max))
    if out is None:
        out = torch.empty((h, w), dtype=torch.float32, device=science.device)
    s = stream if stream is not None else torch.cuda.current_stream(science.device)
    _check(_lib.dg_kernel_difference_device(w, h, _tp(science), _tp(reference), _tp(params),
                                            int(n_max), int(radius), _tp(out),
                                            _vp(s.cuda_stream)))
    return out


def kernel_difference_vjp_device(reference, grad_difference, params, n_max, radius, out=None,
                                 stream=None):
    """dL/dparams given dL/dD for the constant-kernel model (torch CUDA
    tensors); returns [0, dL/dparams...] (nc + 5 float64)."""
    import torch

    h, w = reference.shape
    _dev_image(reference, h, w, "reference")
    _dev_image(grad_difference, h, w, "grad_difference")
    nc = n_components(n_max)
    _dev_params(params, nc)
    if out is None:
        out = torch.zeros(nc + 5, dtype=torch.float64, device=reference.device)
    s = stream if stream is not None else torch.cuda.current_stream(reference.device)
    _check(_lib.dg_kernel_difference_vjp_device(w, h, _tp(reference), _tp(grad_difference),
                                                _tp(params), int(n_max), int(radius), _tp(out),
                                                _vp(s.cuda_stream)))
    return out
