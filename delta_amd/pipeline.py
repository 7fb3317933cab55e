"""Device-resident counterpart of the reference's ``Subtractor._fit/_subtract``.

The reference's Python layer (python/delta/pipeline.py:430-806) orchestrates
the ``_core`` calls; this module makes the same calls, in the same order and
with the same choices, on torch CUDA tensors that stay in HBM:

  * direction: the reference image is convolved to match the science image
    (the ``target`` is the science frame, pipeline.py:477-487);
  * beta from the two median FWHMs (``_default_beta``, :267-282) unless given;
  * knots ``grid_knots(0, 0, W-1, H-1, n, n)`` (:509); lambda by GCV
    (``cv_folds=0``, the default here) or leave-stamp-out CV (``cv_folds``);
  * decorrelation only when both variance maps exist (:720);
  * score PSF: median unit-sum stamp of the target at the stamp centres,
    radius ``min(stamp_radius, 8)``, clipped at 0 (``_estimate_psf``,
    :285-302, :756); whitened by ``whiten_score_psf`` when the difference was
    decorrelated, scored against the post-whitening variance, else against
    Var(D) (:758-778);
  * candidates: seeds ``score >= threshold`` and their 8-connected labels,
    and the source catalog of the score (``build_catalog(score, difference,
    psf=psf)``, :792-796, src/catalog.cpp:600-697) built on device.

Stamps come from ``select_stamps_device`` (detection on both frames, or a
catalog) unless positions are given; the selection's direction decides which
frame is convolved, as ``Subtractor._fit`` does (pipeline.py:451-491): for
"science" the reference is the fit target and the difference is negated so
transients stay positive.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import _core, grid_knots
from .solution import KernelSolution
from .synth import default_beta

_FWHM_TO_SIGMA = 1.0 / 2.354820045030949


def estimate_psf(target_host_cutouts, radius):
    """Median unit-sum stamp (reference ``_estimate_psf``, pipeline.py:285-302).

    ``target_host_cutouts``: (n, 2r+1, 2r+1) float arrays of the target image
    at the stamp centres (already restricted to stamps fully inside the frame).
    """
    stamps = []
    for cut in target_host_cutouts:
        c = cut.astype(np.float64)
        c = c - np.median(c)
        total = c.sum()
        if total > 0:
            stamps.append(c / total)
    if not stamps:
        raise ValueError("no usable sources for PSF estimation")
    psf = np.median(np.stack(stamps), axis=0)
    psf = np.clip(psf, 0.0, None)
    psf /= psf.sum()
    return np.ascontiguousarray(psf, dtype=np.float32)


def _cutouts(img, xs, ys, radius):
    """Host copies of the (2r+1)^2 cutouts of a device image at integer centres."""
    import torch

    h, w = img.shape
    keep = [(int(x), int(y)) for x, y in zip(xs, ys)
            if x - radius >= 0 and x + radius < w and y - radius >= 0 and y + radius < h]
    if not keep:
        return np.zeros((0, 2 * radius + 1, 2 * radius + 1), np.float32)
    cuts = torch.stack([img[y - radius:y + radius + 1, x - radius:x + radius + 1]
                        for x, y in keep])
    return cuts.cpu().numpy()


@dataclasses.dataclass
class DiffResult:
    difference: object          # torch tensor (H, W) float32
    variance: object            # post-whitening variance when decorrelated, else Var(D)
    score: object = None
    labels: object = None       # int32 (H, W)
    n_candidates: int = 0
    source_catalog: np.ndarray | None = None  # delta_amd.catalog.CATALOG_DTYPE rows
    psf: np.ndarray | None = None
    fit: dict | None = None
    solution: KernelSolution | None = None


@dataclasses.dataclass
class _Layers:
    """The fit target and the convolved frame (with variances), the sign that
    keeps transients positive, and the stamps the score PSF is cut from."""
    target: object
    conv: object
    tvar: object
    cvar: object
    sign: float
    stamp_x: np.ndarray
    stamp_y: np.ndarray


def _layers(science, reference, science_var, reference_var, direction, stamp_x, stamp_y):
    if direction not in ("science", "reference"):
        raise ValueError(f"direction must be 'science', 'reference', or None, got {direction!r}")
    # The convolved frame is the sharper one; the other is the fit target.
    if direction == "science":
        return _Layers(reference, science, reference_var, science_var, -1.0, stamp_x, stamp_y)
    return _Layers(science, reference, science_var, reference_var, 1.0, stamp_x, stamp_y)


def fit_pair(science, reference, science_var, reference_var, stamp_x=None, stamp_y=None, *,
             n_knots=5, n_max=6, radius=0, stamp_radius=15, beta=None, fwhm=None,
             lambda_grid=None, clip_sigma=4.0, clip_iterations=5, min_stamps=5,
             precision="fp64", stream=None, direction=None, catalog=None, detect_sigma=5.0,
             max_stamps=200, saturation=0.0, match_radius=2, fwhm_radius=0, cv_folds=0):
    """Stamps, direction, beta and the kernel fit of one pair of CUDA tensors
    (``Subtractor._fit``, pipeline.py:430-600) -> (KernelSolution, fit dict, layers).

    Without ``stamp_x``/``stamp_y`` the stamps, direction and FWHMs come from
    ``select_stamps_device`` (``catalog``: an (n, 2) int array of x, y;
    ``detect_sigma`` its threshold).  With given positions the direction
    defaults to "reference" and ``fwhm`` = (reference, science) median FWHMs
    feeds the beta heuristic.
    """
    import torch

    h, w = science.shape
    if reference.shape != science.shape:
        raise ValueError("science and reference must share a shape")
    sel = None
    if stamp_x is None or stamp_y is None:
        cx = cy = None
        if catalog is not None:
            cat = np.asarray(catalog, dtype=np.int32)
            cx = torch.from_numpy(np.ascontiguousarray(cat[:, 0])).to(science.device)
            cy = torch.from_numpy(np.ascontiguousarray(cat[:, 1])).to(science.device)
        sel = _core.select_stamps_device(science, reference, catalog_x=cx, catalog_y=cy,
                                         stamp_radius=stamp_radius, threshold_sigma=detect_sigma,
                                         max_stamps=max_stamps, saturation=saturation,
                                         match_radius=match_radius, fwhm_radius=fwhm_radius,
                                         stream=stream)
        if sel["x"].numel() == 0:
            raise ValueError("no stamps selected; lower threshold_sigma, supply a catalog, or "
                             "widen match_radius if science/reference are astrometrically "
                             "misregistered")
        stamp_x = sel["x"].cpu().numpy()
        stamp_y = sel["y"].cpu().numpy()
        if direction is None:
            direction = sel["direction"]
        if beta is None:
            beta = default_beta(sel["median_fwhm_science"], sel["median_fwhm_reference"])
    if direction is None:
        direction = "reference"
    if beta is None:
        beta = default_beta(*fwhm) if fwhm is not None else 2.0
    lay = _layers(science, reference, science_var, reference_var, direction,
                  np.asarray(stamp_x), np.asarray(stamp_y))
    knots = grid_knots(0.0, 0.0, w - 1.0, h - 1.0, n_knots, n_knots)
    grid = np.logspace(-6, 6, 25) if lambda_grid is None else np.asarray(lambda_grid)
    rad = radius if radius > 0 else _core.default_radius(beta, n_max)
    fit = _core.fit_kernel_device(lay.target, lay.conv, knots, stamp_x, stamp_y,
                                  stamp_radius=stamp_radius, beta=beta, n_max=n_max,
                                  lambda_grid=grid, radius=rad, clip_sigma=clip_sigma,
                                  clip_iterations=clip_iterations, min_stamps=min_stamps,
                                  science_var=lay.tvar, reference_var=lay.cvar,
                                  precision=precision, stream=stream, cv_folds=cv_folds)
    fit["direction"] = direction
    fit["beta"] = beta
    if sel is not None:
        fit["selection"] = sel
    solution = KernelSolution.from_fit(fit, beta=beta, n_max=n_max, radius=radius, knots=knots,
                                       direction=direction, shape=(h, w))
    return solution, fit, lay


def _subtract(solution, lay, *, decorrelate=True, block=256, score=True, threshold=5.0,
              stamp_radius=15, source_catalog=True, stream=None):
    """Subtract, decorrelate, score, label and catalogue with a solution
    (``Subtractor._subtract``, pipeline.py:654-806)."""
    import torch

    knots, theta, beta, n_max = solution.knots, solution.theta, solution.beta, solution.n_max
    rad = solution.radius if solution.radius > 0 else _core.default_radius(beta, n_max)
    sub = _core.subtract_device(lay.target, lay.conv, knots, theta, beta, n_max, rad, 0.0,
                                lay.tvar, lay.cvar, stream=stream)
    diff, var = sub["difference"], sub["variance"]
    if lay.sign < 0:
        diff.neg_()
    white_var = None
    if decorrelate and lay.tvar is not None and lay.cvar is not None:
        dec = _core.decorrelate_device(diff, knots, theta, beta, n_max, lay.tvar, lay.cvar,
                                       block=block, radius=rad, stream=stream)
        diff, white_var = dec["difference"], dec["variance"]
    res = DiffResult(difference=diff, variance=white_var if white_var is not None else var,
                     solution=solution)
    if not score:
        return res
    if lay.stamp_x is None or np.asarray(lay.stamp_x).size == 0:
        raise ValueError("score requested but the solution carries no stamp positions "
                         "for PSF estimation")
    psf_radius = min(stamp_radius, 8)
    psf = estimate_psf(_cutouts(lay.target, lay.stamp_x, lay.stamp_y, psf_radius), psf_radius)
    if white_var is not None:
        score_var = white_var
        psf_dev = _core.whiten_score_psf_device(psf, knots, theta, beta, n_max, lay.tvar,
                                                lay.cvar, block, radius=rad, stream=stream)
    else:
        score_var = var if var is not None else torch.full_like(diff, float(diff.double().var()))
        psf_dev = torch.from_numpy(psf).to(diff.device)
    sc = _core.matched_filter_device(diff, psf_dev, score_var, stream=stream)
    labels, n = _core.label_components_device(sc, threshold, 1, stream=stream)
    res.score, res.labels, res.n_candidates, res.psf = sc, labels, n, psf_dev.cpu().numpy()
    if source_catalog:
        from .catalog import _pack, measure_psf_fwhm

        cat_kw = dict(threshold_sigma=threshold, expected_fwhm=measure_psf_fwhm(res.psf),
                      stream=stream)
        rows, nrow = _core.build_catalog_device(sc, diff, capacity=max(4 * n, 1024), **cat_kw)
        if nrow > rows.shape[0]:
            rows, nrow = _core.build_catalog_device(sc, diff, capacity=nrow, **cat_kw)
        res.source_catalog = _pack(_core.catalog_rows_to_dict(rows, nrow))
    return res


def subtract_pair(science, reference, science_var, reference_var, stamp_x=None, stamp_y=None,
                  *, n_knots=5, n_max=6, radius=0, stamp_radius=15, beta=None, fwhm=None,
                  lambda_grid=None, clip_sigma=4.0, clip_iterations=5, min_stamps=5,
                  decorrelate=True, block=256, score=True, threshold=5.0,
                  precision="fp64", stream=None, direction=None, catalog=None,
                  detect_sigma=5.0, max_stamps=200, saturation=0.0, match_radius=2,
                  fwhm_radius=0, source_catalog=True, cv_folds=0):
    """Fit, then subtract, decorrelate, score and label one pair of CUDA
    tensors (``Subtractor.subtract``); ``threshold`` is the score threshold.
    The result carries the fit dict and the :class:`KernelSolution`."""
    solution, fit, lay = fit_pair(
        science, reference, science_var, reference_var, stamp_x, stamp_y, n_knots=n_knots,
        n_max=n_max, radius=radius, stamp_radius=stamp_radius, beta=beta, fwhm=fwhm,
        lambda_grid=lambda_grid, clip_sigma=clip_sigma, clip_iterations=clip_iterations,
        min_stamps=min_stamps, precision=precision, stream=stream, direction=direction,
        catalog=catalog, detect_sigma=detect_sigma, max_stamps=max_stamps,
        saturation=saturation, match_radius=match_radius, fwhm_radius=fwhm_radius,
        cv_folds=cv_folds)
    res = _subtract(solution, lay, decorrelate=decorrelate, block=block, score=score,
                    threshold=threshold, stamp_radius=stamp_radius,
                    source_catalog=source_catalog, stream=stream)
    res.fit = fit
    return res


def apply_solution(solution, science, reference, science_var=None, reference_var=None, *,
                   decorrelate=True, block=256, score=True, threshold=5.0, stamp_radius=15,
                   source_catalog=True, stream=None):
    """A frozen solution on a new pair, no re-fit (``Subtractor.apply``,
    pipeline.py:872-908, layers as ``_coerce_apply_layers`` :603-652): the
    direction, basis, knots and coefficients come from ``solution``, the score
    PSF from its stamps."""
    if science.shape != reference.shape:
        raise ValueError("science and reference must share a shape")
    if tuple(science.shape) != tuple(solution.shape):
        raise ValueError(f"solution shape {tuple(solution.shape)} does not match input frame "
                         f"shape {tuple(science.shape)}")
    lay = _layers(science, reference, science_var, reference_var, solution.direction,
                  solution.stamp_x, solution.stamp_y)
    return _subtract(solution, lay, decorrelate=decorrelate, block=block, score=score,
                     threshold=threshold, stamp_radius=stamp_radius,
                     source_catalog=source_catalog, stream=stream)
