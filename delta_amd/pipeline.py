"""The reference's difference-imaging pipeline (python/delta/pipeline.py) on
the GPU, in two layers.

``fit_pair`` / ``subtract_pair`` / ``apply_solution`` run ``Subtractor._fit``
and ``Subtractor._subtract`` (pipeline.py:430-806) on torch CUDA tensors that
stay in HBM, with every option explicit (GCV unless ``cv_folds`` > 1);
``Subtractor``, ``subtract()`` and ``apply()`` wrap them with the reference's
constructor knobs, defaults (5-fold CV, no decorrelation or score unless
asked) and input coercion, so host arrays in give host arrays out.  The
choices are the reference's:

  * direction from the stamp selection (or given): the convolved frame is the
    sharper one, the other the fit target; for "science" the difference is
    negated so transients stay positive (:451-491, :688-689);
  * beta from the two median FWHMs (``_default_beta``, :267-282) unless given;
  * knots ``grid_knots(0, 0, W-1, H-1, nx, ny)``, fewer per axis when
    ``spatial_scale`` asks for it (:504-509);
  * output bright-core masking at the lowest positive of ``saturation`` and
    ``bright_mask`` (:663-664); the optional variance rescale to a unit
    reduced chi^2 (:66-86, :693-705);
  * decorrelation only when both variance maps exist (:720);
  * score PSF: median unit-sum stamp of the target at the stamp centres,
    radius ``min(stamp_radius, 8)``, clipped at 0 (``_estimate_psf``,
    :285-302, :756); whitened by ``whiten_score_psf`` when the difference was
    decorrelated, scored against the post-whitening variance, else against
    Var(D), else a constant (:758-778);
  * the source catalog of the score with the output mask (``build_catalog
    (score, difference, mask=mask, psf=psf)``, :792-796) built on device;
    the seeds ``score >= threshold`` and their 8-connected labels are kept too.
"""
from __future__ import annotations

import dataclasses
import logging
import time

import numpy as np

from . import _core, grid_knots
from .solution import KernelSolution
from .synth import default_beta

# largest kernel radius the device subtraction runs (the radius-30 kernels;
# radii 21-30 run zero-padded on the radius-25 / radius-30 instantiations)
MAX_SUBTRACT_RADIUS = 30

logger = logging.getLogger("delta_amd")


@dataclasses.dataclass
class DiffResult:
    """Products of a subtraction (the reference's ``DiffResult``,
    pipeline.py:104-264).  Images are torch CUDA tensors from the device
    functions and numpy arrays from :class:`Subtractor` on host inputs."""

    difference: object          # float32 (H, W); whitened when decorrelated
    variance: object            # post-whitening variance when decorrelated, else Var(D)
    mask: object = None         # uint8 (H, W) or None
    score: object = None
    solution: KernelSolution | None = None
    psf: np.ndarray | None = None
    source_catalog: np.ndarray | None = None  # delta_amd.catalog.CATALOG_DTYPE rows
    config: dict = dataclasses.field(default_factory=dict)
    elapsed: float = float("nan")
    variance_scale: float | None = None
    labels: object = None       # int32 (H, W): 8-connected seeds score >= threshold
    n_candidates: int = 0
    fit: dict | None = None

    def header_cards(self) -> dict:
        """Fit and configuration provenance (pipeline.py:135-148); the
        reference's host/user/git environment cards are not produced here."""
        return {**self.solution.header_cards(), **self.config, "DLTELAP": self.elapsed}

    def build_catalog(self, *, threshold_sigma=5.0, threshold_sigma_dipole=3.0,
                      fwhm_tolerance_lo=0.3, fwhm_tolerance_hi=3.0, aperture_radius=0,
                      exclude_bad_pixels=True, fit_psf_shape=False, shape_chi2_tolerance=3.0,
                      bright_x=None, bright_y=None, bright_star_radius=0.0, max_shape_fit=5000,
                      polarity="positive", as_table=False):
        """Rebuild the source catalog from ``score`` with custom thresholds
        (pipeline.py:196-264); bright stars default to the solution's stamps."""
        if self.score is None or self.psf is None:
            raise ValueError("build_catalog() requires a score image and PSF stamp; pass "
                             "score=True to subtract()")
        if fit_psf_shape and self.variance is None:
            raise ValueError("fit_psf_shape=True requires a variance map (self.variance)")
        from .catalog import build_catalog as _build_catalog

        if bright_x is None and bright_y is None and self.solution is not None:
            bright_x, bright_y = self.solution.stamp_x, self.solution.stamp_y

        def host(a):
            return a.cpu().numpy() if hasattr(a, "cpu") else a

        return _build_catalog(
            host(self.score), host(self.difference), mask=host(self.mask), psf=self.psf,
            threshold_sigma=threshold_sigma, threshold_sigma_dipole=threshold_sigma_dipole,
            fwhm_tolerance_lo=fwhm_tolerance_lo, fwhm_tolerance_hi=fwhm_tolerance_hi,
            aperture_radius=aperture_radius, exclude_bad_pixels=exclude_bad_pixels,
            fit_psf_shape=fit_psf_shape, shape_chi2_tolerance=shape_chi2_tolerance,
            variance=host(self.variance) if fit_psf_shape else None, bright_x=bright_x,
            bright_y=bright_y, bright_star_radius=bright_star_radius,
            max_shape_fit=max_shape_fit, polarity=polarity, as_table=as_table)


@dataclasses.dataclass
class _Layers:
    """The fit target and the convolved frame (with variances and masks), the
    sign that keeps transients positive, and the stamps the score PSF is cut
    from."""
    target: object
    conv: object
    tvar: object
    cvar: object
    sign: float
    stamp_x: np.ndarray
    stamp_y: np.ndarray
    tmask: object = None
    cmask: object = None


def _layers(science, reference, science_var, reference_var, direction, stamp_x, stamp_y,
            science_mask=None, reference_mask=None):
    if direction not in ("science", "reference"):
        raise ValueError(f"direction must be 'science', 'reference', or None, got {direction!r}")
    # The convolved frame is the sharper one; the other is the fit target.
    if direction == "science":
        return _Layers(reference, science, reference_var, science_var, -1.0, stamp_x, stamp_y,
                       reference_mask, science_mask)
    return _Layers(science, reference, science_var, reference_var, 1.0, stamp_x, stamp_y,
                   science_mask, reference_mask)


def knot_counts(w, h, n_knots, spatial_scale=None):
    """Knots per axis: ``n_knots``, or fewer so that knots sit no closer than
    ``spatial_scale`` pixels, at least 3 (pipeline.py:504-509)."""
    nkx = nky = n_knots
    if spatial_scale is not None and spatial_scale > 0:
        nkx = int(np.clip(round((w - 1) / spatial_scale) + 1, 3, n_knots))
        nky = int(np.clip(round((h - 1) / spatial_scale) + 1, 3, n_knots))
    return nkx, nky


def _fit_warnings(fit, cv_folds):
    selector = f"{cv_folds}-fold CV" if cv_folds > 1 else "GCV"
    grid = np.asarray(fit["lambda_grid"])
    if grid.size > 1 and fit["lambda"] in (grid[0], grid[-1]):
        edge = "lower" if fit["lambda"] == grid[0] else "upper"
        logger.warning("selected lambda=%.3g (%s) sits at the %s edge of lambda_grid "
                       "[%.3g, %.3g] -- consider widening lambda_grid", fit["lambda"], selector,
                       edge, grid[0], grid[-1])
    design_bytes = fit.get("cv_exact_design_bytes", 0)
    if design_bytes > _CV_EXACT_DESIGN_WARN_BYTES:
        logger.warning("exact (non-stamped) k-fold CV path used a whitened design equivalent to "
                       "%.2f GB, rebuilt every IRLS pass -- consider fewer/coarser knots, a "
                       "larger spatial_scale, or a smaller stamp_radius", design_bytes / 1e9)
    if fit["reduced_chi2"] > 3.0:
        logger.warning("kernel fit reduced chi2=%.2f >> 1: stellar residuals likely. Check "
                       "registration (dipoles), basis (n_max/beta), or knots.",
                       fit["reduced_chi2"])


def fit_pair(science, reference, science_var, reference_var, stamp_x=None, stamp_y=None, *,
             n_knots=5, n_max=6, radius=0, stamp_radius=15, beta=None, fwhm=None,
             lambda_grid=None, clip_sigma=4.0, clip_iterations=5, min_stamps=5,
             precision="fp64", stream=None, direction=None, catalog=None, detect_sigma=5.0,
             max_stamps=200, saturation=0.0, match_radius=2, fwhm_radius=0, cv_folds=0,
             science_mask=None, reference_mask=None, spatial_scale=None):
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
        sel = _core.select_stamps_device(science, reference, science_mask=science_mask,
                                         reference_mask=reference_mask, catalog_x=cx,
                                         catalog_y=cy, stamp_radius=stamp_radius,
                                         threshold_sigma=detect_sigma, max_stamps=max_stamps,
                                         saturation=saturation, match_radius=match_radius,
                                         fwhm_radius=fwhm_radius, stream=stream)
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
                  np.asarray(stamp_x), np.asarray(stamp_y), science_mask, reference_mask)
    nkx, nky = knot_counts(w, h, n_knots, spatial_scale)
    knots = grid_knots(0.0, 0.0, w - 1.0, h - 1.0, nkx, nky)
    grid = np.logspace(-6, 6, 25) if lambda_grid is None else np.asarray(lambda_grid)
    rad = radius if radius > 0 else _core.default_radius(beta, n_max)
    if rad > MAX_SUBTRACT_RADIUS:
        # the fit would succeed and the subtraction then refuse the kernel
        raise ValueError(f"kernel radius {rad} (beta {beta:.3f}, n_max {n_max}) exceeds the "
                         f"GPU subtraction's limit of {MAX_SUBTRACT_RADIUS} px; pass a smaller "
                         "beta or radius")
    fit = _core.fit_kernel_device(lay.target, lay.conv, knots, stamp_x, stamp_y,
                                  stamp_radius=stamp_radius, beta=beta, n_max=n_max,
                                  lambda_grid=grid, radius=rad, clip_sigma=clip_sigma,
                                  clip_iterations=clip_iterations, min_stamps=min_stamps,
                                  science_var=lay.tvar, reference_var=lay.cvar,
                                  science_mask=lay.tmask, reference_mask=lay.cmask,
                                  precision=precision, stream=stream, cv_folds=cv_folds)
    _fit_warnings(fit, cv_folds)
    fit["direction"] = direction
    fit["beta"] = beta
    if sel is not None:
        fit["selection"] = sel
    solution = KernelSolution.from_fit(fit, beta=beta, n_max=n_max, radius=radius, knots=knots,
                                       direction=direction, shape=(h, w))
    return solution, fit, lay


def _subtract(solution, lay, *, decorrelate=True, block=256, score=True, threshold=5.0,
              stamp_radius=15, source_catalog=True, stream=None, saturation=0.0,
              bright_mask=0.0, rescale_variance=False):
    """Subtract, rescale, decorrelate, score, label and catalogue with a
    solution (``Subtractor._subtract``, pipeline.py:654-806)."""
    import torch

    knots, theta, beta, n_max = solution.knots, solution.theta, solution.beta, solution.n_max
    rad = solution.radius if solution.radius > 0 else _core.default_radius(beta, n_max)
    # output bright-core level: the lowest positive of saturation and bright_mask
    levels = [v for v in (saturation, bright_mask) if v and v > 0]
    out_saturation = min(levels) if levels else 0.0
    sub = _core.subtract_device(lay.target, lay.conv, knots, theta, beta, n_max, rad,
                                out_saturation, lay.tvar, lay.cvar, lay.tmask, lay.cmask,
                                stream=stream)
    diff, var, mask = sub["difference"], sub["variance"], sub["mask"]
    if lay.sign < 0:
        diff.neg_()
    scale = None
    if rescale_variance and var is not None:
        scale = _core.variance_rescale_factor_device(diff, var, mask, stream=stream)
        var = var * scale
        logger.info("variance rescaled by %.4f to force diff-image reduced chi2=1", scale)
    elif rescale_variance:
        logger.warning("variance rescale requested but no variance available; skipping")
    white_var = None
    if decorrelate and lay.tvar is not None and lay.cvar is not None:
        # the reference decorrelates with the default footprint of (beta, n_max),
        # whatever radius the solution was fitted with (pipeline.py:721-730)
        dec = _core.decorrelate_device(diff, knots, theta, beta, n_max, lay.tvar, lay.cvar,
                                       block=block, radius=0, stream=stream)
        diff, white_var = dec["difference"], dec["variance"]
        if scale is not None:
            white_var.mul_(scale)
    elif decorrelate:
        logger.warning("decorrelation requested but no variance available; skipping")
    res = DiffResult(difference=diff, variance=white_var if white_var is not None else var,
                     mask=mask, solution=solution, variance_scale=scale)
    if not score:
        return res
    if lay.stamp_x is None or np.asarray(lay.stamp_x).size == 0:
        raise ValueError("score requested but the solution carries no stamp positions "
                         "for PSF estimation")
    psf_radius = min(stamp_radius, 8)
    psf = _core.estimate_psf_device(lay.target, lay.stamp_x, lay.stamp_y, psf_radius,
                                    stream=stream)
    if white_var is not None:
        score_var = white_var
        psf_dev = _core.whiten_score_psf_device(psf, knots, theta, beta, n_max, lay.tvar,
                                                lay.cvar, block, radius=rad, stream=stream)
    elif var is not None:
        score_var = var
        psf_dev = torch.from_numpy(psf).to(diff.device)
    else:
        # constant variance of the difference itself (population variance)
        score_var = torch.full_like(diff, float(diff.double().var(unbiased=False)))
        psf_dev = torch.from_numpy(psf).to(diff.device)
    sc = _core.matched_filter_device(diff, psf_dev, score_var, stream=stream)
    labels, n = _core.label_components_device(sc, threshold, 1, stream=stream)
    res.score, res.labels, res.n_candidates, res.psf = sc, labels, n, psf_dev.cpu().numpy()
    if source_catalog:
        from .catalog import _pack, measure_psf_fwhm

        cat_kw = dict(mask=mask, threshold_sigma=threshold,
                      expected_fwhm=measure_psf_fwhm(res.psf), stream=stream)
        if mask is None:
            # the catalog's positive seeds are exactly the ones just labelled
            # (same score, threshold and polarity, no seed mask): reuse them
            cat_kw.update(labels=labels, n_labels=n)
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
                  fwhm_radius=0, source_catalog=True, cv_folds=0, science_mask=None,
                  reference_mask=None, spatial_scale=None, bright_mask=0.0,
                  rescale_variance=False):
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
        cv_folds=cv_folds, science_mask=science_mask, reference_mask=reference_mask,
        spatial_scale=spatial_scale)
    res = _subtract(solution, lay, decorrelate=decorrelate, block=block, score=score,
                    threshold=threshold, stamp_radius=stamp_radius,
                    source_catalog=source_catalog, stream=stream, saturation=saturation,
                    bright_mask=bright_mask, rescale_variance=rescale_variance)
    res.fit = fit
    return res


def apply_solution(solution, science, reference, science_var=None, reference_var=None, *,
                   decorrelate=True, block=256, score=True, threshold=5.0, stamp_radius=15,
                   source_catalog=True, stream=None, science_mask=None, reference_mask=None,
                   saturation=0.0, bright_mask=0.0, rescale_variance=False):
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
                  solution.stamp_x, solution.stamp_y, science_mask, reference_mask)
    return _subtract(solution, lay, decorrelate=decorrelate, block=block, score=score,
                     threshold=threshold, stamp_radius=stamp_radius,
                     source_catalog=source_catalog, stream=stream, saturation=saturation,
                     bright_mask=bright_mask, rescale_variance=rescale_variance)


# --------------------------------------------------------------------------- Subtractor
def _split(obj):
    """(data, variance, mask) of an array, a torch tensor or an astropy-like
    object with .data / .uncertainty / .mask (python/delta/_inputs.py:18-44)."""
    if hasattr(obj, "is_cuda"):
        return obj, None, None
    data = getattr(obj, "data", None)
    # a plain ndarray's .data is its raw buffer; a MaskedArray (an ndarray
    # subclass) carries its data and per-pixel mask as the reference reads them
    if data is None or (isinstance(obj, np.ndarray) and not isinstance(obj, np.ma.MaskedArray)):
        return np.asarray(obj), None, None
    variance = None
    unc = getattr(obj, "uncertainty", None)
    if unc is not None and getattr(unc, "array", None) is not None:
        arr = np.asarray(unc.array, dtype=np.float64)
        variance = arr ** 2 if getattr(unc, "uncertainty_type", "std") == "std" else arr
    mask = getattr(obj, "mask", None)
    if mask is np.ma.nomask:
        mask = None
    return np.asarray(data), variance, mask


def as_layers(obj, variance=None, mask=None):
    """(data float32, variance float32 | None, mask uint8 | None), explicit
    arguments overriding the object's own layers; a boolean mask is True ==
    bad (python/delta/_inputs.py:47-72).  Host inputs give numpy arrays, CUDA
    tensors stay tensors."""
    data, v, m = _split(obj)
    if variance is not None:
        v = variance
    if mask is not None:
        m = mask
    if hasattr(data, "is_cuda"):
        import torch

        data = data.to(torch.float32).contiguous()
        if data.dim() != 2:
            raise ValueError(f"expected a 2-D image, got shape {tuple(data.shape)}")
        if v is not None:
            v = torch.as_tensor(v, device=data.device).to(torch.float32).contiguous()
            if v.shape != data.shape:
                raise ValueError("variance shape does not match data")
        if m is not None:
            m = torch.as_tensor(m, device=data.device).to(torch.uint8).contiguous()
            if m.shape != data.shape:
                raise ValueError("mask shape does not match data")
        return data, v, m
    data = np.ascontiguousarray(data, dtype=np.float32)
    if data.ndim != 2:
        raise ValueError(f"expected a 2-D image, got shape {data.shape}")
    if v is not None:
        v = np.ascontiguousarray(v, dtype=np.float32)
        if v.shape != data.shape:
            raise ValueError("variance shape does not match data")
    if m is not None:
        m = np.asarray(m)
        if m.dtype == bool:
            m = m.astype(np.uint8)
        m = np.ascontiguousarray(m, dtype=np.uint8)
        if m.shape != data.shape:
            raise ValueError("mask shape does not match data")
    return data, v, m


def synth_variance(data, gain, read_noise=0.0):
    """Var = max(data, 0) / gain + read_noise^2 (python/delta/_inputs.py:75-86),
    computed on the GPU; numpy in, numpy out (CUDA tensor in, tensor out)."""
    if hasattr(data, "is_cuda"):
        return _core.synth_variance_device(data, gain, read_noise)
    return _core.synth_variance(data, gain, read_noise)


_CONFIG_KEYS = frozenset({
    "n_max", "beta", "n_knots", "radius", "stamp_radius", "threshold_sigma", "max_stamps",
    "saturation", "match_radius", "fwhm_radius", "bright_mask", "lambda_grid", "clip_sigma",
    "clip_iterations", "min_stamps", "cv_folds", "spatial_scale", "decorrelate", "score",
    "source_catalog", "block", "rescale_variance", "precision"})

# warn when the exact (per-row) CV design exceeds this many bytes
_CV_EXACT_DESIGN_WARN_BYTES = 256 * 1024 * 1024


class Subtractor:
    """Reusable difference-imaging configuration with the reference's knobs
    and defaults (python/delta/pipeline.py:305-406); every stage runs on the
    GPU.  Host (numpy / astropy-like) inputs give numpy products, CUDA tensors
    give tensors.  ``precision`` selects the normal-equation build ("fp64",
    the reference's arithmetic, by default)."""

    def __init__(self, *, n_max=6, beta=None, n_knots=5, radius=0, stamp_radius=15,
                 threshold_sigma=5.0, max_stamps=200, saturation=0.0, match_radius=2,
                 fwhm_radius=0, bright_mask=0.0, lambda_grid=None, clip_sigma=4.0,
                 clip_iterations=5, min_stamps=5, cv_folds=5, spatial_scale=None,
                 decorrelate=False, score=False, source_catalog=True, block=256,
                 rescale_variance=False, precision="fp64"):
        self.n_max = n_max
        self.beta = beta
        self.n_knots = n_knots
        self.radius = radius
        self.stamp_radius = stamp_radius
        self.threshold_sigma = threshold_sigma
        self.max_stamps = max_stamps
        self.match_radius = match_radius
        self.fwhm_radius = fwhm_radius
        self.saturation = saturation
        self.bright_mask = bright_mask
        self.lambda_grid = (np.logspace(-6.0, 6.0, 25) if lambda_grid is None
                            else np.asarray(lambda_grid, dtype=np.float64))
        self.clip_sigma = clip_sigma
        self.clip_iterations = clip_iterations
        self.min_stamps = min_stamps
        self.cv_folds = cv_folds
        self.spatial_scale = spatial_scale
        self.decorrelate = decorrelate
        self.score = score
        self.source_catalog = source_catalog
        self.block = block
        self.rescale_variance = rescale_variance
        self.precision = precision

    def config_cards(self) -> dict:
        """Reproducibility cards for the knobs not in the solution (pipeline.py:386-406)."""
        return {
            "DLTSRAD": self.stamp_radius, "DLTTHSIG": self.threshold_sigma,
            "DLTMAXST": self.max_stamps, "DLTSAT": self.saturation, "DLTMRAD": self.match_radius,
            "DLTFRAD": self.fwhm_radius, "DLTBMASK": self.bright_mask,
            "DLTCLPSG": self.clip_sigma, "DLTCLPIT": self.clip_iterations,
            "DLTMINST": self.min_stamps, "DLTCVF": self.cv_folds,
            "DLTSSCL": self.spatial_scale if self.spatial_scale is not None else 0.0,
            "DLTDECOR": self.decorrelate, "DLTSCORE": self.score,
            "DLTSCAT": self.source_catalog, "DLTBLK": self.block,
        }

    def _device_inputs(self, science, reference, science_var, reference_var, science_mask,
                       reference_mask, gain, read_noise):
        import torch

        sci, svar, smask = as_layers(science, science_var, science_mask)
        ref, rvar, rmask = as_layers(reference, reference_var, reference_mask)
        if tuple(sci.shape) != tuple(ref.shape):
            raise ValueError("science and reference must share a shape")
        host = not hasattr(sci, "is_cuda")
        dev = torch.device("cuda", torch.cuda.current_device()) if host else sci.device

        def up(a):
            if a is None or hasattr(a, "is_cuda"):
                return a
            return torch.from_numpy(a).to(dev)

        sci, svar, smask, ref, rvar, rmask = map(up, (sci, svar, smask, ref, rvar, rmask))
        if svar is None and gain is not None:
            svar = _core.synth_variance_device(sci, gain, read_noise)
        if rvar is None and gain is not None:
            rvar = _core.synth_variance_device(ref, gain, read_noise)
        return host, (sci, ref, svar, rvar, smask, rmask)

    def _fit(self, science, reference, *, science_var, reference_var, science_mask,
             reference_mask, gain, read_noise, catalog, direction):
        host, (sci, ref, svar, rvar, smask, rmask) = self._device_inputs(
            science, reference, science_var, reference_var, science_mask, reference_mask, gain,
            read_noise)
        solution, fit, lay = fit_pair(
            sci, ref, svar, rvar, n_knots=self.n_knots, n_max=self.n_max, radius=self.radius,
            stamp_radius=self.stamp_radius, beta=self.beta, lambda_grid=self.lambda_grid,
            clip_sigma=self.clip_sigma, clip_iterations=self.clip_iterations,
            min_stamps=self.min_stamps, precision=self.precision, direction=direction,
            catalog=catalog, detect_sigma=self.threshold_sigma, max_stamps=self.max_stamps,
            saturation=self.saturation, match_radius=self.match_radius,
            fwhm_radius=self.fwhm_radius, cv_folds=self.cv_folds, science_mask=smask,
            reference_mask=rmask, spatial_scale=self.spatial_scale)
        return host, solution, fit, lay

    def _run(self, solution, lay, host):
        res = _subtract(solution, lay, decorrelate=self.decorrelate, block=self.block,
                        score=self.score, stamp_radius=self.stamp_radius,
                        source_catalog=self.source_catalog, saturation=self.saturation,
                        bright_mask=self.bright_mask, rescale_variance=self.rescale_variance)
        res.config = self.config_cards()
        if host:
            for name in ("difference", "variance", "mask", "score", "labels"):
                v = getattr(res, name)
                if v is not None:
                    setattr(res, name, v.cpu().numpy())
        return res

    def fit(self, science, reference, *, science_var=None, reference_var=None,
            science_mask=None, reference_mask=None, gain=None, read_noise=0.0, catalog=None,
            direction=None) -> KernelSolution:
        """Select stamps and fit the kernel and background (pipeline.py:810-837)."""
        _, solution, _, _ = self._fit(science, reference, science_var=science_var,
                                      reference_var=reference_var, science_mask=science_mask,
                                      reference_mask=reference_mask, gain=gain,
                                      read_noise=read_noise, catalog=catalog, direction=direction)
        return solution

    def subtract(self, science, reference, *, science_var=None, reference_var=None,
                 science_mask=None, reference_mask=None, gain=None, read_noise=0.0,
                 catalog=None, direction=None) -> DiffResult:
        """Fit and produce the difference (+ decorrelation / score / catalog)
        (pipeline.py:839-870)."""
        start = time.perf_counter()
        host, solution, fit, lay = self._fit(
            science, reference, science_var=science_var, reference_var=reference_var,
            science_mask=science_mask, reference_mask=reference_mask, gain=gain,
            read_noise=read_noise, catalog=catalog, direction=direction)
        res = self._run(solution, lay, host)
        res.fit = fit
        res.elapsed = time.perf_counter() - start
        return res

    def apply(self, solution, science, reference, *, science_var=None, reference_var=None,
              science_mask=None, reference_mask=None, gain=None, read_noise=0.0) -> DiffResult:
        """A frozen solution on a new pair without re-fitting (pipeline.py:872-908)."""
        start = time.perf_counter()
        host, (sci, ref, svar, rvar, smask, rmask) = self._device_inputs(
            science, reference, science_var, reference_var, science_mask, reference_mask, gain,
            read_noise)
        if tuple(sci.shape) != tuple(solution.shape):
            raise ValueError(f"solution shape {tuple(solution.shape)} does not match input "
                             f"frame shape {tuple(sci.shape)}")
        lay = _layers(sci, ref, svar, rvar, solution.direction, solution.stamp_x,
                      solution.stamp_y, smask, rmask)
        res = self._run(solution, lay, host)
        res.elapsed = time.perf_counter() - start
        return res


def subtract(science, reference, **kwargs) -> DiffResult:
    """Configure a :class:`Subtractor` and run it once (pipeline.py:911-924):
    configuration keywords are split from the per-call inputs."""
    config = {k: v for k, v in kwargs.items() if k in _CONFIG_KEYS}
    call = {k: v for k, v in kwargs.items() if k not in _CONFIG_KEYS}
    return Subtractor(**config).subtract(science, reference, **call)


def apply(solution, science, reference, **kwargs) -> DiffResult:
    """Apply a frozen ``solution`` without re-fitting (pipeline.py:927-940)."""
    config = {k: v for k, v in kwargs.items() if k in _CONFIG_KEYS}
    call = {k: v for k, v in kwargs.items() if k not in _CONFIG_KEYS}
    return Subtractor(**config).apply(solution, science, reference, **call)
