"""delta_amd -- MI355X-native implementation of delta's PSF-matching hot path.

``delta_amd._core`` mirrors the reference's ``delta._core`` entry points
(src/bindings.cpp) on top of libdeltagpu (include/deltagpu.h), whose kernels
run on gfx950; the package namespace mirrors the reference's ``delta``
(python/delta/__init__.py): the ``_core`` functions, ``Subtractor`` /
``subtract`` / ``apply``, ``KernelSolution`` and ``catalog``.  Importing the
package requires the built library.
"""
from . import _core  # noqa: F401  (fails loudly if libdeltagpu.so is missing)
from ._core import (  # noqa: F401
    basis_convolve,
    decorrelate,
    decorrelate_block,
    decorrelation_kernel,
    default_radius,
    detect_stamps,
    estimate_background,
    fit_kernel,
    gauss_hermite_basis1d,
    gauss_hermite_kernels,
    grid_knots,
    matched_filter,
    photometric_scale,
    photometric_scale_at,
    select_stamps,
    solve_gls,
    solve_gls_cv,
    solve_gls_gcv,
    subtract_device,
    tps_design,
    tps_evaluate,
    tps_fit,
    tps_penalty,
    weighted_mean,
    whiten_score_psf,
)
from ._core import subtract as subtract_model  # noqa: F401
from . import catalog  # noqa: F401,E402
from .pipeline import DiffResult, Subtractor, apply, as_layers, subtract, synth_variance  # noqa: F401,E402
from .solution import KernelSolution  # noqa: F401,E402

__version__ = "%d.%d.%d" % (_core.version() // 10000, _core.version() // 100 % 100,
                            _core.version() % 100)

__all__ = [
    "DiffResult", "KernelSolution", "Subtractor", "__version__", "apply", "as_layers",
    "basis_convolve", "catalog", "decorrelate", "decorrelate_block", "decorrelation_kernel",
    "default_radius", "detect_stamps", "estimate_background", "fit_kernel",
    "gauss_hermite_basis1d", "gauss_hermite_kernels", "grid_knots", "matched_filter",
    "photometric_scale", "photometric_scale_at", "select_stamps", "solve_gls", "solve_gls_cv",
    "solve_gls_gcv", "subtract", "subtract_device", "subtract_model", "synth_variance",
    "tps_design", "tps_evaluate", "tps_fit", "tps_penalty", "weighted_mean", "whiten_score_psf",
]
