"""delta_amd -- MI355X-native implementation of delta's PSF-matching hot path.

``delta_amd._core`` mirrors the reference's ``delta._core`` entry points
(src/bindings.cpp) on top of libdeltagpu (include/deltagpu.h), whose kernels
run on gfx950.  Importing the package requires the built library.
"""
from . import _core  # noqa: F401  (fails loudly if libdeltagpu.so is missing)
from ._core import default_radius, grid_knots, subtract, subtract_device  # noqa: F401

__all__ = ["_core", "default_radius", "subtract", "subtract_device", "grid_knots"]
