"""delta_amd -- MI355X-native implementation of delta's PSF-matching hot path.

``delta_amd._core`` mirrors the reference's ``delta._core`` entry points
(src/bindings.cpp) on top of libdeltagpu (include/deltagpu.h), whose kernels
run on gfx950.  Importing the package requires the built library.
"""
from . import _core  # noqa: F401  (fails loudly if libdeltagpu.so is missing)
from ._core import default_radius, subtract, subtract_device  # noqa: F401

__all__ = ["_core", "default_radius", "subtract", "subtract_device", "grid_knots"]


def grid_knots(x0: float, y0: float, x1: float, y1: float, nx: int, ny: int):
    """Regular nx*ny knot grid over [x0,x1]x[y0,y1], shape (nx*ny, 2).

    Same layout as the reference's grid_knots (src/spatial.cpp:113-129):
    x varies fastest; a single knot along an axis sits at the midpoint.
    """
    import numpy as np

    if nx < 1 or ny < 1:
        raise ValueError("nx, ny must be >= 1")
    fx = np.full(1, 0.5) if nx == 1 else np.arange(nx) / (nx - 1)
    fy = np.full(1, 0.5) if ny == 1 else np.arange(ny) / (ny - 1)
    xs = x0 + fx * (x1 - x0)
    ys = y0 + fy * (y1 - y0)
    gx, gy = np.meshgrid(xs, ys)
    return np.column_stack([gx.ravel(), gy.ravel()]).astype(np.float64)
