"""The fitted kernel/background solution and its .npz file format, the
counterpart of the reference's python/delta/solution.py:13-143.

The file holds the same keys, dtypes and shapes as the reference writes
(``np.savez``), so solutions move between the two builds in both directions;
the photometric-scale methods run on the device.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _core


@dataclass
class KernelSolution:
    """Gauss-Hermite basis (beta, n_max, radius), thin-plate knots, coefficient
    vector theta ordered [c_0 | ... | c_{nc-1} | b] and the GCV diagnostics;
    ``direction`` is the convolved image ("reference" or "science"), ``shape``
    the full frame (height, width)."""

    beta: float
    n_max: int
    radius: int
    knots: np.ndarray
    theta: np.ndarray
    n_components: int
    n_spatial: int
    lam: float
    gcv: float
    effective_dof: float
    rss: float
    component_sums: np.ndarray
    direction: str
    shape: tuple
    reduced_chi2: float = float("nan")
    n_stamps_used: int = 0
    n_stamps_rejected: int = 0
    stamp_x: np.ndarray | None = None
    stamp_y: np.ndarray | None = None
    stamp_chi2: np.ndarray | None = None
    stamp_accepted: np.ndarray | None = None

    @classmethod
    def from_fit(cls, fit, *, beta, n_max, radius, knots, direction, shape):
        """From a ``fit_kernel`` result dict, as ``Subtractor._fit`` builds it
        (python/delta/pipeline.py:578-600)."""
        return cls(beta=float(beta), n_max=int(n_max), radius=int(radius),
                   knots=np.asarray(knots, np.float64), theta=np.asarray(fit["theta"]),
                   n_components=int(fit["n_components"]), n_spatial=int(fit["n_spatial"]),
                   lam=float(fit["lambda"]), gcv=float(fit["gcv"]),
                   effective_dof=float(fit["effective_dof"]), rss=float(fit["rss"]),
                   component_sums=np.asarray(fit["component_sums"]), direction=str(direction),
                   shape=(int(shape[0]), int(shape[1])),
                   reduced_chi2=float(fit["reduced_chi2"]),
                   n_stamps_used=int(fit["n_stamps_used"]),
                   n_stamps_rejected=int(fit["n_stamps_rejected"]),
                   stamp_x=fit["stamp_x"], stamp_y=fit["stamp_y"],
                   stamp_chi2=fit["stamp_chi2"], stamp_accepted=fit["stamp_accepted"])

    def photometric_scale(self) -> np.ndarray:
        """Per-pixel photometric scale sum_n a_n(x, y) S_n, shape (H, W)."""
        h, w = self.shape
        return _core.photometric_scale(self.knots, self.theta, self.component_sums, h, w)

    def photometric_scale_at(self, points) -> np.ndarray:
        """Photometric scale at (m, 2) pixel coordinates."""
        return _core.photometric_scale_at(self.knots, self.theta, self.component_sums,
                                          np.asarray(points, np.float64))

    def save(self, path: str) -> None:
        """Write the reference's .npz layout; absent per-stamp diagnostics are
        stored as empty arrays (python/delta/solution.py:60-91)."""
        def arr(a, like):
            return like if a is None else a

        np.savez(path, beta=self.beta, n_max=self.n_max, radius=self.radius, knots=self.knots,
                 theta=self.theta, n_components=self.n_components, n_spatial=self.n_spatial,
                 lam=self.lam, gcv=self.gcv, effective_dof=self.effective_dof, rss=self.rss,
                 component_sums=self.component_sums, direction=self.direction,
                 shape=np.asarray(self.shape), reduced_chi2=self.reduced_chi2,
                 n_stamps_used=self.n_stamps_used, n_stamps_rejected=self.n_stamps_rejected,
                 stamp_x=arr(self.stamp_x, np.empty(0, np.int32)),
                 stamp_y=arr(self.stamp_y, np.empty(0, np.int32)),
                 stamp_chi2=arr(self.stamp_chi2, np.empty(0, np.float64)),
                 stamp_accepted=arr(self.stamp_accepted, np.empty(0, np.uint8)))

    @classmethod
    def load(cls, path: str) -> KernelSolution:
        """Read a file written by :meth:`save` or by the reference; keys added
        after the first format version are optional (solution.py:93-119)."""
        d = np.load(path, allow_pickle=False)
        return cls(beta=float(d["beta"]), n_max=int(d["n_max"]), radius=int(d["radius"]),
                   knots=d["knots"], theta=d["theta"], n_components=int(d["n_components"]),
                   n_spatial=int(d["n_spatial"]), lam=float(d["lam"]), gcv=float(d["gcv"]),
                   effective_dof=float(d["effective_dof"]), rss=float(d["rss"]),
                   component_sums=d["component_sums"], direction=str(d["direction"]),
                   shape=(int(d["shape"][0]), int(d["shape"][1])),
                   reduced_chi2=float(d["reduced_chi2"]) if "reduced_chi2" in d else float("nan"),
                   n_stamps_used=int(d["n_stamps_used"]) if "n_stamps_used" in d else 0,
                   n_stamps_rejected=(int(d["n_stamps_rejected"]) if "n_stamps_rejected" in d
                                      else 0),
                   stamp_x=d.get("stamp_x"), stamp_y=d.get("stamp_y"),
                   stamp_chi2=d.get("stamp_chi2"), stamp_accepted=d.get("stamp_accepted"))

    def header_cards(self) -> dict:
        """Provenance key/value pairs for FITS headers (solution.py:121-134)."""
        return {"DLTBETA": self.beta, "DLTNMAX": self.n_max, "DLTNKNOT": int(self.knots.shape[0]),
                "DLTLAM": self.lam, "DLTGCV": self.gcv, "DLTEDOF": self.effective_dof,
                "DLTCHI2": self.reduced_chi2, "DLTNSTU": self.n_stamps_used,
                "DLTNSTR": self.n_stamps_rejected, "DLTCONV": self.direction}
