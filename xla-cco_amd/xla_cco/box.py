"""Periodic cell: volume, fractional/Cartesian maps, force conversion, wrapping.

Rows of ``basis`` are lattice vectors (density 1: V = N).  Host-side NumPy
helpers restating src/xla_cco/box.py:34-139; none of these is on the device
hot loop except ``wrap_fractional``'s rule, which the HIP relax kernels apply
to their outputs with the same ``nextafter(1, 0)`` clamp.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from xla_cco._types import require_float64


@dataclasses.dataclass(frozen=True)
class Box:
    """Periodic box with a (d, d) float64 basis (box.py:34-109)."""

    basis: np.ndarray

    def __post_init__(self) -> None:
        basis = self.basis
        if isinstance(basis, np.ndarray):
            require_float64(basis, "basis")

    @property
    def dim(self) -> int:
        return int(self.basis.shape[0])

    @property
    def volume(self) -> float:
        """|det(basis)| (box.py:65-67)."""
        return float(abs(np.linalg.det(self.basis)))

    def reciprocal(self) -> np.ndarray:
        """Rows b_i with b_i . a_j = 2 pi delta_ij."""
        return 2.0 * np.pi * np.linalg.inv(self.basis).T

    def to_cartesian(self, s: np.ndarray) -> np.ndarray:
        require_float64(s, "s")
        return np.asarray(s) @ self.basis

    def to_fractional(self, r: np.ndarray) -> np.ndarray:
        require_float64(r, "r")
        return np.asarray(r) @ np.linalg.inv(self.basis)

    def cartesian_forces(self, grad_s: np.ndarray) -> np.ndarray:
        """F = -grad_s @ inv(B)^T (box.py:98-109)."""
        require_float64(grad_s, "grad_s")
        return -np.asarray(grad_s) @ np.linalg.inv(self.basis).T


def cubic_box(n_particles: int, dim: int) -> Box:
    """Cubic box of side N^(1/d), V = N (box.py:111-124)."""
    assert n_particles >= 1, "n_particles must be >= 1"
    assert dim in (1, 2, 3), "dim must be 1, 2 or 3"
    side = float(n_particles) ** (1.0 / dim)
    return Box(basis=np.asarray(side * np.eye(dim), dtype=np.float64))


_BELOW_ONE = float(np.nextafter(1.0, 0.0))


def wrap_fractional(s: np.ndarray) -> np.ndarray:
    """Wrap into [0, 1) with the nextafter(1, 0) clamp (box.py:127-139)."""
    require_float64(s, "s")
    s = np.asarray(s)
    return np.minimum(s - np.floor(s), _BELOW_ONE)
