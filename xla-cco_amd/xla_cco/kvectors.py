"""Host construction of the stealthy half-space k-set (NumPy, bit-exact).

The enumeration must agree bit for bit with src/xla_cco/kvectors.py:60-101,
because the success masks and the k-count M depend on strict float64 shell
tests at the boundary.  The arithmetic therefore keeps the reference's
operation order: the per-axis bound ``ceil(K |a_i| / 2 pi) + 1``, the
Cartesian ``k_j = sum_i n_i recip_ij`` summed in index order (no BLAS), and
``|k|^2`` summed component by component.  Rows are sorted by (k_sq, n
lexicographic) exactly as ``np.lexsort`` with k_sq as the primary key.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass(frozen=True)
class KSet:
    """Half-space k-vectors with K1^2 < |k|^2 < K2^2 (kvectors.py:33-57).

    n (M, d) int64, k (M, d) float64 Cartesian, k_sq (M,), weights (M,) = 1.
    """

    n: np.ndarray
    k: np.ndarray
    k_sq: np.ndarray
    weights: np.ndarray
    k_inner: float
    k_outer: float


def _validate_inputs(basis: np.ndarray, k_inner: float, k_outer: float) -> None:
    """ValueError for a malformed basis or cutoff pair (kvectors.py:104-116)."""
    if basis.ndim != 2 or basis.shape[0] != basis.shape[1] or basis.shape[0] < 1:
        raise ValueError(f"basis must be square (d, d), got {basis.shape}")
    if not (
        np.all(np.isfinite(basis)) and math.isfinite(k_inner) and math.isfinite(k_outer)
    ):
        raise ValueError("basis and cutoffs must be finite")
    if not 0.0 <= k_inner < k_outer:
        raise ValueError(f"need 0 <= k_inner < k_outer, got {k_inner}, {k_outer}")
    det = float(np.linalg.det(basis))
    if not abs(det) > 0.0 or not math.isfinite(det):
        raise ValueError("basis must have positive |det|")


def _halfspace_mask(n: np.ndarray) -> np.ndarray:
    """Rows whose first nonzero coefficient is positive (kvectors.py:119-126)."""
    keep = np.zeros(n.shape[0], dtype=bool)
    open_rows = np.ones(n.shape[0], dtype=bool)
    for axis in range(n.shape[1]):
        col = n[:, axis]
        keep |= open_rows & (col > 0)
        open_rows &= col == 0
    return keep


def enumerate_kvectors(basis: np.ndarray, k_inner: float, k_outer: float) -> KSet:
    """Enumerate the half-space k-set of a periodic box (kvectors.py:60-101)."""
    basis = np.asarray(basis, dtype=np.float64)
    _validate_inputs(basis, k_inner, k_outer)
    dim = basis.shape[0]
    recip = 2.0 * math.pi * np.linalg.inv(basis).T
    bounds = [
        int(math.ceil(k_outer * float(np.linalg.norm(basis[i])) / (2.0 * math.pi))) + 1
        for i in range(dim)
    ]
    grids = np.meshgrid(
        *[np.arange(-b, b + 1, dtype=np.int64) for b in bounds], indexing="ij"
    )
    n = np.stack(grids, axis=-1).reshape(-1, dim)
    n = n[_halfspace_mask(n)]
    k = np.zeros((n.shape[0], dim))
    for j in range(dim):
        for i in range(dim):  # index order, one rounding per term
            k[:, j] = k[:, j] + n[:, i] * recip[i, j]
    k_sq = np.zeros(n.shape[0])
    for i in range(dim):
        k_sq = k_sq + k[:, i] * k[:, i]
    inside = (k_inner * k_inner < k_sq) & (k_sq < k_outer * k_outer)
    n, k, k_sq = n[inside], k[inside], k_sq[inside]
    keys = tuple(n[:, i] for i in range(dim - 1, -1, -1)) + (k_sq,)
    order = np.lexsort(keys)
    n, k, k_sq = n[order], k[order], k_sq[order]
    return KSet(
        n=n,
        k=k,
        k_sq=k_sq,
        weights=np.ones(n.shape[0]),
        k_inner=float(k_inner),
        k_outer=float(k_outer),
    )


# V_d(1) as the C++ HyperSphere_C literals round them (kvectors.py:131-136).
_UNIT_BALL_LITERALS = {
    1: 2.0,
    2: math.pi,
    3: 4.1887902047863909846,
    4: 4.9348022005446793094,
}


def unit_ball_volume(dim: int) -> float:
    """Volume of the unit d-ball, bit-equal to the C++ literals for d <= 4."""
    assert dim >= 1, "dim must be >= 1"
    if dim in _UNIT_BALL_LITERALS:
        return _UNIT_BALL_LITERALS[dim]
    return float(math.pi ** (dim / 2.0) / math.gamma(dim / 2.0 + 1.0))


def k_outer_from_ka(ka: float, phi: float, dim: int) -> float:
    """K2 = Ka / a with a = (phi / V_d(1))^(1/d) (kvectors.py:152-156)."""
    assert phi > 0.0 and dim >= 1, "need phi > 0 and dim >= 1"
    a = (phi / unit_ball_volume(dim)) ** (1.0 / dim)
    return float(ka / a)


def ka_from_chi(chi: float, phi: float, dim: int) -> float:
    """Ka = (2^(d+1) d chi Gamma(d/2+1)^2 phi)^(1/d) (kvectors.py:159-163)."""
    assert chi >= 0.0 and phi > 0.0 and dim >= 1, "need chi >= 0, phi > 0, dim >= 1"
    arg = 2.0 ** (dim + 1) * dim * chi * math.gamma(dim / 2.0 + 1.0) ** 2 * phi
    return float(arg ** (1.0 / dim))


def chi_from_count(m: int, n_particles: int, dim: int) -> float:
    """chi = M / (d (N - 1)) (kvectors.py:166-169)."""
    assert n_particles >= 2 and dim >= 1, "need N >= 2 and dim >= 1"
    return float(m / (dim * (n_particles - 1)))
