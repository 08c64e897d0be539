"""The stealthy half-space k-set of a periodic box (host NumPy).

The set must equal the reference's bit for bit (src/xla_cco/kvectors.py:60-101):
the k-count M and the success masks hinge on strict float64 shell tests at the
boundary.  What the arithmetic has to keep, and why:

* the per-axis search bound ceil(K |a_i| / 2 pi) + 1 (|n_i| = |k . a_i| / 2 pi);
* Cartesian k_j = n_1 r_1j + n_2 r_2j + ... accumulated in axis order with one
  rounding per term (no BLAS reassociation), r = 2 pi inv(B)^T;
* |k|^2 accumulated component by component in the same way;
* the order: by |k|^2, ties by n lexicographically.

The candidate lattice points are generated here as offsets of an index grid and
the half-space (first nonzero coefficient positive) is selected through the sign
of each row's leading nonzero entry; neither choice changes a bit of the result.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass(frozen=True)
class KSet:
    """Half-space k-vectors with K1^2 < |k|^2 < K2^2 (the reference's KSet, kvectors.py:33-57).

    n (M, d) int64 integer coefficients, k (M, d) Cartesian, k_sq (M,), weights (M,) = 1.
    """

    n: np.ndarray
    k: np.ndarray
    k_sq: np.ndarray
    weights: np.ndarray
    k_inner: float
    k_outer: float


def _check_box(basis: np.ndarray, k_inner: float, k_outer: float) -> None:
    """ValueError unless basis is a finite, invertible (d, d) matrix and 0 <= K1 < K2 < inf."""
    if basis.ndim != 2 or basis.shape[0] < 1 or basis.shape[0] != basis.shape[1]:
        raise ValueError(f"the basis must be a square (d, d) matrix, not shape {basis.shape}")
    finite = bool(np.isfinite(basis).all()) and math.isfinite(k_inner) and math.isfinite(k_outer)
    if not finite:
        raise ValueError("the basis and both cutoffs must be finite numbers")
    if not (0.0 <= k_inner < k_outer):
        raise ValueError(f"the cutoffs must satisfy 0 <= k_inner < k_outer (got {k_inner}, {k_outer})")
    volume = abs(float(np.linalg.det(basis)))
    if not (volume > 0.0 and math.isfinite(volume)):
        raise ValueError(f"the basis is singular (|det| = {volume})")


def _candidates(bounds: list) -> np.ndarray:
    """Every integer vector with |n_i| <= bounds[i], one row each (any order)."""
    span = [2 * b + 1 for b in bounds]
    grid = np.indices(span, dtype=np.int64).reshape(len(span), -1).T
    return grid - np.asarray(bounds, dtype=np.int64)


def _upper_half(n: np.ndarray) -> np.ndarray:
    """Rows whose leading nonzero coefficient is positive (one of each pair +-n; 0 excluded)."""
    lead = np.argmax(n != 0, axis=1)
    return n[np.arange(n.shape[0]), lead] > 0


def _accumulate(rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
    """sum_i rows[:, i] * cols[i] accumulated in i order, one rounding per term."""
    acc = 0.0 + rows[:, 0:1] * cols[0]
    for i in range(1, rows.shape[1]):
        acc = acc + rows[:, i:i + 1] * cols[i]
    return acc


def enumerate_kvectors(basis: np.ndarray, k_inner: float, k_outer: float) -> KSet:
    """Half-space k-set of the box ``basis`` (rows = lattice vectors) with K1 < |k| < K2."""
    basis = np.asarray(basis, dtype=np.float64)
    _check_box(basis, k_inner, k_outer)
    dim = basis.shape[0]
    recip = 2.0 * math.pi * np.linalg.inv(basis).T
    bounds = [int(math.ceil(k_outer * float(np.linalg.norm(row)) / (2.0 * math.pi))) + 1 for row in basis]
    n = _candidates(bounds)
    n = n[_upper_half(n)]
    k = _accumulate(n, recip)
    k_sq = _accumulate(k * k, np.ones(dim))[:, 0]
    shell = (k_sq > k_inner * k_inner) & (k_sq < k_outer * k_outer)
    n, k, k_sq = n[shell], k[shell], k_sq[shell]
    # |k|^2 first, then n_1, n_2, ... (np.lexsort takes its primary key last)
    order = np.lexsort([n[:, i] for i in reversed(range(dim))] + [k_sq])
    return KSet(
        n=n[order],
        k=k[order],
        k_sq=k_sq[order],
        weights=np.ones(order.size),
        k_inner=float(k_inner),
        k_outer=float(k_outer),
    )


# V_d(1) for d <= 4 as literals, rounded to float64 as the C++ HyperSphere_C
# table rounds them (the reference keeps the same values, kvectors.py:131-136).
_BALL_VOLUME = {1: 2.0, 2: math.pi, 3: 4.1887902047863909846, 4: 4.9348022005446793094}


def unit_ball_volume(dim: int) -> float:
    """Volume of the unit d-ball (the C++ literals for d <= 4)."""
    if dim < 1:
        raise ValueError(f"dimension must be >= 1, got {dim}")
    return _BALL_VOLUME.get(dim, float(math.pi ** (0.5 * dim) / math.gamma(0.5 * dim + 1.0)))


def k_outer_from_ka(ka: float, phi: float, dim: int) -> float:
    """K2 = Ka / a for the particle radius a with phi = V_d(1) a^d (kvectors.py:152-156)."""
    if not (phi > 0.0 and dim >= 1):
        raise ValueError(f"need phi > 0 and dim >= 1, got phi={phi}, dim={dim}")
    radius = (phi / unit_ball_volume(dim)) ** (1.0 / dim)
    return float(ka / radius)


def ka_from_chi(chi: float, phi: float, dim: int) -> float:
    """Ka of constraint ratio chi: (2^(d+1) d chi Gamma(d/2+1)^2 phi)^(1/d) (kvectors.py:159-163)."""
    if not (chi >= 0.0 and phi > 0.0 and dim >= 1):
        raise ValueError(f"need chi >= 0, phi > 0, dim >= 1, got {chi}, {phi}, {dim}")
    return float((2.0 ** (dim + 1) * dim * chi * math.gamma(dim / 2.0 + 1.0) ** 2 * phi) ** (1.0 / dim))


def chi_from_count(m: int, n_particles: int, dim: int) -> float:
    """chi = M / (d (N - 1)): constrained over free degrees of freedom (kvectors.py:166-169)."""
    if not (n_particles >= 2 and dim >= 1):
        raise ValueError(f"need N >= 2 and dim >= 1, got N={n_particles}, dim={dim}")
    return float(m / (dim * (n_particles - 1)))
