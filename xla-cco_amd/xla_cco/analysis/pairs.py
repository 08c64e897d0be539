"""Pair statistics: minimum-image pair-distance histogram, g2 and NN distances.

The distance work (all pairs over the 3^d neighbouring images, ds - round(ds)
on fractional differences) runs on the GPU (``cco_pair_stats``); g2's shell
normalisation is host arithmetic (src/xla_cco/analysis/pairs.py:42-131).
Positions must be wrapped into [0, 1) and the cell reduced, as in the reference.
"""

from __future__ import annotations

import numpy as np

from xla_cco import _native
from xla_cco._types import require_float64
from xla_cco.box import Box
from xla_cco.kvectors import unit_ball_volume


def _check_edges(edges: np.ndarray) -> None:
    edges = np.asarray(edges)
    if edges.ndim != 1 or edges.shape[0] < 2:
        raise ValueError("edges must be 1-D with at least 2 entries")
    if not np.all(np.isfinite(edges)) or edges[0] < 0.0:
        raise ValueError("edges must be finite and non-negative")
    if not np.all(np.diff(edges) > 0.0):
        raise ValueError("edges must be strictly increasing")


def _pair_stats(s: np.ndarray, box: Box, edges: np.ndarray, want_nn: bool, device: int = 0):
    s = np.ascontiguousarray(s)
    b, n, d = s.shape
    e = np.ascontiguousarray(edges, dtype=np.float64)
    counts = np.zeros(e.size - 1, dtype=np.int64)
    nn = np.empty((b, n)) if want_nn else None
    ctx = _native.context(device)
    basis = np.ascontiguousarray(np.asarray(box.basis, dtype=np.float64))
    _native.check(_native.lib().cco_pair_stats(
        ctx.handle, _native.dptr(s), b, n, d, _native.dptr(basis), _native.dptr(e), int(e.size),
        _native.i64ptr(counts), _native.dptr(nn) if nn is not None else None))
    return counts, nn


def pair_distance_histogram(s: np.ndarray, box: Box, edges: np.ndarray) -> np.ndarray:
    """Pair-distance counts summed over the batch (b, n, d), each pair once."""
    require_float64(s, "s")
    _check_edges(edges)
    s = np.asarray(s)
    assert s.ndim == 3 and s.shape[2] == box.dim, "s must have shape (b, n, d)"
    return _pair_stats(s, box, edges, False)[0]


def nearest_neighbor_distances(s: np.ndarray, box: Box) -> np.ndarray:
    """Minimum-image nearest-neighbour distance of every particle of one configuration."""
    require_float64(s, "s")
    s = np.asarray(s)
    assert s.ndim == 2 and s.shape[1] == box.dim, "s must have shape (n, d)"
    _, nn = _pair_stats(s[None], box, np.array([0.0, 1.0]), True)
    return nn[0]


def nearest_neighbor_distances_batch(s: np.ndarray, box: Box) -> np.ndarray:
    require_float64(s, "s")
    _, nn = _pair_stats(np.asarray(s), box, np.array([0.0, 1.0]), True)
    return nn


def g2_from_histogram(
    counts: np.ndarray, edges: np.ndarray, n_particles: int, n_configs: int, volume: float, dim: int
) -> np.ndarray:
    """g2 = counts / [(V_d r_hi^d - V_d r_lo^d) * rho N n_configs / 2], rho = N / V."""
    _check_edges(edges)
    if counts.shape != (edges.shape[0] - 1,):
        raise ValueError("counts must have length len(edges) - 1")
    v_d = unit_ball_volume(dim)
    shell = v_d * edges[1:] ** dim - v_d * edges[:-1] ** dim
    rho = n_particles / volume
    return np.asarray(counts, dtype=np.float64) / (shell * 0.5 * rho * n_particles * n_configs)
