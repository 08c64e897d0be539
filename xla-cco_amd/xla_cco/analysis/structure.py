"""Structure factor: S(n) on a k-set and the binned isotropic S(|k|).

The isotropic k-set follows the C++ ``IsotropicStructureFactor`` as the
reference restates it (src/xla_cco/analysis/structure.py:111-144): primitive
directions of the half-space shell |k| < k_dir_max (gcd of the coefficients
1, in k-set order), each followed by its multiples i k with |i k|^2 <= k_max^2
(|.|^2 summed in index order).  The per-k sums of S and S^2 over the batch
run on the GPU (``cco_sk_moments``); binning by floor(|k| / bin_width) with
the C++ error columns is host NumPy (:176-201).
"""

from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from xla_cco import _native
from xla_cco._types import require_float64
from xla_cco.box import Box
from xla_cco.energy import collective_density
from xla_cco.kvectors import enumerate_kvectors


def structure_factor(s: np.ndarray, n_vectors: np.ndarray) -> np.ndarray:
    """S(n) = |rho(n)|^2 / N for integer vectors (structure.py:44-52)."""
    require_float64(s, "s")
    require_float64(n_vectors, "n_vectors")
    re, im = collective_density(s, n_vectors)
    return (re * re + im * im) / np.asarray(s).shape[0]


def isotropic_directions(
    basis: np.ndarray, k_dir_max: float, k_max: float
) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(primitive directions (D, d) int64, multiples per direction (D,), |k|^2 (M,))."""
    assert basis.ndim == 2 and basis.shape[0] == basis.shape[1], "square basis"
    if basis.shape[0] == 1:  # d = 1: k = i b regardless of k_dir_max
        b = 2.0 * math.pi / basis[0, 0]
        n_prim = np.ones((1, 1), dtype=np.int64)
        k_prim = np.array([[b]])
        ksq_prim = k_prim[:, 0] ** 2
    else:
        ks = enumerate_kvectors(basis, 0.0, k_dir_max)
        prim = np.gcd.reduce(np.abs(ks.n), axis=1) == 1
        n_prim, k_prim, ksq_prim = ks.n[prim], ks.k[prim], ks.k_sq[prim]
    limit = k_max * k_max
    counts, ksq_parts = [], []
    for k_row, q in zip(k_prim, ksq_prim):
        mult = np.arange(1, int(math.sqrt(limit / q)) + 3, dtype=np.float64)
        kk = mult[:, None] * k_row[None, :]
        m2 = np.zeros(mult.shape[0])
        for c in range(kk.shape[1]):
            m2 = m2 + kk[:, c] * kk[:, c]
        ok = m2 <= limit
        # the multiples that pass form a prefix 1..L (|i k| grows with i)
        counts.append(int(ok.sum()))
        ksq_parts.append(m2[ok])
    if not counts:
        raise ValueError("no reciprocal-lattice vector below k_dir_max")
    return n_prim.astype(np.int64), np.array(counts, dtype=np.int32), np.concatenate(ksq_parts)


def sk_moments(s: np.ndarray, n_dir: np.ndarray, n_mult: np.ndarray, device: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """Configuration means of S and S^2 per k (direction-major, multiples contiguous)."""
    s = np.ascontiguousarray(s)
    b, n, d = s.shape
    m = int(n_mult.sum())
    tot = np.empty(m)
    tot2 = np.empty(m)
    ctx = _native.context(device)
    nd = np.ascontiguousarray(n_dir, dtype=np.int64)
    nm = np.ascontiguousarray(n_mult, dtype=np.int32)
    _native.check(_native.lib().cco_sk_moments(
        ctx.handle, _native.dptr(s), b, n, d, _native.i64ptr(nd),
        nm.ctypes.data_as(_native._PI32), int(nd.shape[0]), _native.dptr(tot), _native.dptr(tot2)))
    return tot / b, tot2 / b


def _bin_moments(
    k_mod: np.ndarray, s_mean: np.ndarray, s2_mean: np.ndarray, n_cfg: int, bin_width: float
) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """C++ binning: sums over configs and k, empty bins dropped (structure.py:176-201)."""
    idx = np.floor(k_mod / bin_width).astype(np.int64)
    size = int(idx.max()) + 1
    cnt = np.bincount(idx, minlength=size) * float(n_cfg)
    keep = cnt != 0
    cnt = cnt[keep]

    def tot(values: np.ndarray) -> np.ndarray:
        return np.bincount(idx, weights=values, minlength=size)[keep] * n_cfg

    mean_k = tot(k_mod) / cnt
    mean_s = tot(s_mean) / cnt
    var_k = (tot(k_mod * k_mod) / cnt - mean_k**2) / cnt
    err_k = np.sqrt(np.where(var_k > 0.0, var_k, 0.0))
    with np.errstate(invalid="ignore"):
        err_s = np.sqrt((tot(s2_mean) / cnt - mean_s**2) / cnt)
    return mean_k, mean_s, err_k, err_s


def isotropic_structure_factor(
    s: np.ndarray,
    box: Box,
    k_dir_max: float = 5.0,
    k_max: float = 400.0,
    bin_width: float = 0.01,
) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Binned isotropic S(|k|) of a batch (b, n, d): mean |k|, mean S, their errors."""
    require_float64(s, "s")
    s = np.asarray(s)
    if s.ndim != 3 or s.shape[2] != box.dim:
        raise ValueError(f"s must have shape (b, n, {box.dim}), got {s.shape}")
    if not (0.0 < k_dir_max < k_max and bin_width > 0.0):
        raise ValueError("need 0 < k_dir_max < k_max and bin_width > 0")
    n_dir, n_mult, ksq = isotropic_directions(np.asarray(box.basis), k_dir_max, k_max)
    s_mean, s2_mean = sk_moments(s, n_dir, n_mult)
    return _bin_moments(np.sqrt(ksq), s_mean, s2_mean, int(s.shape[0]), bin_width)
