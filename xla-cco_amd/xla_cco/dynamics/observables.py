"""Density waves along MD trajectories: rho_k, centre-of-mass correction, winding.

src/xla_cco/dynamics/observables.py:37-125, one lane on the host.  The MD
kernels apply the same rules to every lane on the device each step; these
functions serve the host-side bookkeeping (initial rho, pending rows) and
users.  rho_k = sum_j exp(i k . r_j) for Cartesian r and k, stored as
(..., 2) = (Re, Im).
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from xla_cco._types import require_float64


def wave_densities(r: np.ndarray, kvec: np.ndarray, chunk_size: Optional[int] = None) -> np.ndarray:
    """rho_k (Re, Im) of every wave at Cartesian positions r: (w, 2) (observables.py:56-78)."""
    require_float64(r, "r")
    require_float64(kvec, "kvec")
    assert r.ndim == 2 and kvec.ndim == 2 and r.shape[1] == kvec.shape[1], "shapes"
    assert chunk_size is None or chunk_size >= 1, "chunk_size must be >= 1"
    step = kvec.shape[0] if chunk_size is None else int(chunk_size)
    out = np.empty((kvec.shape[0], 2))
    for lo in range(0, kvec.shape[0], max(step, 1)):
        th = kvec[lo : lo + step] @ r.T
        out[lo : lo + step, 0] = np.cos(th).sum(axis=1)
        out[lo : lo + step, 1] = np.sin(th).sum(axis=1)
    return out


def center_of_mass(r: np.ndarray) -> np.ndarray:
    """Mean position (mass 1, unwrapped)."""
    require_float64(r, "r")
    assert r.ndim == 2 and r.shape[0] >= 1, "r must have shape (n, d)"
    return np.mean(r, axis=0)


def com_correct(rho: np.ndarray, kvec: np.ndarray, r_cm: np.ndarray) -> np.ndarray:
    """rho_k exp(-i k . R_cm): translation-invariant densities (observables.py:89-98)."""
    assert rho.ndim == 2 and rho.shape[1] == 2, "rho must have shape (w, 2)"
    assert kvec.shape == (rho.shape[0], r_cm.shape[0]), "kvec must be (w, d)"
    phase = kvec @ r_cm
    c, s = np.cos(phase), np.sin(phase)
    re, im = rho[:, 0], rho[:, 1]
    return np.stack([re * c + im * s, im * c - re * s], axis=-1)


def winding_update(
    rho_prev: np.ndarray,
    rho_new: np.ndarray,
    wind_abs: np.ndarray,
    wind_net: np.ndarray,
    n_valid: np.ndarray,
    active: np.ndarray,
    n_part: int,
    phase_floor: float,
) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """(rho_prev', wind_abs', wind_net', n_valid') (observables.py:101-125).

    Delta = atan2(Im p, Re p) with p = rho_new conj(rho_old); a wave counts
    where ``active`` and S = |rho|^2 / n_part > ``phase_floor`` at both ends;
    rho_prev is replaced where ``active``.
    """
    assert rho_prev.shape == rho_new.shape == (wind_abs.shape[0], 2), "shapes"
    assert n_part >= 1, "n_part must be positive"
    o_re, o_im = rho_prev[:, 0], rho_prev[:, 1]
    n_re, n_im = rho_new[:, 0], rho_new[:, 1]
    delta = np.arctan2(n_im * o_re - n_re * o_im, n_re * o_re + n_im * o_im)
    s_old = (o_re * o_re + o_im * o_im) / n_part
    s_new = (n_re * n_re + n_im * n_im) / n_part
    ok = active & (s_old > phase_floor) & (s_new > phase_floor)
    return (
        np.where(active[:, None], rho_new, rho_prev),
        wind_abs + np.where(ok, np.abs(delta), 0.0),
        wind_net + np.where(ok, delta, 0.0),
        n_valid + ok.astype(n_valid.dtype),
    )
