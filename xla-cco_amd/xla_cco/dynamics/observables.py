"""Host forms of the MD density-wave observables (one lane).

The MD kernels evaluate these quantities for every lane on the device each step;
the host forms serve the set-up of a run (initial densities, pending rows) and
users of the reference's API (src/xla_cco/dynamics/observables.py:37-125, same
names and arguments).  They are written with complex arithmetic:

* rho_k = sum_j exp(i k . r_j) for Cartesian r and k, returned as (w, 2) = (Re, Im);
* the centre-of-mass correction multiplies rho_k by exp(-i k . R_cm), which makes
  it invariant under a rigid translation of every particle;
* the winding of wave k over a step is the angle of rho_new conj(rho_old), in
  (-pi, pi], counted only while S_k = |rho_k|^2 / n stays above a floor at both ends
  (below it the phase carries no signal).
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from xla_cco._types import require_float64


def _as_complex(rho: np.ndarray) -> np.ndarray:
    return rho[:, 0] + 1j * rho[:, 1]


def _as_pairs(z: np.ndarray) -> np.ndarray:
    return np.stack([z.real, z.imag], axis=-1)


def wave_densities(r: np.ndarray, kvec: np.ndarray, chunk_size: Optional[int] = None) -> np.ndarray:
    """rho_k of every wave k (rows of ``kvec``, (w, d)) at positions r (n, d): (w, 2).

    ``chunk_size`` bounds the number of waves whose phases are held at once.
    """
    require_float64(r, "r")
    require_float64(kvec, "kvec")
    if r.ndim != 2 or kvec.ndim != 2 or r.shape[1] != kvec.shape[1]:
        raise ValueError(f"need r (n, d) and kvec (w, d) of one d, got {r.shape} and {kvec.shape}")
    if chunk_size is not None and chunk_size < 1:
        raise ValueError(f"chunk_size must be >= 1, got {chunk_size}")
    w = kvec.shape[0]
    width = w if chunk_size is None else int(chunk_size)
    rho = np.empty(w, dtype=np.complex128)
    for lo in range(0, w, max(width, 1)):
        rho[lo: lo + width] = np.exp(1j * (kvec[lo: lo + width] @ r.T)).sum(axis=1)
    return _as_pairs(rho)


def center_of_mass(r: np.ndarray) -> np.ndarray:
    """Mean of the (unwrapped) positions, unit masses."""
    require_float64(r, "r")
    if r.ndim != 2 or r.shape[0] == 0:
        raise ValueError(f"r must be (n, d) with n >= 1, got {r.shape}")
    return r.sum(axis=0) / r.shape[0]


def com_correct(rho: np.ndarray, kvec: np.ndarray, r_cm: np.ndarray) -> np.ndarray:
    """rho_k exp(-i k . R_cm) as (w, 2)."""
    if rho.ndim != 2 or rho.shape[1] != 2 or kvec.shape != (rho.shape[0], r_cm.shape[0]):
        raise ValueError(f"need rho (w, 2), kvec (w, d), r_cm (d,), got {rho.shape}, {kvec.shape}, {r_cm.shape}")
    return _as_pairs(_as_complex(rho) * np.exp(-1j * (kvec @ r_cm)))


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
    """One step of the winding tallies: (rho_prev', wind_abs', wind_net', n_valid').

    Waves where ``active`` is set take rho_new as their new reference; of those, the
    ones with S above ``phase_floor`` before and after add |angle| to ``wind_abs``,
    the signed angle to ``wind_net`` and one to ``n_valid``.
    """
    w = wind_abs.shape[0]
    if rho_prev.shape != (w, 2) or rho_new.shape != (w, 2):
        raise ValueError(f"rho_prev and rho_new must be ({w}, 2)")
    if n_part < 1:
        raise ValueError(f"n_part must be >= 1, got {n_part}")
    old, new = _as_complex(rho_prev), _as_complex(rho_new)
    turn = np.angle(new * np.conj(old))
    s_old = (old.real**2 + old.imag**2) / n_part
    s_new = (new.real**2 + new.imag**2) / n_part
    lit = active & (s_old > phase_floor) & (s_new > phase_floor)
    return (
        np.where(active[:, None], rho_new, rho_prev),
        wind_abs + np.where(lit, np.abs(turn), 0.0),
        wind_net + np.where(lit, turn, 0.0),
        n_valid + lit.astype(n_valid.dtype),
    )
