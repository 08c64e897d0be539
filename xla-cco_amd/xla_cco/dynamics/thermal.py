"""Kinetic energy, Maxwell-Boltzmann velocities and harmonic draws (mass 1, k_B T = T).

src/xla_cco/dynamics/thermal.py:47-153.  Velocities come from a Philox key
(xla_cco.dynamics.rng, the functional key threading of the reference's JAX
keys; the distribution is the reference's, the stream is not).
``harmonic_sample`` draws with the caller's NumPy generator, as the
reference does.
"""

from __future__ import annotations

import math

import numpy as np

from xla_cco._types import require_float64
from xla_cco.dynamics import rng as _rng

require_key = _rng.require_key


def maxwell_boltzmann(key: np.ndarray, n: int, d: int, temperature: float, remove_com: bool) -> np.ndarray:
    """Velocities (n, d) ~ N(0, T) per component; zero total momentum if asked (thermal.py:73-97).

    The components are sqrt(T) normals(key, n d) in row-major order; with
    ``remove_com`` the mean velocity is subtracted (d (n - 1) degrees of
    freedom remain).  Raises ValueError unless n, d >= 1 and TypeError on a
    key that is not uint32 of shape (2,).
    """
    if n < 1 or d < 1:
        raise ValueError(f"n and d must be positive, got n={n}, d={d}")
    require_key(key)
    v = math.sqrt(float(temperature)) * _rng.normals(np.asarray(key), n * d).reshape(n, d)
    if remove_com:
        v = v - np.mean(v, axis=0, keepdims=True)
    return v


def kinetic_energy(v: np.ndarray) -> float:
    """sum_j |v_j|^2 / 2 (mass 1)."""
    require_float64(v, "v")
    return float(0.5 * np.sum(v * v))


def kinetic_temperature(v: np.ndarray, dof: int) -> float:
    """Instantaneous temperature 2 KE / dof (dof = d (n - 1) without COM motion)."""
    if dof < 1:
        raise ValueError(f"dof must be positive, got {dof}")
    return 2.0 * kinetic_energy(v) / dof


def cartesian_hessian(energy, s: np.ndarray, basis: np.ndarray) -> np.ndarray:  # type: ignore[no-untyped-def]
    """Cartesian Hessian d^2 Phi / dr dr at r = s B, row-major (n d, n d).

    thermal.py:47-65: with s = r inv(B), H_r = (I (x) inv(B)) H_s (I (x) inv(B))^T
    from ``energy.hessian`` (on the GPU for a ``StealthyEnergy``), symmetrized
    against roundoff.
    """
    s = np.asarray(s, dtype=np.float64)
    n, d = s.shape
    basis = np.asarray(basis, dtype=np.float64)
    if basis.shape != (d, d) or not np.isfinite(basis).all():
        raise ValueError(f"basis must be a finite ({d}, {d}) array")
    inv = np.linalg.inv(basis)
    h_s = np.asarray(energy.hessian(s))
    assert h_s.shape == (n * d, n * d), f"hessian shape {h_s.shape}"
    h_r = np.einsum("ac,jcle,be->jalb", inv, h_s.reshape(n, d, n, d), inv).reshape(n * d, n * d)
    assert np.isfinite(h_r).all(), "non-finite Cartesian Hessian"
    return np.asarray(0.5 * (h_r + h_r.T), dtype=np.float64)


def harmonic_sample(  # type: ignore[no-untyped-def]
    rng: np.random.Generator,
    energy,
    s_gs: np.ndarray,
    basis: np.ndarray,
    temperature: float,
    zero_rel: float = 1e-9,
) -> np.ndarray:
    """Cartesian positions (n, d): ground state plus a canonical harmonic draw.

    thermal.py:118-153.  Modes of the Cartesian Hessian at ``s_gs`` with
    lambda > zero_rel * lambda_max get amplitude N(0, T / lambda), the others
    none; ``rng.standard_normal(k)`` is called once for the k kept modes
    (ascending eigenvalue order).  Returns s_gs @ B + displacement (unwrapped).
    Raises ValueError on a negative or non-finite temperature, a basis that is
    not (d, d), or a Hessian without positive eigenvalues; TypeError when rng
    is not a ``numpy.random.Generator``.
    """
    if not (math.isfinite(temperature) and temperature >= 0.0):
        raise ValueError(f"temperature must be finite and >= 0, got {temperature}")
    if not isinstance(rng, np.random.Generator):
        raise TypeError("rng must be a numpy.random.Generator")
    s = np.asarray(s_gs, dtype=np.float64)
    b = np.asarray(basis, dtype=np.float64)
    if s.ndim != 2:
        raise ValueError(f"s_gs must have shape (n, d), got {s.shape}")
    lam, vec = np.linalg.eigh(cartesian_hessian(energy, s, b))
    if not lam[-1] > 0.0:
        raise ValueError("Hessian has no positive eigenvalue")
    keep = lam > zero_rel * lam[-1]
    amp = np.sqrt(temperature / lam[keep]) * rng.standard_normal(int(keep.sum()))
    out = s @ b + (vec[:, keep] @ amp).reshape(s.shape)
    assert out.shape == s.shape and np.isfinite(out).all(), "bad harmonic draw"
    return np.asarray(out, dtype=np.float64)
