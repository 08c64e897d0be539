"""Kinetic energy and Maxwell-Boltzmann velocities (mass 1, k_B T = T).

src/xla_cco/dynamics/thermal.py:73-106.  The reference draws with JAX's
threefry generator; here draws come from a NumPy ``Generator`` (the
distribution is the same, the streams are not).
"""

from __future__ import annotations

from typing import Union

import numpy as np

from xla_cco._types import require_float64


def maxwell_boltzmann(
    rng: Union[np.random.Generator, int], n: int, d: int, temperature: float, remove_com: bool
) -> np.ndarray:
    """Velocities (n, d) ~ N(0, T) per component; zero total momentum if asked."""
    if n < 1 or d < 1:
        raise ValueError(f"n and d must be positive, got n={n}, d={d}")
    gen = rng if isinstance(rng, np.random.Generator) else np.random.default_rng(rng)
    v = np.sqrt(float(temperature)) * gen.standard_normal((n, d))
    if remove_com:
        v = v - np.mean(v, axis=0, keepdims=True)
    return v


def kinetic_energy(v: np.ndarray) -> float:
    """sum_j |v_j|^2 / 2 (mass 1)."""
    require_float64(v, "v")
    return float(0.5 * np.sum(v * v))


def kinetic_temperature(v: np.ndarray, dof: int) -> float:
    """2 KE / dof."""
    return 2.0 * kinetic_energy(v) / dof
