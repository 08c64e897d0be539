"""Adaptive normalised gradient descent settings and single-configuration driver.

The C++ rule (src/xla_cco/minimize/descent.py:52-197): g_hat = g/|g| (0 when
|g| = 0), x_n = s - g_hat * step; reject when |(f - f2) - step |g|| /
(step |g|) > rel_tol (NaN accepts); step *= shrink on reject, grow on accept.
Status precedence energy tol > zero gradient > step floor > budget;
``n_evals`` counts 2 per iteration and the budget is ``max_evals // 2``
iterations.  The iteration runs on the GPU (descent_advance_kernel).
"""

from __future__ import annotations

import dataclasses

import numpy as np

from xla_cco._types import require_float64
from xla_cco.energy import AbstractEnergy
from xla_cco.minimize.result import RelaxResult

DEFAULT_CHUNK_ITERS: int = 2_000


@dataclasses.dataclass(frozen=True)
class DescentSettings:
    """Adaptive-descent settings (descent.py:52-63), same fields and defaults."""

    step_init: float = 1e-3
    step_min: float = 1e-15
    shrink: float = 0.5
    grow: float = 1.0
    rel_tol: float = 0.01
    max_evals: int = 10_000
    energy_tol: float = 0.0
    record_every: int = 0
    max_records: int = 0


def max_rounds(settings: DescentSettings, chunk_iters: int) -> int:
    """Upper bound on host rounds (descent.py:212-215)."""
    assert chunk_iters >= 1, "chunk_iters must be positive"
    return (settings.max_evals // 2) // chunk_iters + 2


def adaptive_descent(
    energy: AbstractEnergy, s0: np.ndarray, settings: DescentSettings
) -> RelaxResult:
    """Relax one configuration by adaptive descent (descent.py:218-255)."""
    from xla_cco.minimize import relax

    require_float64(s0, "s0")
    s0 = np.asarray(s0)
    assert s0.ndim == 2, "s0 must have shape (n, d)"
    return relax(energy, s0[None], settings, chunk_iters=DEFAULT_CHUNK_ITERS).lane(0)
