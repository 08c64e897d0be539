"""Three-pass L-BFGS settings and single-configuration driver.

Pass p minimises Phi(x / R_p) over x = R_p s with R = (1e4, 1, 1e-2)
(src/xla_cco/minimize/lbfgs.py:55-350).  Each iteration: memory update and
two-loop recursion over ``memory`` (s, y) pairs, direction u = -H g, a zoom
line search (strong Wolfe with Hager-Zhang approximate decrease, initial
step 1, at most ``max_linesearch_steps`` evaluations), then the reference's
pass logic.  The iteration runs on the GPU (lbfgs_advance_kernel).
"""

from __future__ import annotations

import dataclasses
from typing import Tuple

import numpy as np

from xla_cco._types import require_float64
from xla_cco.energy import AbstractEnergy
from xla_cco.minimize.result import RelaxResult

DEFAULT_CHUNK_ITERS: int = 2_000


@dataclasses.dataclass(frozen=True)
class LBFGSSettings:
    """L-BFGS settings (lbfgs.py:55-67), same fields and defaults."""

    memory: int = 100
    rescales: Tuple[float, ...] = (1e4, 1.0, 1e-2)
    max_evals_per_pass: int = 500_000
    energy_tol: float = 0.0
    grad_tol: float = 0.0
    stagnation_patience: int = 50
    max_linesearch_steps: int = 30
    record_every: int = 0
    max_records: int = 0
    # split-precision evaluations while the energy is above precision_switch
    # (an extension: "fp64" keeps the reference's arithmetic)
    precision: str = "fp64"
    precision_switch: float = 1e-8
    # (s, y) pairs of split-precision lanes: "fp64", or "fp32" (half the history
    # bytes per two-loop on the register-resident advance, widened at the switch)
    split_history: str = "fp64"


def update_stagnation(f_new: float, f_old: float, stag: int) -> int:
    """Consecutive non-decrease counter: reset on strict decrease (lbfgs.py:228-232)."""
    return 0 if f_new < f_old else stag + 1


def max_rounds(settings: LBFGSSettings, chunk_iters: int) -> int:
    """Upper bound on host rounds (lbfgs.py:344-350)."""
    assert chunk_iters >= 1, "chunk_iters must be positive"
    total = len(settings.rescales) * (
        settings.max_evals_per_pass + settings.max_linesearch_steps + 1
    )
    return total // chunk_iters + 2


def lbfgs(energy: AbstractEnergy, s0: np.ndarray, settings: LBFGSSettings) -> RelaxResult:
    """Relax one configuration with the three-pass schedule (lbfgs.py:353-390)."""
    from xla_cco.minimize import relax

    require_float64(s0, "s0")
    s0 = np.asarray(s0)
    assert s0.ndim == 2, "s0 must have shape (n, d)"
    return relax(energy, s0[None], settings, chunk_iters=DEFAULT_CHUNK_ITERS).lane(0)
