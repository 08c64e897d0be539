"""Status codes and result containers (src/xla_cco/minimize/result.py:35-75).

0 running, 1 energy below tolerance, 2 step below floor, 3 evaluation budget
exhausted, 4 stagnated, 5 line-search failure, 6 gradient norm below
tolerance, 7 MD energy drift above tolerance.
"""

from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

STATUS_RUNNING: int = 0
STATUS_ENERGY_TOL: int = 1
STATUS_STEP_FLOOR: int = 2
STATUS_BUDGET: int = 3
STATUS_STAGNATED: int = 4
STATUS_LINESEARCH_FAIL: int = 5
STATUS_GRAD_TOL: int = 6
STATUS_DRIFT: int = 7


@dataclasses.dataclass
class Trajectory:
    """Recorded rows (capacity t, ``count`` valid); ``s`` rows are wrapped."""

    s: np.ndarray
    energy: np.ndarray
    step: np.ndarray
    iteration: np.ndarray
    count: np.ndarray


@dataclasses.dataclass
class RelaxResult:
    """Outcome of a relaxation; batched results carry a leading lane axis."""

    s: np.ndarray
    energy: np.ndarray
    grad_norm: np.ndarray
    n_evals: np.ndarray
    n_iters: np.ndarray
    status: np.ndarray
    trajectory: Optional[Trajectory]

    def lane(self, i: int) -> "RelaxResult":
        """Result of configuration ``i`` of a batched relax."""
        tr = self.trajectory
        sub = None
        if tr is not None:
            sub = Trajectory(tr.s[i], tr.energy[i], tr.step[i], tr.iteration[i], tr.count[i])
        return RelaxResult(
            self.s[i],
            self.energy[i],
            self.grad_norm[i],
            self.n_evals[i],
            self.n_iters[i],
            self.status[i],
            sub,
        )
