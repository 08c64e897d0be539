"""One velocity-Verlet step with a caller-supplied force function.

src/xla_cco/dynamics/integrators.py:83-96: v' = v + dt/2 F; r' = r + dt v';
F' = F(r'); v'' = v' + dt/2 F'.  The batched device loop lives in
``xla_cco.dynamics.run.simulate``; this host form serves single steps with
``make_force_fn`` (GPU forces) or any other force function.
"""

from __future__ import annotations

import math
from typing import Callable, Tuple

import numpy as np

from xla_cco._types import require_float64

ForceOut = Tuple[float, np.ndarray, np.ndarray, np.ndarray]
ForceFn = Callable[[np.ndarray], ForceOut]


def _check_phase_space(r: np.ndarray, v: np.ndarray, f: np.ndarray, dt: float) -> None:
    for x, name in ((r, "r"), (v, "v"), (f, "f")):
        require_float64(x, name)
    if r.ndim != 2 or v.shape != r.shape or f.shape != r.shape:
        raise ValueError(f"r, v, f must share an (n, d) shape: {r.shape}, {v.shape}, {f.shape}")
    if not (math.isfinite(dt) and dt > 0.0):
        raise ValueError(f"dt must be finite and positive, got {dt}")


def verlet_step(
    r: np.ndarray, v: np.ndarray, f: np.ndarray, force_fn: ForceFn, dt: float
) -> Tuple[np.ndarray, np.ndarray, ForceOut]:
    """One velocity-Verlet step; returns (r', v', force_fn(r'))."""
    _check_phase_space(r, v, f, dt)
    half = 0.5 * dt
    v_half = v + half * f
    r_new = r + dt * v_half
    out = force_fn(r_new)
    assert len(out) == 4, "force_fn must return (Phi, F, re, im)"
    v_new = v_half + half * out[1]
    return r_new, v_new, out
