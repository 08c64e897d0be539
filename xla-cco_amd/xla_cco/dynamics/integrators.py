"""Single MD steps on the host with a caller-supplied force function.

src/xla_cco/dynamics/integrators.py:83-176.  ``verlet_step``: v' = v + dt/2 F;
r' = r + dt v'; F' = F(r'); v'' = v' + dt/2 F'.  ``andersen_reset`` and
``baoab_step`` thread a Philox key (xla_cco.dynamics.rng) as the reference
threads its JAX key.  The batched device loop lives in
``xla_cco.dynamics.run.simulate`` and applies the same rules with the same
draws; these host forms serve single steps with ``make_force_fn`` (GPU
forces) or any other force function.
"""

from __future__ import annotations

import math
from typing import Callable, Tuple

import numpy as np

from xla_cco._types import require_float64
from xla_cco.dynamics import rng
from xla_cco.dynamics.state import ANDERSEN_MODES

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


def andersen_reset(
    key: np.ndarray, v: np.ndarray, reset_acc: float, temperature: float, p: float, mode: str
) -> Tuple[np.ndarray, np.ndarray, float, int]:
    """Andersen resets after a step; returns (key', v', reset_acc', n_reset) (integrators.py:99-140).

    Keys: ``key', k_pick, k_draw, k_coin = split(key, 4)``.  "count": the
    accumulator gains p and fires (losing 1) when it reaches 1; "bernoulli":
    fires when uniform(k_coin) < p, the accumulator unchanged.  A firing
    redraws particle randint(k_pick, n) as sqrt(T) normals(k_draw, d).  The
    MD kernels draw the same numbers (xla_cco.dynamics.rng).
    """
    if mode not in ANDERSEN_MODES:
        raise ValueError(f"mode must be one of {ANDERSEN_MODES}, got {mode!r}")
    if not 0.0 <= p <= 1.0:
        raise ValueError(f"p must lie in [0, 1], got {p}")
    rng.require_key(key)
    require_float64(v, "v")
    if v.ndim != 2:
        raise ValueError(f"v must have shape (n, d), got {v.shape}")
    n, d = v.shape
    key, k_pick, k_draw, k_coin = rng.split(key, 4)
    acc = float(reset_acc)
    if mode == "count":
        acc = acc + p
        fire = acc >= 1.0
        if fire:
            acc = acc - 1.0
    else:
        fire = rng.uniform(k_coin) < p
    v_new = np.array(v, copy=True)
    if fire:
        v_new[rng.randint(k_pick, n)] = math.sqrt(float(temperature)) * rng.normals(k_draw, d)
    return key, v_new, acc, int(fire)


def baoab_step(
    key: np.ndarray,
    r: np.ndarray,
    v: np.ndarray,
    f: np.ndarray,
    force_fn: ForceFn,
    dt: float,
    gamma: float,
    temperature: float,
) -> Tuple[np.ndarray, np.ndarray, np.ndarray, ForceOut]:
    """One BAOAB Langevin step; returns (key', r', v', force_fn(r')) (integrators.py:143-176).

    B (half kick), A (half drift), O (v = c v + sqrt((1 - c^2) T) xi with
    c = exp(-gamma dt)), A, force, B; ``key', k_xi = split(key)`` and xi the
    first n d normals of k_xi in row-major (particle, axis) order.
    """
    _check_phase_space(r, v, f, dt)
    if not (math.isfinite(gamma) and gamma >= 0.0):
        raise ValueError(f"gamma must be finite and >= 0, got {gamma}")
    rng.require_key(key)
    c = math.exp(-gamma * dt)
    noise_var = -math.expm1(-2.0 * gamma * dt)  # 1 - c^2 without cancellation
    half = 0.5 * dt
    key, k_xi = rng.split(key, 2)
    v = v + half * f
    r = r + half * v
    xi = rng.normals(k_xi, r.size).reshape(r.shape)
    v = c * v + math.sqrt(noise_var * float(temperature)) * xi
    r = r + half * v
    out = force_fn(r)
    assert len(out) == 4, "force_fn must return (Phi, F, re, im)"
    v = v + half * out[1]
    return key, r, v, out
