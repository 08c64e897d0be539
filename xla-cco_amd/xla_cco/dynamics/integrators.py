"""Single MD steps on the host, with a caller-supplied force function.

The steps are compositions of three maps on (r, v) with unit masses:

* kick  B(h): v <- v + h F
* drift A(h): r <- r + h v
* Ornstein-Uhlenbeck O: v <- c v + sqrt((1 - c^2) T) xi, c = exp(-gamma dt)

velocity Verlet is B(dt/2) A(dt) [force] B(dt/2); BAOAB is B(dt/2) A(dt/2) O A(dt/2)
[force] B(dt/2).  The reference's names and arguments are kept
(src/xla_cco/dynamics/integrators.py:83-176).  Random draws come from the Philox
streams of xla_cco.dynamics.rng in the order the MD kernels use, so a host step and
a device step of one lane draw the same numbers; the reference threads a JAX key the
same way.  The batched device loop is ``xla_cco.dynamics.run.simulate``.
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


def _kick(v: np.ndarray, f: np.ndarray, h: float) -> np.ndarray:
    return v + h * f


def _drift(r: np.ndarray, v: np.ndarray, h: float) -> np.ndarray:
    return r + h * v


def _forces(force_fn: ForceFn, r: np.ndarray) -> ForceOut:
    out = force_fn(r)
    if len(out) != 4:
        raise ValueError("a force function returns (Phi, F, re, im)")
    return out


def _phase_space(r: np.ndarray, v: np.ndarray, f: np.ndarray, dt: float) -> None:
    """float64 (n, d) positions, velocities and forces of one shape; finite dt > 0."""
    require_float64(r, "r")
    require_float64(v, "v")
    require_float64(f, "f")
    if r.ndim != 2 or r.shape != v.shape or r.shape != f.shape:
        raise ValueError(f"r, v and f must all be (n, d); got {r.shape}, {v.shape}, {f.shape}")
    if not (dt > 0.0 and math.isfinite(dt)):
        raise ValueError(f"the time step must be positive and finite, got {dt}")


def verlet_step(
    r: np.ndarray, v: np.ndarray, f: np.ndarray, force_fn: ForceFn, dt: float
) -> Tuple[np.ndarray, np.ndarray, ForceOut]:
    """Velocity Verlet: (r', v', force_fn(r'))."""
    _phase_space(r, v, f, dt)
    v = _kick(v, f, 0.5 * dt)
    r = _drift(r, v, dt)
    out = _forces(force_fn, r)
    return r, _kick(v, out[1], 0.5 * dt), out


def andersen_reset(
    key: np.ndarray, v: np.ndarray, reset_acc: float, temperature: float, p: float, mode: str
) -> Tuple[np.ndarray, np.ndarray, float, int]:
    """Andersen thermostat after a step: (key', v', reset_acc', number of resets).

    The key splits into (key', pick, draw, coin).  Mode "count" fires
    deterministically every 1/p steps (an accumulator gains p per step and gives
    up 1 when it fires); mode "bernoulli" fires when uniform(coin) < p.  A firing
    gives particle randint(pick, n) the velocity sqrt(T) normals(draw, d).
    """
    if mode not in ANDERSEN_MODES:
        raise ValueError(f"unknown Andersen mode {mode!r} (expected one of {ANDERSEN_MODES})")
    if not (0.0 <= p <= 1.0):
        raise ValueError(f"the reset probability must be in [0, 1], got {p}")
    rng.require_key(key)
    require_float64(v, "v")
    if v.ndim != 2:
        raise ValueError(f"v must be (n, d), got {v.shape}")
    key, k_pick, k_draw, k_coin = rng.split(key, 4)
    acc = float(reset_acc)
    if mode == "count":
        acc += p
        fire = acc >= 1.0
        acc = acc - 1.0 if fire else acc
    else:
        fire = bool(rng.uniform(k_coin) < p)
    out = v.copy()
    if fire:
        n, d = v.shape
        out[rng.randint(k_pick, n)] = math.sqrt(float(temperature)) * rng.normals(k_draw, d)
    return key, out, acc, int(fire)


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
    """BAOAB Langevin step: (key', r', v', force_fn(r')).

    The key splits into (key', xi); the O map uses the first n d normals of xi in
    (particle, axis) order, and 1 - c^2 is taken as -expm1(-2 gamma dt).
    """
    _phase_space(r, v, f, dt)
    if not (gamma >= 0.0 and math.isfinite(gamma)):
        raise ValueError(f"the friction must be finite and >= 0, got {gamma}")
    rng.require_key(key)
    key, k_xi = rng.split(key, 2)
    c = math.exp(-gamma * dt)
    sigma = math.sqrt(-math.expm1(-2.0 * gamma * dt) * float(temperature))
    h = 0.5 * dt
    v = _kick(v, f, h)
    r = _drift(r, v, h)
    v = c * v + sigma * rng.normals(k_xi, r.size).reshape(r.shape)
    r = _drift(r, v, h)
    out = _forces(force_fn, r)
    return key, r, _kick(v, out[1], h), out
