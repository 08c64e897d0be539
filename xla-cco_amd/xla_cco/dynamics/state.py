"""MD containers: settings, tracked waves, snapshots, the rho trace and the lane state.

src/xla_cco/dynamics/state.py:20-172, as NumPy dataclasses.  An ``MDState``
holds one lane (r of shape (n, d)) or a batch (a leading lane axis on every
field); the field order is the reference's, so checkpoints list the same
leaves in the same order (dynamics/checkpoint.py).

Andersen "count" mode: an accumulator gains p per step and one uniformly
chosen particle is redrawn from N(0, T) each time it reaches 1 (the
accumulator is carried across calls); "bernoulli" redraws one particle with
probability p per step.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Iterator, Optional, Tuple

import numpy as np

SCHEMES: Tuple[str, ...] = ("verlet", "andersen", "baoab")
ANDERSEN_MODES: Tuple[str, ...] = ("count", "bernoulli")


@dataclasses.dataclass(frozen=True)
class MDSettings:
    """Integration controls (state.py:57-88), same fields and defaults.

    ``dt`` time step; ``scheme`` in SCHEMES; ``andersen_p`` reset rate per step
    and ``andersen_mode`` in ANDERSEN_MODES; ``gamma`` BAOAB friction;
    ``remove_com`` zeroes the total momentum when a segment begins;
    ``chunk_steps`` steps per device call; ``wind_every`` cadence of the
    unconstrained waves; ``record_steps`` absolute step numbers of the
    snapshots (strictly increasing, at most 256); ``record_rho_every`` /
    ``max_rho_records`` the dense rho trace (0: off); ``phase_floor`` on
    S = |rho|^2 / N; ``drift_tol`` aborts a "verlet" lane (STATUS_DRIFT) when
    |E - E0| exceeds drift_tol d (N - 1) T / 2 (inf disables the guard).
    ``precision`` "mixed" evaluates the forces with the split-precision MFMA
    kernels (an extension; "fp64" is the reference's arithmetic).
    """

    dt: float = 0.01
    scheme: str = "verlet"
    andersen_p: float = 0.01
    andersen_mode: str = "count"
    gamma: float = 0.0
    remove_com: bool = True
    chunk_steps: int = 10_000
    wind_every: int = 1
    record_steps: Tuple[int, ...] = ()
    record_rho_every: int = 0
    max_rho_records: int = 0
    phase_floor: float = 1e-16
    drift_tol: float = 1e-3
    precision: str = "fp64"


@dataclasses.dataclass
class Waves:
    """Tracked density waves, the constrained set first (state.py:91-104).

    ``n_vectors`` (w, d) integer-valued reciprocal coefficients, ``kvec``
    (w, d) Cartesian wavevectors n @ 2 pi inv(B)^T, ``constrained`` (w,) True
    for the first ``n_constrained`` rows, which equal the energy's first
    n_vectors (their rho comes from the force pass).
    """

    n_vectors: np.ndarray
    kvec: np.ndarray
    constrained: np.ndarray
    n_constrained: int


@dataclasses.dataclass
class Snapshots:
    """Snapshots at ``record_steps`` (``count`` valid rows; step -1 = empty)."""

    step: np.ndarray
    time: np.ndarray
    wind_abs: np.ndarray
    wind_net: np.ndarray
    rho: np.ndarray
    r: np.ndarray
    phi: np.ndarray
    kinetic: np.ndarray
    count: np.ndarray


@dataclasses.dataclass
class RhoTrace:
    """Dense rho of the tracked waves every ``record_rho_every`` steps."""

    rho: np.ndarray
    time: np.ndarray
    count: np.ndarray


@dataclasses.dataclass
class MDState:
    """Lane state (state.py:129-165); a batch carries a leading lane axis.

    ``key`` (2,) uint32 (xla_cco.dynamics.rng); ``reset_acc`` the Andersen
    count accumulator; ``rho_prev`` (w, 2) the last rho used for winding
    (COM-corrected unless verlet; NaN rows are pending, recomputed before the
    next step); ``e0`` the total energy at the segment start; ``sum_*``
    running sums over ``n_sum`` steps; ``status`` STATUS_RUNNING or
    STATUS_DRIFT.
    """

    r: np.ndarray
    v: np.ndarray
    f: np.ndarray
    phi: np.ndarray
    temperature: np.ndarray
    key: np.ndarray
    step: np.ndarray
    time: np.ndarray
    reset_acc: np.ndarray
    rho_prev: np.ndarray
    wind_abs: np.ndarray
    wind_net: np.ndarray
    n_valid: np.ndarray
    e0: np.ndarray
    max_drift: np.ndarray
    sum_kinetic: np.ndarray
    sum_phi: np.ndarray
    sum_phi2: np.ndarray
    sum_s: np.ndarray
    n_sum: np.ndarray
    status: np.ndarray
    snap: Snapshots
    trace: RhoTrace

    @property
    def batched(self) -> bool:
        return np.ndim(self.r) == 3


@dataclasses.dataclass
class MDResult:
    """Final (batched) state, its snapshots and the dense rho trace (None when off)."""

    state: MDState
    snapshots: Snapshots
    trace: Optional[RhoTrace]


def leaves(state: MDState) -> Iterator[np.ndarray]:
    """The arrays of a state in field order (nested snapshots and trace in theirs)."""
    for f in dataclasses.fields(state):
        x = getattr(state, f.name)
        if dataclasses.is_dataclass(x):
            for g in dataclasses.fields(x):
                yield getattr(x, g.name)
        else:
            yield x


def with_leaves(state: MDState, arrays) -> MDState:  # type: ignore[no-untyped-def]
    """A state of the same structure whose arrays are ``arrays`` in leaf order."""
    it = iter(arrays)
    kw = {}
    for f in dataclasses.fields(state):
        x = getattr(state, f.name)
        if dataclasses.is_dataclass(x):
            kw[f.name] = type(x)(**{g.name: next(it) for g in dataclasses.fields(x)})
        else:
            kw[f.name] = next(it)
    return MDState(**kw)


def stack(states) -> MDState:  # type: ignore[no-untyped-def]
    """A batch from single lanes (the reference builds it with jax.vmap)."""
    states = list(states)
    assert states and not states[0].batched, "stack single lanes"
    cols = zip(*(list(leaves(s)) for s in states))
    return with_leaves(states[0], [np.stack(c) for c in cols])


def lane(state: MDState, i: int) -> MDState:
    """Lane i of a batch."""
    assert state.batched, "lane() of a batched state"
    return with_leaves(state, [np.array(x[i]) for x in leaves(state)])


def check_settings(s: MDSettings) -> None:
    """Raise ValueError on settings the driver cannot run (run.py:108-133)."""
    if s.scheme not in SCHEMES or s.andersen_mode not in ANDERSEN_MODES:
        raise ValueError(f"unknown scheme {s.scheme!r} or mode {s.andersen_mode!r}")
    if not (math.isfinite(s.dt) and s.dt > 0.0):
        raise ValueError(f"dt must be finite and positive, got {s.dt}")
    if not (0.0 <= s.andersen_p <= 1.0) or not (0.0 <= s.gamma < math.inf):
        raise ValueError(f"need 0 <= andersen_p <= 1 and gamma >= 0: {s}")
    if s.chunk_steps < 1 or s.wind_every < 1:
        raise ValueError("chunk_steps and wind_every must be >= 1")
    rs = tuple(s.record_steps)
    if len(rs) > 256 or any(int(x) != x or x < 1 for x in rs):
        raise ValueError("record_steps: at most 256 integers >= 1")
    if any(b <= a for a, b in zip(rs, rs[1:])):
        raise ValueError("record_steps must be strictly increasing")
    if s.record_rho_every < 0 or s.max_rho_records < 0:
        raise ValueError("record_rho_every and max_rho_records must be >= 0")
    if s.record_rho_every > 0 and s.max_rho_records < 1:
        raise ValueError("a rho trace (record_rho_every > 0) needs max_rho_records")
    if not (s.phase_floor >= 0.0) or not (s.drift_tol > 0.0):
        raise ValueError("need phase_floor >= 0 and drift_tol > 0")
    if s.precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {s.precision!r}")
