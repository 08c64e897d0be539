"""Batched NVE velocity-Verlet MD on the stealthy potential, stepped on the GPU.

Restates the verlet scheme of src/xla_cco/dynamics/run.py: forces at
s = r inv(B) from ``value_grad_density`` with F = -g inv(B)^T
(make_force_fn, :219-244); per step the kinetic energy, the drift guard
|Phi + KE - e0| <= drift_tol * d (N - 1) T / 2 (else STATUS_DRIFT and the
lane freezes), and running sums of KE, Phi and Phi^2 (_step, :415-442);
e0 = Phi + KE(v0) at the start (init_state, :300-381).  Positions are
Cartesian and never wrapped.

The reference's winding, COM-corrected density tracking, snapshots,
Andersen/BAOAB thermostats and checkpoints are not part of this build.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import Optional, Tuple

import numpy as np

from xla_cco import _native
from xla_cco._types import require_float64
from xla_cco.dynamics.integrators import ForceFn
from xla_cco.energy import StealthyEnergy


@dataclasses.dataclass(frozen=True)
class MDSettings:
    """Integration controls of the verlet scheme (dynamics/state.py:57-88 subset)."""

    dt: float = 0.01
    scheme: str = "verlet"
    remove_com: bool = True
    chunk_steps: int = 10_000
    drift_tol: float = 1e-3
    # "mixed": forces from the split-precision MFMA kernels (positions and
    # velocities stay FP64); an extension, "fp64" is the reference's arithmetic
    precision: str = "fp64"


@dataclasses.dataclass
class MDState:
    """Batched lane state (leading lane axis) plus the device session holding it."""

    r: np.ndarray
    v: np.ndarray
    f: np.ndarray
    phi: np.ndarray
    kinetic: np.ndarray
    temperature: np.ndarray
    step: np.ndarray
    time: np.ndarray
    e0: np.ndarray
    max_drift: np.ndarray
    sum_kinetic: np.ndarray
    sum_phi: np.ndarray
    sum_phi2: np.ndarray
    n_sum: np.ndarray
    status: np.ndarray
    session: Optional["_MDSession"] = dataclasses.field(default=None, repr=False)


@dataclasses.dataclass
class MDResult:
    state: MDState


def make_force_fn(energy: StealthyEnergy, basis: np.ndarray) -> ForceFn:
    """r -> (Phi, F, re, im) on the GPU at s = r inv(B), F = -g inv(B)^T (run.py:219-244)."""
    b = np.asarray(basis, dtype=np.float64)
    d = energy.dim
    if b.shape != (d, d):
        raise ValueError(f"basis must have shape ({d}, {d}), got {b.shape}")
    inv = np.linalg.inv(b)
    inv_t = inv.T

    def force_fn(r: np.ndarray) -> Tuple[float, np.ndarray, np.ndarray, np.ndarray]:
        assert r.ndim == 2 and r.shape[1] == d, "r must have shape (n, d)"
        phi, g, re, im = energy.value_grad_density(np.ascontiguousarray(r @ inv))
        return phi, -g @ inv_t, re, im

    return force_fn


class _MDSession:
    def __init__(self, energy: StealthyEnergy, basis: np.ndarray, b: int, n: int, d: int,
                 settings: MDSettings) -> None:
        ctx = _native.context(energy.device)
        h = C.c_void_p()
        bas = np.ascontiguousarray(np.asarray(basis, dtype=np.float64))
        tol = settings.drift_tol if math.isfinite(settings.drift_tol) else math.inf
        _native.check(_native.lib().cco_md_create(ctx.handle, energy.handle(), b, n, d, _native.dptr(bas),
                                                  float(settings.dt), tol, C.byref(h)))
        self.handle = h
        if settings.precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {settings.precision!r}")
        if settings.precision == "mixed":
            _native.check(_native.lib().cco_relax_set_precision(h, 1, 0.0))
        self.b, self.n, self.d = b, n, d

    def close(self) -> None:
        if self.handle is not None and _native._lib is not None:
            _native._lib.cco_relax_destroy(self.handle)
        self.handle = None

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def read(self, session_ref: "Optional[_MDSession]" = None) -> MDState:
        b, n, d = self.b, self.n, self.d
        r = np.empty((b, n, d))
        v = np.empty((b, n, d))
        f = np.empty((b, n, d))
        sc = np.empty((b, 9))
        it = np.empty((b, 3), dtype=np.int64)
        _native.check(_native.lib().cco_md_state(self.handle, _native.dptr(r), _native.dptr(v), _native.dptr(f),
                                                 _native.dptr(sc), _native.i64ptr(it)))
        return MDState(
            r=r, v=v, f=f, phi=sc[:, 0].copy(), kinetic=sc[:, 1].copy(), temperature=sc[:, 8].copy(),
            step=it[:, 0].copy(), time=sc[:, 7].copy(), e0=sc[:, 2].copy(), max_drift=sc[:, 3].copy(),
            sum_kinetic=sc[:, 4].copy(), sum_phi=sc[:, 5].copy(), sum_phi2=sc[:, 6].copy(),
            n_sum=it[:, 1].copy(), status=it[:, 2].copy(), session=self,
        )


def init_state(
    energy: StealthyEnergy,
    basis: np.ndarray,
    r0: np.ndarray,
    v0: np.ndarray,
    temperature,
    settings: MDSettings = MDSettings(),
) -> MDState:
    """Batched lanes at step 0: forces at r0, e0 = Phi + KE(v0) (run.py:300-381).

    With ``settings.remove_com`` the mean velocity of each lane is removed
    first (begin_segment, run.py:384-412).
    """
    require_float64(r0, "r0")
    require_float64(v0, "v0")
    r0 = np.asarray(r0)
    v0 = np.asarray(v0)
    if r0.ndim != 3 or v0.shape != r0.shape:
        raise ValueError(f"r0 and v0 must share a (b, n, d) shape: {r0.shape}, {v0.shape}")
    if settings.scheme != "verlet":
        raise NotImplementedError(f"scheme {settings.scheme!r}: only 'verlet' (NVE) is built")
    b, n, d = r0.shape
    if energy.dim != d:
        raise ValueError(f"positions have dimension {d}, the energy {energy.dim}")
    if settings.remove_com:
        v0 = v0 - np.mean(v0, axis=1, keepdims=True)
    temp = np.ascontiguousarray(np.broadcast_to(np.asarray(temperature, dtype=np.float64), (b,)))
    sess = _MDSession(energy, basis, b, n, d, settings)
    _native.check(_native.lib().cco_md_init(sess.handle, _native.dptr(np.ascontiguousarray(r0)),
                                            _native.dptr(np.ascontiguousarray(v0)), _native.dptr(temp)))
    return sess.read()


def simulate(state: MDState, n_steps: int, settings: MDSettings = MDSettings()) -> MDResult:
    """Advance every running lane by n_steps Verlet steps on the device (run.py:592-638)."""
    if int(n_steps) != n_steps or n_steps < 0:
        raise ValueError(f"need n_steps >= 0, got {n_steps}")
    sess = state.session
    if sess is None or sess.handle is None:
        raise ValueError("state has no device session (build it with init_state)")
    na = C.c_int(0)
    _native.check(_native.lib().cco_md_run(sess.handle, int(n_steps), max(1, min(int(settings.chunk_steps), 1024)),
                                           C.byref(na)))
    return MDResult(state=sess.read())
