"""Gradient flow dr/dt = -grad_r Phi on the GPU: Dormand-Prince 5(4) or ROS2.

Restates src/xla_cco/minimize/flow.py:93-481 (``method="dopri5"``): error
control on the RMS Cartesian displacement (atol + rtol |step|) and on the
clocks (flow time, arclength), FSAL, controller factor
clip(0.9 err^-1/5, 0.2, 5), optional arclength parametrisation, stops by
energy / gradient / time / budget / step floor.  Every stage of every lane
runs in ``flow_advance_kernel``; energies are HIP kernels for a
``StealthyEnergy`` or host callbacks for another ``AbstractEnergy``.

``method="rosenbrock"`` (flow.py:367-398): L-stable ROS2 with the exact
Hessian, time parametrization only.  Each round factors W = I - gamma h J
(J = -H (I (x) inv(B B^T))) for every running lane on the GPU (closed-form
Hessian, rocSOLVER LU, three solves, two energy evaluations) and applies the
same step control with the order-2 exponent.  It needs a ``StealthyEnergy``
(another ``AbstractEnergy`` raises NotImplementedError: its energy and
Hessian live on the host).
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import Optional

import numpy as np

from xla_cco import _native
from xla_cco._types import require_float64
from xla_cco.energy import AbstractEnergy, StealthyEnergy


@dataclasses.dataclass(frozen=True)
class FlowSettings:
    """Integration controls (flow.py:93-125), same fields and defaults."""

    rtol: float = 1e-6
    atol: float = 1e-14
    parametrization: str = "time"
    method: str = "dopri5"
    dt_init: float = 1e-3
    dt_min: float = 1e-14
    dt_max: float = math.inf
    t_max: float = math.inf
    max_steps: int = 200_000
    energy_tol: float = 0.0
    grad_tol: float = 0.0
    record_every: int = 0
    max_records: int = 0
    # Split precision (no reference counterpart; the reference is FP64): "mixed"
    # evaluates rho and the forces on the matrix cores (fp16 hi/lo operands, FP32
    # accumulation) until a step starts at Phi <= precision_switch, then FP64.
    precision: str = "fp64"
    precision_switch: float = 1e-8


@dataclasses.dataclass
class FlowTrajectory:
    """Recorded flow (capacity max_records, ``count`` valid rows; s unwrapped)."""

    s: np.ndarray
    t: np.ndarray
    length: np.ndarray
    energy: np.ndarray
    count: np.ndarray


@dataclasses.dataclass
class FlowResult:
    """Final state (flow.py:138-154); batched results carry a leading lane axis."""

    s: np.ndarray
    energy: np.ndarray
    t: np.ndarray
    length: np.ndarray
    dt: np.ndarray
    n_steps: np.ndarray
    n_evals: np.ndarray
    status: np.ndarray
    trajectory: Optional[FlowTrajectory]

    def lane(self, i: int) -> "FlowResult":
        tr = self.trajectory
        sub = None
        if tr is not None:
            sub = FlowTrajectory(tr.s[i], tr.t[i], tr.length[i], tr.energy[i], tr.count[i])
        return FlowResult(self.s[i], self.energy[i], self.t[i], self.length[i], self.dt[i],
                          self.n_steps[i], self.n_evals[i], self.status[i], sub)


def _validate(s0: np.ndarray, basis: np.ndarray, settings: FlowSettings) -> None:
    """Raise on bad inputs before any device work (flow.py:186-209)."""
    require_float64(s0, "s0")
    if s0.ndim != 3 or basis.shape != (s0.shape[2], s0.shape[2]):
        raise ValueError(f"s0 must be (n, d) with a (d, d) basis, got {s0.shape[1:]}")
    positive = (settings.atol, settings.rtol, settings.dt_init, settings.max_steps)
    if not all(x > 0 for x in positive):
        raise ValueError("atol, rtol, dt_init and max_steps must be positive")
    if settings.parametrization not in ("time", "arclength"):
        raise ValueError(f"unknown parametrization {settings.parametrization!r}")
    if settings.method not in ("dopri5", "rosenbrock"):
        raise ValueError(f"unknown method {settings.method!r}")
    if settings.method == "rosenbrock" and settings.parametrization != "time":
        raise ValueError("method='rosenbrock' integrates the time parametrization")
    if settings.parametrization == "arclength" and math.isfinite(settings.t_max):
        raise ValueError("t_max is supported only with parametrization='time'")
    if settings.record_every < 0 or settings.max_records < 0:
        raise ValueError("record_every and max_records must be >= 0")
    if (settings.record_every > 0) != (settings.max_records > 1):
        raise ValueError("recording needs record_every > 0 and max_records >= 2")
    if settings.precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {settings.precision!r}")
    if settings.precision == "mixed" and settings.method != "dopri5":
        raise ValueError("precision='mixed' is available for method='dopri5'")
    if not settings.precision_switch >= 0.0:
        raise ValueError("precision_switch must be >= 0")



def gradient_flow_batch(
    energy: AbstractEnergy,
    s0: np.ndarray,
    basis: np.ndarray,
    settings: FlowSettings,
    dt0: Optional[np.ndarray] = None,
    device: int = 0,
) -> FlowResult:
    """Integrate every configuration of s0 (b, n, d) until its own stop."""
    s0 = np.asarray(s0) if hasattr(s0, "dtype") else s0
    basis = np.asarray(basis, dtype=np.float64)
    _validate(s0, basis, settings)
    s0 = np.ascontiguousarray(s0)
    b, n, d = s0.shape
    builtin = isinstance(energy, StealthyEnergy)
    if builtin and energy.dim != d:  # type: ignore[union-attr]
        raise ValueError("dimension mismatch")
    ctx = _native.context(energy.device if builtin else device)  # type: ignore[union-attr]
    fs = _native.FlowSettingsC(
        settings.rtol, settings.atol, settings.dt_init, settings.dt_min, settings.dt_max,
        settings.t_max, settings.energy_tol, settings.grad_tol, int(settings.max_steps),
        1 if settings.parametrization == "arclength" else 0,
        int(settings.record_every), int(settings.max_records),
    )
    L = _native.lib()
    h = C.c_void_p()
    bas = np.ascontiguousarray(basis)
    _native.check(
        L.cco_flow_create(ctx.handle, energy.handle() if builtin else None, b, n, d,  # type: ignore[union-attr]
                          C.byref(fs), _native.dptr(bas), C.byref(h))
    )
    try:
        if settings.method == "rosenbrock":
            if not builtin:
                raise NotImplementedError(
                    "method='rosenbrock' runs on the GPU Hessian of a StealthyEnergy; "
                    f"{type(energy).__name__} is evaluated on the host"
                )
            _native.check(L.cco_flow_set_method(h, 1))
        if settings.precision == "mixed":
            if not builtin:
                raise ValueError("precision='mixed' needs a StealthyEnergy (its kernels run on the GPU)")
            _native.check(L.cco_relax_set_precision(h, 1, float(settings.precision_switch)))
        dt0a = None
        if dt0 is not None:
            dt0a = np.ascontiguousarray(np.broadcast_to(np.asarray(dt0, dtype=np.float64), (b,)))
        _native.check(L.cco_flow_init(h, _native.dptr(s0), _native.dptr(dt0a) if dt0a is not None else None))
        bound = 1 + 6 * 4 * int(settings.max_steps) + 6  # evaluations (ROS2 uses 2 per attempt)
        if builtin:
            _native.check(L.cco_relax_run(h, bound, 256))
        else:
            _external_loop(L, h, energy, b, n, d, bound)
        s = np.empty((b, n, d))
        out = {k: np.empty(b) for k in ("energy", "t", "length", "dt")}
        ints = {k: np.empty(b, dtype=np.int64) for k in ("n_steps", "n_evals", "status")}
        _native.check(
            L.cco_flow_result(h, _native.dptr(s), _native.dptr(out["energy"]), _native.dptr(out["t"]),
                              _native.dptr(out["length"]), _native.dptr(out["dt"]),
                              _native.i64ptr(ints["n_steps"]), _native.i64ptr(ints["n_evals"]),
                              _native.i64ptr(ints["status"]))
        )
        traj = None
        if settings.record_every > 0:
            cap = settings.max_records
            cnt = np.empty(b, dtype=np.int64)
            ts = np.zeros((b, cap, n, d))
            tt = np.zeros((b, cap))
            tl = np.zeros((b, cap))
            te = np.zeros((b, cap))
            _native.check(L.cco_flow_trajectory(h, _native.i64ptr(cnt), _native.dptr(ts), _native.dptr(tt),
                                                 _native.dptr(tl), _native.dptr(te)))
            # append the final state; overwrite the last row if full or already final (flow.py:454-468)
            for i in range(b):
                last = int(cnt[i]) - 1
                final_is_last = tl[i, last] == out["length"][i] and tt[i, last] == out["t"][i]
                idx = last if final_is_last else min(int(cnt[i]), cap - 1)
                ts[i, idx] = s[i]
                tt[i, idx] = out["t"][i]
                tl[i, idx] = out["length"][i]
                te[i, idx] = out["energy"][i]
                cnt[i] = idx + 1
            traj = FlowTrajectory(ts, tt, tl, te, cnt)
        return FlowResult(s, out["energy"], out["t"], out["length"], out["dt"], ints["n_steps"],
                          ints["n_evals"], ints["status"], traj)
    finally:
        L.cco_relax_destroy(h)


def _external_loop(L, h, energy: AbstractEnergy, b: int, n: int, d: int, bound: int) -> None:
    lanes = np.empty(b, dtype=np.int32)
    trial = np.empty((b, n, d))
    for _ in range(bound):
        na = C.c_int(0)
        _native.check(L.cco_relax_pending(h, C.byref(na), lanes.ctypes.data_as(C.POINTER(C.c_int32)),
                                          _native.dptr(trial)))
        k = na.value
        if k == 0:
            return
        phi = np.empty(k)
        grad = np.empty((k, n, d))
        for i in range(k):
            f, g = energy.value_and_grad(trial[i].copy())
            phi[i] = float(f)
            grad[i] = np.asarray(g, dtype=np.float64).reshape(n, d)
        _native.check(L.cco_relax_submit(h, _native.dptr(phi), _native.dptr(grad)))
    raise RuntimeError("gradient flow did not finish within its step bound")


def gradient_flow(
    energy: AbstractEnergy,
    s0: np.ndarray,
    basis: np.ndarray,
    settings: FlowSettings,
    dt0: Optional[float] = None,
) -> FlowResult:
    """Integrate one configuration (n, d) from s0 until a stop (flow.py:258-323)."""
    require_float64(s0, "s0")
    s0 = np.asarray(s0)
    if s0.ndim != 2:
        raise ValueError(f"s0 must be (n, d), got {s0.shape}")
    res = gradient_flow_batch(energy, s0[None], basis, settings, None if dt0 is None else np.array([dt0]))
    return res.lane(0)
