"""Batched relaxation on the GPU (src/xla_cco/minimize/__init__.py:102-151).

``relax(energy, s0, settings)`` dispatches on the settings type
(``DescentSettings`` -> adaptive descent, ``LBFGSSettings`` -> L-BFGS) and
returns a ``RelaxResult`` whose leaves carry a leading lane axis, in input
order.  Every lane's minimizer state lives in device memory; lanes advance
one energy evaluation per step, a finished lane leaves the active list on
the device and is never written again, so results do not depend on which
other lanes share the batch or on ``chunk_iters`` (which only sets how often
the host reads the active count).

For a ``StealthyEnergy`` every evaluation is a HIP kernel.  For any other
``AbstractEnergy`` the host evaluates ``energy.value_and_grad`` on the
pending trial points each step and hands the values back to the device
minimizer.
"""

from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple, Union

import numpy as np

from xla_cco import _native
from xla_cco._types import require_float64
from xla_cco.energy import AbstractEnergy, StealthyEnergy
from xla_cco.minimize.descent import DescentSettings
from xla_cco.minimize.descent import max_rounds as _descent_rounds
from xla_cco.minimize.lbfgs import LBFGSSettings
from xla_cco.minimize.lbfgs import max_rounds as _lbfgs_rounds
from xla_cco.minimize.result import (
    STATUS_BUDGET,
    STATUS_ENERGY_TOL,
    STATUS_GRAD_TOL,
    STATUS_LINESEARCH_FAIL,
    STATUS_RUNNING,
    STATUS_STAGNATED,
    STATUS_STEP_FLOOR,
    RelaxResult,
    Trajectory,
)

__all__ = [
    "relax",
    "relax_uneven",
    "UnevenSession",
    "DescentSettings",
    "LBFGSSettings",
    "RelaxResult",
    "Trajectory",
    "RelaxSession",
    "STATUS_RUNNING",
    "STATUS_ENERGY_TOL",
    "STATUS_STEP_FLOOR",
    "STATUS_BUDGET",
    "STATUS_STAGNATED",
    "STATUS_LINESEARCH_FAIL",
    "STATUS_GRAD_TOL",
]

Settings = Union[DescentSettings, LBFGSSettings]


def _c_settings(settings: Settings):  # type: ignore[no-untyped-def]
    if isinstance(settings, DescentSettings):
        if settings.max_evals < 0:
            raise ValueError("max_evals must be >= 0")
        ds = _native.DescentSettingsC(
            settings.step_init,
            settings.step_min,
            settings.shrink,
            settings.grow,
            settings.rel_tol,
            settings.energy_tol,
            int(settings.max_evals),
            int(settings.record_every),
            int(settings.max_records),
        )
        return _native.METHOD_DESCENT, ds, None
    if isinstance(settings, LBFGSSettings):
        if not 1 <= len(settings.rescales) <= 8:
            raise ValueError("need 1 to 8 rescale passes")
        arr = (C.c_double * 8)(*([float(r) for r in settings.rescales] + [0.0] * (8 - len(settings.rescales))))
        ls = _native.LBFGSSettingsC(
            int(settings.memory),
            len(settings.rescales),
            arr,
            int(settings.max_evals_per_pass),
            float(settings.energy_tol),
            float(settings.grad_tol),
            int(settings.stagnation_patience),
            int(settings.max_linesearch_steps),
            int(settings.record_every),
            int(settings.max_records),
        )
        return _native.METHOD_LBFGS, None, ls
    raise TypeError(f"unsupported settings type {type(settings).__name__}")


def _eval_bound(settings: Settings, chunk_iters: int) -> int:
    """Evaluation steps allowed before the round bound (>= the budget)."""
    if isinstance(settings, DescentSettings):
        return _descent_rounds(settings, chunk_iters) * chunk_iters + 1
    return _lbfgs_rounds(settings, chunk_iters) * chunk_iters + 1


class RelaxSession:
    """One device relax session (cco_relax) over B lanes of shape (N, d)."""

    def __init__(
        self,
        energy: AbstractEnergy,
        b: int,
        n: int,
        d: int,
        settings: Settings,
        device: int = 0,
    ) -> None:
        self.method, ds, ls = _c_settings(settings)
        self.settings = settings
        self.energy = energy
        self.b, self.n, self.d = b, n, d
        self.builtin = isinstance(energy, StealthyEnergy)
        ctx = _native.context(device if not self.builtin else energy.device)  # type: ignore[union-attr]
        eh = energy.handle() if self.builtin else None  # type: ignore[union-attr]
        h = C.c_void_p()
        _native.check(
            _native.lib().cco_relax_create(
                ctx.handle,
                eh,
                b,
                n,
                d,
                self.method,
                C.byref(ds) if ds is not None else None,
                C.byref(ls) if ls is not None else None,
                C.byref(h),
            )
        )
        self.handle = h
        if isinstance(settings, LBFGSSettings):
            if settings.precision not in ("fp64", "mixed"):
                raise ValueError(f"precision must be 'fp64' or 'mixed', got {settings.precision!r}")
            if not settings.precision_switch >= 0.0:
                raise ValueError("precision_switch must be >= 0")
            if settings.precision == "mixed":
                _native.check(
                    _native.lib().cco_relax_set_precision(h, 1, float(settings.precision_switch))
                )
            _apply_split_history(h, settings)

    def precision_state(self) -> Tuple[bool, int]:
        """(split precision enabled by the last init, lanes still split now)."""
        mode, lanes = C.c_int(0), C.c_int(0)
        _native.check(
            _native.lib().cco_relax_precision_state(self.handle, C.byref(mode), C.byref(lanes))
        )
        return bool(mode.value), int(lanes.value)

    def close(self) -> None:
        if self.handle is not None and _native._lib is not None:
            _native._lib.cco_relax_destroy(self.handle)
        self.handle = None

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def init(self, s0: np.ndarray) -> None:
        s0 = np.ascontiguousarray(s0)
        _native.check(_native.lib().cco_relax_init(self.handle, _native.dptr(s0)))

    def run(self, max_evals: int, sync_every: int) -> None:
        if self.builtin:
            _native.check(_native.lib().cco_relax_run(self.handle, int(max_evals), int(sync_every)))
            return
        L = _native.lib()
        steps = 0
        lanes = np.empty(self.b, dtype=np.int32)
        trial = np.empty((self.b, self.n, self.d))
        while True:
            na = C.c_int(0)
            _native.check(
                L.cco_relax_pending(
                    self.handle,
                    C.byref(na),
                    lanes.ctypes.data_as(C.POINTER(C.c_int32)),
                    _native.dptr(trial),
                )
            )
            k = na.value
            if k == 0:
                return
            if steps >= max_evals:
                raise RuntimeError("relax did not finish within the round bound")
            phi = np.empty(k)
            grad = np.empty((k, self.n, self.d))
            for i in range(k):
                f, g = self.energy.value_and_grad(trial[i].copy())
                phi[i] = float(f)
                grad[i] = np.asarray(g, dtype=np.float64).reshape(self.n, self.d)
            _native.check(L.cco_relax_submit(self.handle, _native.dptr(phi), _native.dptr(grad)))
            steps += 1

    def result(self) -> RelaxResult:
        b, n, d = self.b, self.n, self.d
        s = np.empty((b, n, d))
        e = np.empty(b)
        gn = np.empty(b)
        ne = np.empty(b, dtype=np.int64)
        ni = np.empty(b, dtype=np.int64)
        st = np.empty(b, dtype=np.int64)
        _native.check(
            _native.lib().cco_relax_result(
                self.handle,
                _native.dptr(s),
                _native.dptr(e),
                _native.dptr(gn),
                _native.i64ptr(ne),
                _native.i64ptr(ni),
                _native.i64ptr(st),
            )
        )
        traj = None
        every = self.settings.record_every
        cap = self.settings.max_records
        if every > 0 and cap > 0:
            cnt = np.empty(b, dtype=np.int64)
            it = np.zeros((b, cap), dtype=np.int64)
            ts = np.zeros((b, cap, n, d))
            te = np.zeros((b, cap))
            tt = np.zeros((b, cap))
            _native.check(
                _native.lib().cco_relax_trajectory(
                    self.handle,
                    _native.i64ptr(cnt),
                    _native.i64ptr(it),
                    _native.dptr(ts),
                    _native.dptr(te),
                    _native.dptr(tt),
                )
            )
            traj = Trajectory(s=ts, energy=te, step=tt, iteration=it, count=cnt)
        return RelaxResult(s=s, energy=e, grad_norm=gn, n_evals=ne, n_iters=ni, status=st, trajectory=traj)

    def stats(self) -> int:
        steps = C.c_int64(0)
        _native.check(_native.lib().cco_relax_stats(self.handle, C.byref(steps), None))
        return int(steps.value)


def relax(
    energy: AbstractEnergy,
    s0: np.ndarray,
    settings: Settings,
    chunk_iters: int = 2_000,
    autotune: bool = False,
    tune_exact: bool = True,
) -> RelaxResult:
    """Relax a batch (b, n, d) of configurations; leaves have a leading lane axis.

    ``autotune`` (no reference counterpart) runs the built-in energy with the
    launch and tile knobs xla_cco.minimize.tuning picked for this workload class,
    timing the candidates on ``s0`` the first time; results are bit-identical.
    ``tune_exact=False`` lets it also pick the rho split length and the L-BFGS
    advance variant, which change results at rounding level (the tolerance is
    stated in xla_cco.minimize.tuning); the energy's options are restored after.

    Raises TypeError if ``s0`` is not float64 or the settings type is unknown,
    RuntimeError if lanes are still running after the round bound.
    """
    require_float64(s0, "s0")
    s0 = np.asarray(s0)
    assert s0.ndim == 3, "s0 must have shape (b, n, d)"
    assert chunk_iters >= 1, "chunk_iters must be positive"
    assert s0.shape[0] >= 1, "need at least one configuration"
    if not isinstance(settings, (DescentSettings, LBFGSSettings)):
        raise TypeError(f"unsupported settings type {type(settings).__name__}")
    if isinstance(energy, StealthyEnergy) and s0.shape[2] != energy.dim:
        raise ValueError("dimension mismatch")
    b, n, d = s0.shape
    choice = None
    if autotune and isinstance(energy, StealthyEnergy):
        from xla_cco.minimize import tuning

        choice = tuning.choice_for(energy, s0, settings, exact=tune_exact)
    sess = RelaxSession(energy, b, n, d, settings)
    saved = None
    try:
        if choice is not None:
            saved = tuning.energy_options(energy)
            tuning.apply(sess, choice)
        sess.init(s0)
        sess.run(_eval_bound(settings, chunk_iters), sync_every=max(1, min(chunk_iters, 256)))
        return sess.result()
    finally:
        sess.close()
        if saved is not None:
            tuning.restore_energy(energy, saved)


def _apply_split_history(h, settings: LBFGSSettings) -> None:
    """CCO_OPT_SPLIT_HISTORY from ``settings.split_history`` ("fp64" or "fp32")."""
    if settings.split_history not in ("fp64", "fp32"):
        raise ValueError(f"split_history must be 'fp64' or 'fp32', got {settings.split_history!r}")
    _native.check(_native.lib().cco_relax_set_option(h, 5, 1 if settings.split_history == "fp32" else 0))


class UnevenSession:
    """Device session over configurations of different N, d and k-set.

    ``energies[i]`` and ``shapes[i] = (n_i, d_i)`` describe lane i.  Positions
    are packed lane after lane (``pack`` / ``unpack``); ``init``, ``run`` and
    ``result`` mirror :class:`RelaxSession`.
    """

    def __init__(self, energies: Sequence[StealthyEnergy], shapes: Sequence[Tuple[int, int]],
                 settings: Settings) -> None:
        if len(energies) != len(shapes) or len(shapes) == 0:
            raise ValueError("need one energy per configuration and at least one configuration")
        if not isinstance(settings, (DescentSettings, LBFGSSettings)):
            raise TypeError(f"unsupported settings type {type(settings).__name__}")
        if isinstance(settings, LBFGSSettings):
            if settings.precision not in ("fp64", "mixed"):
                raise ValueError(f"precision must be 'fp64' or 'mixed', got {settings.precision!r}")
            if not settings.precision_switch >= 0.0:
                raise ValueError("precision_switch must be >= 0")
        uniq: List[StealthyEnergy] = []
        index = {}
        self.b = len(shapes)
        self.lane_e = np.empty(self.b, dtype=np.int32)
        self.lane_n = np.empty(self.b, dtype=np.int32)
        self.shapes = [(int(n), int(d)) for n, d in shapes]
        for i, (e, (n, d)) in enumerate(zip(energies, self.shapes)):
            if not isinstance(e, StealthyEnergy):
                raise TypeError("uneven batches need StealthyEnergy instances")
            if d != e.dim or n < 1:
                raise ValueError(f"configuration {i}: shape ({n}, {d}) does not fit its energy (d={e.dim})")
            if id(e) not in index:
                index[id(e)] = len(uniq)
                uniq.append(e)
            self.lane_e[i] = index[id(e)]
            self.lane_n[i] = n
        if len({e.device for e in uniq}) != 1:
            raise ValueError("all energies must live on one device")
        self.settings = settings
        self.energies = uniq
        self.size = int(sum(n * d for n, d in self.shapes))
        method, ds, ls = _c_settings(settings)
        ctx = _native.context(uniq[0].device)
        handles = (C.c_void_p * len(uniq))(*[e.handle().value for e in uniq])
        h = C.c_void_p()
        _native.check(
            _native.lib().cco_relax_uneven_create(
                ctx.handle,
                len(uniq),
                C.cast(handles, C.POINTER(C.c_void_p)),
                self.b,
                self.lane_e.ctypes.data_as(C.POINTER(C.c_int32)),
                self.lane_n.ctypes.data_as(C.POINTER(C.c_int32)),
                method,
                C.byref(ds) if ds is not None else None,
                C.byref(ls) if ls is not None else None,
                C.byref(h),
            )
        )
        self.handle = h
        # split precision: lanes whose k-set and N fit the one-workgroup-per-lane
        # kernel start split and turn FP64 below the switch; the others run FP64
        if isinstance(settings, LBFGSSettings) and settings.precision == "mixed":
            _native.check(_native.lib().cco_relax_set_precision(h, 1, float(settings.precision_switch)))
        if isinstance(settings, LBFGSSettings):
            _apply_split_history(h, settings)

    def precision_state(self) -> Tuple[bool, int]:
        """(split precision enabled by the last init, lanes still split now)."""
        mode, lanes = C.c_int(0), C.c_int(0)
        _native.check(
            _native.lib().cco_relax_precision_state(self.handle, C.byref(mode), C.byref(lanes))
        )
        return bool(mode.value), int(lanes.value)

    def pack(self, s0: Sequence[np.ndarray]) -> np.ndarray:
        if len(s0) != self.b:
            raise ValueError("need one configuration per lane")
        parts = []
        for i, s in enumerate(s0):
            require_float64(s, "s0")
            s = np.asarray(s)
            if s.shape != self.shapes[i]:
                raise ValueError(f"s0[{i}] must have shape {self.shapes[i]}")
            parts.append(np.ascontiguousarray(s).ravel())
        return np.ascontiguousarray(np.concatenate(parts))

    def unpack(self, flat: np.ndarray) -> List[np.ndarray]:
        out, off = [], 0
        for n, d in self.shapes:
            out.append(flat[off: off + n * d].reshape(n, d).copy())
            off += n * d
        return out

    def init(self, s0: Sequence[np.ndarray]) -> None:
        packed = self.pack(s0)
        _native.check(_native.lib().cco_relax_init(self.handle, _native.dptr(packed)))

    def run(self, max_evals: int, sync_every: int) -> None:
        _native.check(_native.lib().cco_relax_run(self.handle, int(max_evals), int(sync_every)))

    def result(self) -> List[RelaxResult]:
        b = self.b
        L = _native.lib()
        s = np.empty(self.size)
        e = np.empty(b)
        gn = np.empty(b)
        ne = np.empty(b, dtype=np.int64)
        ni = np.empty(b, dtype=np.int64)
        st = np.empty(b, dtype=np.int64)
        _native.check(
            L.cco_relax_result(
                self.handle, _native.dptr(s), _native.dptr(e), _native.dptr(gn), _native.i64ptr(ne),
                _native.i64ptr(ni), _native.i64ptr(st),
            )
        )
        every, cap = self.settings.record_every, self.settings.max_records
        tr = None
        if every > 0 and cap > 0:
            cnt = np.empty(b, dtype=np.int64)
            it = np.zeros((b, cap), dtype=np.int64)
            ts = np.zeros(cap * self.size)
            te = np.zeros((b, cap))
            tt = np.zeros((b, cap))
            _native.check(
                L.cco_relax_trajectory(
                    self.handle, _native.i64ptr(cnt), _native.i64ptr(it), _native.dptr(ts),
                    _native.dptr(te), _native.dptr(tt),
                )
            )
            tr = (cnt, it, ts, te, tt)
        out: List[RelaxResult] = []
        off = 0
        for i, (n_i, d_i) in enumerate(self.shapes):
            nd = n_i * d_i
            traj = None
            if tr is not None:
                cnt, it, ts, te, tt = tr
                rows = ts[cap * off: cap * (off + nd)].reshape(cap, n_i, d_i)
                traj = Trajectory(s=rows, energy=te[i], step=tt[i], iteration=it[i], count=cnt[i])
            out.append(
                RelaxResult(
                    s=s[off: off + nd].reshape(n_i, d_i).copy(), energy=e[i], grad_norm=gn[i],
                    n_evals=ne[i], n_iters=ni[i], status=st[i], trajectory=traj,
                )
            )
            off += nd
        return out

    def close(self) -> None:
        if self.handle is not None and _native._lib is not None:
            _native._lib.cco_relax_destroy(self.handle)
        self.handle = None

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def relax_uneven(
    energies: Sequence[StealthyEnergy],
    s0: Sequence[np.ndarray],
    settings: Settings,
    chunk_iters: int = 2_000,
) -> List[RelaxResult]:
    """Relax configurations of different size, dimension and k-set in one batch.

    ``s0[i]`` (shape (n_i, d_i)) relaxes under ``energies[i]`` (the same
    energy object may serve several configurations).  Every configuration is
    a lane of one device session, so each kernel launch covers all of them;
    per-lane freezing and compaction work as in :func:`relax`.  Returns one
    :class:`RelaxResult` per configuration (leaves without a lane axis).

    The reference relaxes one (N, d, k-set) per call (its CLI loops over
    them, cli/relax.py:339-358); the per-configuration arithmetic is the same
    as :func:`relax` with a single lane, up to summation order.
    """
    if len(energies) != len(s0):
        raise ValueError("need one energy per configuration")
    for i, s in enumerate(s0):
        require_float64(s, "s0")
        if np.ndim(s) != 2:
            raise ValueError(f"s0[{i}] must have shape (n, d)")
    assert chunk_iters >= 1, "chunk_iters must be positive"
    sess = UnevenSession(energies, [np.shape(s) for s in s0], settings)
    try:
        sess.init(s0)
        sess.run(_eval_bound(settings, chunk_iters), max(1, min(chunk_iters, 256)))
        return sess.result()
    finally:
        sess.close()
