"""MD driver: tracked waves, forces, lane state, segments and chunked device runs.

Restates src/xla_cco/dynamics/run.py.  ``simulate`` hands a batch of lanes to
one device session (``cco_md_*``) that steps every running lane on the GPU:
the scheme ("verlet", "andersen", "baoab"), forces at s = r inv(B)
(F = -g inv(B)^T), the tracked densities (constrained waves from the force
pass, the others every ``wind_every`` steps), COM correction unless
"verlet", winding, running sums, snapshots at ``record_steps``, the dense rho
trace and the drift guard ("verlet" only); a lane whose status leaves
STATUS_RUNNING keeps every leaf from then on.  The host loop runs chunks of
``chunk_steps`` and writes checkpoints between them.

Host-side rules kept from the reference:

- ``begin_segment`` with ``reset_winding`` marks rho_prev pending (NaN); the
  next ``simulate`` / ``md_chunk`` that takes a step recomputes those rows
  from the current positions with the COM rule of its scheme first
  (run.py:506-518); stopped lanes are left untouched by both.
- The drift guard bounds |E - e0| by drift_tol d (N - 1) T / 2 at every step;
  drift_tol = inf disables it.
- "Latest rho" (snapshots, trace, S sums) is rho_prev after the step.

Random draws come from Philox keys (xla_cco.dynamics.rng) split per step as
the reference splits its JAX keys: (key', k_xi) for BAOAB and (key', k_pick,
k_draw, k_coin) for Andersen.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import math
import time
from pathlib import Path
from typing import Any, Dict, Optional, Sequence, Tuple

import numpy as np

from xla_cco import _native
from xla_cco._types import require_float64
from xla_cco.dynamics.checkpoint import save_state
from xla_cco.dynamics.integrators import ForceFn
from xla_cco.dynamics.observables import center_of_mass, com_correct, wave_densities
from xla_cco.dynamics.state import (
    MDResult,
    MDSettings,
    MDState,
    RhoTrace,
    Snapshots,
    Waves,
    check_settings,
    lane,
    leaves,
    stack,
    with_leaves,
)
from xla_cco.energy import StealthyEnergy
from xla_cco.kvectors import enumerate_kvectors
from xla_cco.minimize.result import STATUS_RUNNING

MAX_RECORD_STEPS = 256
CHECKPOINT_SOURCE = "xla_cco.dynamics.run.simulate"
_SCHEME_CODE = {"verlet": 0, "andersen": 1, "baoab": 2}
_MODE_CODE = {"count": 0, "bernoulli": 1}


# ----- checks -----
def _check_lane(state: MDState, waves: Waves, settings: MDSettings) -> None:
    """Raise ValueError unless one lane's buffers fit the waves and settings (run.py:136-155)."""
    w = int(waves.kvec.shape[0])
    if state.r.ndim != 2 or state.rho_prev.shape != (w, 2):
        raise ValueError(
            f"need one lane with (n, d) positions and ({w}, 2) rho, got "
            f"{state.r.shape} and {state.rho_prev.shape}"
        )
    if state.snap.step.shape != (len(settings.record_steps),):
        raise ValueError(
            f"snapshot capacity {state.snap.step.shape} != {len(settings.record_steps)} record_steps"
        )
    if state.trace.rho.shape[:2] != (settings.max_rho_records, w):
        raise ValueError(
            f"trace capacity {state.trace.rho.shape[:2]} != ({settings.max_rho_records}, {w})"
        )


def _check_waves(energy: StealthyEnergy, waves: Waves) -> None:
    """The constrained waves must be the energy's first n_constrained n_vectors (run.py:158-180).

    The device reads the constrained densities from the force pass, so any
    other wave set would silently take the energy's rows as its rho.
    """
    e, w, nc = np.asarray(energy.n_vectors), np.asarray(waves.n_vectors), waves.n_constrained
    assert w.ndim == 2 and 1 <= nc <= w.shape[0], "waves hold the constrained rows"
    if e.ndim != 2 or e.shape[0] < nc or e.shape[1] != w.shape[1]:
        raise ValueError(
            f"energy n_vectors {e.shape} cannot hold the {nc} constrained waves of dimension {w.shape[1]}"
        )
    if not np.array_equal(e[:nc], w[:nc]):
        raise ValueError(
            f"the {nc} constrained waves differ from the energy's first {nc} n_vectors (use one k-set)"
        )


# ----- waves and forces -----
def make_waves(basis: np.ndarray, k_constraint: float, k_track: float) -> Waves:
    """Tracked waves: the constrained set (|k| < k_constraint, k-set order), then
    the half-space waves with k_constraint < |k| < k_track by |k| (run.py:188-226).

    Both cutoffs are strict; kvec = n @ 2 pi inv(B)^T.  Raises ValueError
    unless 0 < k_constraint <= k_track < inf, on a bad basis, or when no wave
    has |k| < k_constraint.
    """
    b = np.asarray(basis, dtype=np.float64)
    if not (0.0 < k_constraint <= k_track < math.inf):
        raise ValueError(f"need 0 < k_constraint <= k_track < inf, got {k_constraint}, {k_track}")
    inner = enumerate_kvectors(b, 0.0, k_constraint)
    if inner.n.shape[0] == 0:
        raise ValueError(f"no constrained wave with |k| < {k_constraint}")
    blocks = [inner.n]
    if k_track > k_constraint:
        blocks.append(enumerate_kvectors(b, k_constraint, k_track).n)
    n_vec = np.concatenate(blocks).astype(np.float64)
    kvec = n_vec @ (2.0 * math.pi * np.linalg.inv(b).T)
    m = int(inner.n.shape[0])
    return Waves(
        n_vectors=n_vec,
        kvec=kvec,
        constrained=np.arange(n_vec.shape[0]) < m,
        n_constrained=m,
    )


def _basis_inverse(energy: StealthyEnergy, basis: np.ndarray) -> np.ndarray:
    b = np.asarray(basis, dtype=np.float64)
    d = energy.dim
    if b.shape != (d, d):
        raise ValueError(f"basis must have shape ({d}, {d}), got {b.shape}")
    return np.linalg.inv(b)


def make_force_fn(energy: StealthyEnergy, basis: np.ndarray) -> ForceFn:
    """r -> (Phi, F, re, im) on the GPU at s = r inv(B), F = -g inv(B)^T (run.py:229-256)."""
    inv = _basis_inverse(energy, basis)
    inv_t = inv.T
    d = energy.dim

    def force_fn(r: np.ndarray) -> Tuple[float, np.ndarray, np.ndarray, np.ndarray]:
        assert r.ndim == 2 and r.shape[1] == d, "r must have shape (n, d)"
        phi, g, re, im = energy.value_grad_density(np.ascontiguousarray(r @ inv))
        return phi, -g @ inv_t, re, im

    return force_fn


def _forces_batch(energy: StealthyEnergy, inv: np.ndarray, r: np.ndarray):  # type: ignore[no-untyped-def]
    """(Phi (b,), F (b, n, d), re, im (b, m)) of a batch of positions in one device pass."""
    phi, g, re, im = energy.value_grad_density_batch(np.ascontiguousarray(r @ inv))
    return phi, -g @ inv.T, re, im


def _lattice_rows(basis: np.ndarray, kvec: np.ndarray) -> Optional[np.ndarray]:
    """n = B k / (2 pi) per row when every row is integral (to 1e-9) and nonzero, else None."""
    x = np.asarray(kvec) @ np.asarray(basis, dtype=np.float64).T / (2.0 * math.pi)
    n = np.rint(x)
    if not np.all(np.abs(x - n) <= 1e-9 * np.maximum(1.0, np.abs(n))) or np.any(np.all(n == 0.0, axis=1)):
        return None
    return n


def _unconstrained_rho(energy: StealthyEnergy, waves: Waves, basis: np.ndarray, inv: np.ndarray,
                       r: np.ndarray) -> np.ndarray:
    """rho (b, w - nc, 2) of the unconstrained waves for a batch of positions.

    Waves on the reciprocal lattice go through the device density pass over
    their own k-set (exp(i k.r) = exp(2 pi i n.s)), as the MD step does;
    others, or a set too wide for that pass, are summed on the host.
    """
    nc = waves.n_constrained
    kv = np.asarray(waves.kvec[nc:])
    n_t = _lattice_rows(basis, kv)
    if n_t is not None:
        try:
            t = StealthyEnergy(n_t, np.ones(n_t.shape[0]), 1.0, device=energy.device)
            _, _, re, im = t._eval(np.ascontiguousarray(r @ inv), grad=False, density=True)
        except ValueError:
            pass  # too wide for the density pass's tables
        else:
            assert re is not None and im is not None
            return np.stack([re, im], axis=-1)
    return np.stack([wave_densities(r[i], kv, energy.chunk_size) for i in range(r.shape[0])])


def _lane_rho(r: np.ndarray, re: np.ndarray, im: np.ndarray, waves: Waves, scheme: str,
              rho_u: np.ndarray) -> np.ndarray:
    """rho (w, 2) of one lane with every wave fresh (run.py:259-291, wind = True);
    rho_u holds the unconstrained waves' densities."""
    nc = waves.n_constrained
    w = int(waves.kvec.shape[0])
    rho = np.empty((w, 2))
    rho[:nc, 0] = re[:nc]
    rho[:nc, 1] = im[:nc]
    if w > nc:
        rho[nc:] = rho_u
    if scheme != "verlet":
        rho = com_correct(rho, np.asarray(waves.kvec), center_of_mass(r))
    return rho


# ----- state -----
def _empty_buffers(lead: Tuple[int, ...], n: int, d: int, w: int, settings: MDSettings) -> Tuple[Snapshots, RhoTrace]:
    n_rec, n_tr = len(settings.record_steps), settings.max_rho_records
    assert n_rec <= MAX_RECORD_STEPS and n_tr >= 0, "validated capacities"
    snap = Snapshots(
        step=np.full(lead + (n_rec,), -1, np.int64),
        time=np.zeros(lead + (n_rec,)),
        wind_abs=np.zeros(lead + (n_rec, w)),
        wind_net=np.zeros(lead + (n_rec, w)),
        rho=np.zeros(lead + (n_rec, w, 2)),
        r=np.zeros(lead + (n_rec, n, d)),
        phi=np.zeros(lead + (n_rec,)),
        kinetic=np.zeros(lead + (n_rec,)),
        count=np.zeros(lead, np.int64),
    )
    trace = RhoTrace(rho=np.zeros(lead + (n_tr, w, 2)), time=np.zeros(lead + (n_tr,)),
                     count=np.zeros(lead, np.int64))
    return snap, trace


def init_state(
    energy: StealthyEnergy,
    waves: Waves,
    basis: np.ndarray,
    r0: np.ndarray,
    v0: np.ndarray,
    temperature,
    key: np.ndarray,
    settings: MDSettings,
) -> MDState:
    """Lanes at step 0: forces, rho_prev, zero winding and sums, empty buffers (run.py:300-352).

    One lane for r0 / v0 of shape (n, d) and key (2,); a batch for (b, n, d)
    with keys (b, 2) and a scalar or (b,) temperature (the reference builds
    batches with jax.vmap; here one device pass evaluates every lane).
    e0 = Phi + KE(v0) with the velocities as given (``begin_segment`` removes
    the COM momentum); rho_prev follows the COM rule of ``settings.scheme``.
    Raises ValueError on invalid settings or shapes, or constrained waves
    that are not the energy's first n_vectors; TypeError on non-float64
    positions/velocities or a bad key.
    """
    check_settings(settings)
    r0 = np.asarray(r0)
    v0 = np.asarray(v0)
    require_float64(r0, "r0")
    require_float64(v0, "v0")
    if r0.ndim not in (2, 3) or v0.shape != r0.shape:
        raise ValueError(f"r0 and v0 must share an (n, d) or (b, n, d) shape: {r0.shape}, {v0.shape}")
    single = r0.ndim == 2
    rb = r0[None] if single else r0
    vb = v0[None] if single else v0
    b, n, d = rb.shape
    keys = np.asarray(key)
    keys = keys[None] if single else keys
    if keys.dtype != np.uint32 or keys.shape != (b, 2):
        raise TypeError(f"key must be uint32 of shape {'(2,)' if single else f'({b}, 2)'}, got {keys.dtype} {np.shape(key)}")
    temp = np.broadcast_to(np.asarray(temperature, dtype=np.float64), (b,)).copy()
    w = int(waves.kvec.shape[0])
    if waves.kvec.shape != (w, d):
        raise ValueError(f"waves have dimension {waves.kvec.shape}, positions {d}")
    _check_waves(energy, waves)
    inv = _basis_inverse(energy, basis)
    phi, f, re, im = _forces_batch(energy, inv, rb)
    rho_u = _unconstrained_rho(energy, waves, basis, inv, rb)
    rho = np.stack([_lane_rho(rb[i], re[i], im[i], waves, settings.scheme, rho_u[i]) for i in range(b)])
    snap, trace = _empty_buffers((b,), n, d, w, settings)
    ke = 0.5 * np.sum(vb * vb, axis=(1, 2))
    zf = np.zeros(b)
    zi = np.zeros(b, np.int64)
    st = MDState(
        r=np.array(rb), v=np.array(vb), f=f, phi=phi, temperature=temp, key=keys.copy(), step=zi.copy(),
        time=zf.copy(), reset_acc=zf.copy(), rho_prev=rho, wind_abs=np.zeros((b, w)),
        wind_net=np.zeros((b, w)), n_valid=np.zeros((b, w), np.int64), e0=phi + ke,
        max_drift=zf.copy(), sum_kinetic=zf.copy(), sum_phi=zf.copy(), sum_phi2=zf.copy(),
        sum_s=np.zeros((b, w)), n_sum=zi.copy(), status=np.full(b, STATUS_RUNNING, np.int64),
        snap=snap, trace=trace,
    )
    return lane(st, 0) if single else st


def begin_segment(state: MDState, settings: MDSettings, reset_sums: bool, reset_winding: bool) -> MDState:
    """Start a segment (run.py:355-389), one lane or a batch.

    Optional COM momentum removal (settings.remove_com), e0 = Phi + KE,
    max_drift = 0, optional reset of the sums and of the winding (which marks
    rho_prev pending, NaN, for the next chunk to recompute).  Positions,
    forces, key, step and buffers are unchanged; a lane whose status is not
    STATUS_RUNNING is returned unchanged.
    """
    v = state.v
    if settings.remove_com:
        v = v - np.mean(v, axis=-2, keepdims=True)
    ke = 0.5 * np.sum(v * v, axis=(-2, -1))
    kw: Dict[str, Any] = dict(v=v, e0=state.phi + ke, max_drift=np.zeros_like(state.phi))
    if reset_sums:
        kw.update(sum_kinetic=np.zeros_like(state.sum_kinetic), sum_phi=np.zeros_like(state.sum_phi),
                  sum_phi2=np.zeros_like(state.sum_phi2), sum_s=np.zeros_like(state.sum_s),
                  n_sum=np.zeros_like(state.n_sum))
    if reset_winding:
        kw.update(wind_abs=np.zeros_like(state.wind_abs), wind_net=np.zeros_like(state.wind_net),
                  n_valid=np.zeros_like(state.n_valid), rho_prev=np.full_like(state.rho_prev, np.nan))
    new = dataclasses.replace(state, **kw)
    running = np.asarray(state.status) == STATUS_RUNNING
    out = []
    for old_x, new_x in zip(leaves(state), leaves(new)):
        if old_x is new_x:
            out.append(old_x)
            continue
        sel = running.reshape(running.shape + (1,) * (np.ndim(new_x) - running.ndim))
        out.append(np.where(sel, new_x, old_x).astype(np.asarray(old_x).dtype))
    return with_leaves(state, out)


# ----- the device session -----
class MDSession:
    """One cco_md device session: a batch of lanes with its waves and settings.

    ``put(state)`` uploads a batched MDState, ``run(n)`` advances every running
    lane by n steps on the device, ``get(like)`` reads the state back (``like``
    gives the host arrays' shapes); ``simulate`` drives one per call.
    """

    def __init__(self, energy: StealthyEnergy, waves: Waves, basis: np.ndarray, settings: MDSettings,
                 b: int, n: int, d: int) -> None:
        ctx = _native.context(energy.device)
        bas = np.ascontiguousarray(np.asarray(basis, dtype=np.float64))
        h = C.c_void_p()
        _native.check(_native.lib().cco_md_create(ctx.handle, energy.handle(), b, n, d, _native.dptr(bas),
                                                  float(settings.dt), float(settings.drift_tol), C.byref(h)))
        self.handle: Optional[C.c_void_p] = h
        try:
            if settings.precision == "mixed":
                _native.check(_native.lib().cco_relax_set_precision(h, 1, 0.0))
            ms = _native.MDSettingsC(
                andersen_p=float(settings.andersen_p), gamma=float(settings.gamma),
                phase_floor=float(settings.phase_floor), scheme=_SCHEME_CODE[settings.scheme],
                andersen_mode=_MODE_CODE[settings.andersen_mode], wind_every=int(settings.wind_every),
                record_rho_every=int(settings.record_rho_every), max_rho_records=int(settings.max_rho_records),
                n_record_steps=len(settings.record_steps),
            )
            kvec = np.ascontiguousarray(np.asarray(waves.kvec, dtype=np.float64))
            rec = np.ascontiguousarray(np.asarray(settings.record_steps, dtype=np.int64).reshape(-1))
            _native.check(_native.lib().cco_md_configure(h, C.byref(ms), int(kvec.shape[0]), int(waves.n_constrained),
                                                         _native.dptr(kvec), _native.i64ptr(rec)))
        except BaseException:
            self.close()
            raise
        self.b = b

    def close(self) -> None:
        if self.handle is not None and _native._lib is not None:
            _native._lib.cco_relax_destroy(self.handle)
        self.handle = None

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _io(st: MDState) -> Tuple["_native.MDStateIOC", Dict[str, np.ndarray]]:
        """The C view of a batched state; the arrays it points into (kept alive by the caller)."""
        b = st.r.shape[0]
        kin = 0.5 * np.sum(st.v * st.v, axis=(1, 2))
        a: Dict[str, np.ndarray] = {
            "r": st.r, "v": st.v, "f": st.f,
            "lane_f64": np.stack([st.phi, st.temperature, st.time, st.reset_acc, st.e0, st.max_drift,
                                  st.sum_kinetic, st.sum_phi, st.sum_phi2, kin], axis=1),
            "lane_i64": np.stack([st.step, st.n_sum, st.status, st.snap.count, st.trace.count], axis=1),
            "key": st.key, "rho_prev": st.rho_prev, "wind_abs": st.wind_abs, "wind_net": st.wind_net,
            "n_valid": st.n_valid, "sum_s": st.sum_s, "snap_step": st.snap.step, "snap_time": st.snap.time,
            "snap_phi": st.snap.phi, "snap_kinetic": st.snap.kinetic, "snap_wind_abs": st.snap.wind_abs,
            "snap_wind_net": st.snap.wind_net, "snap_rho": st.snap.rho, "snap_r": st.snap.r,
            "trace_rho": st.trace.rho, "trace_time": st.trace.time,
        }
        for k, x in a.items():
            want = np.uint32 if k == "key" else (np.int64 if k in ("lane_i64", "n_valid", "snap_step") else np.float64)
            x = np.ascontiguousarray(np.array(x, dtype=want, copy=True))
            if x.ndim == 0 or x.shape[0] != b:
                raise ValueError(f"{k} needs a leading axis of {b} lanes, got shape {x.shape}")
            a[k] = x
        io = _native.MDStateIOC()
        for k, x in a.items():
            ptr = _native.u32ptr(x) if k == "key" else (_native.i64ptr(x) if x.dtype == np.int64 else _native.dptr(x))
            setattr(io, k, ptr)
        return io, a

    def put(self, st: MDState) -> None:
        io, keep = self._io(st)
        _native.check(_native.lib().cco_md_put(self.handle, C.byref(io)))
        del keep

    def get(self, like: MDState) -> MDState:
        io, a = self._io(like)
        _native.check(_native.lib().cco_md_get(self.handle, C.byref(io)))
        f64, i64 = a["lane_f64"], a["lane_i64"]
        return MDState(
            r=a["r"], v=a["v"], f=a["f"], phi=f64[:, 0].copy(), temperature=f64[:, 1].copy(), key=a["key"],
            step=i64[:, 0].copy(), time=f64[:, 2].copy(), reset_acc=f64[:, 3].copy(), rho_prev=a["rho_prev"],
            wind_abs=a["wind_abs"], wind_net=a["wind_net"], n_valid=a["n_valid"], e0=f64[:, 4].copy(),
            max_drift=f64[:, 5].copy(), sum_kinetic=f64[:, 6].copy(), sum_phi=f64[:, 7].copy(),
            sum_phi2=f64[:, 8].copy(), sum_s=a["sum_s"], n_sum=i64[:, 1].copy(), status=i64[:, 2].copy(),
            snap=Snapshots(step=a["snap_step"], time=a["snap_time"], wind_abs=a["snap_wind_abs"],
                           wind_net=a["snap_wind_net"], rho=a["snap_rho"], r=a["snap_r"], phi=a["snap_phi"],
                           kinetic=a["snap_kinetic"], count=i64[:, 3].copy()),
            trace=RhoTrace(rho=a["trace_rho"], time=a["trace_time"], count=i64[:, 4].copy()),
        )

    def run(self, n_steps: int) -> int:
        na = C.c_int(0)
        _native.check(_native.lib().cco_md_run(self.handle, int(n_steps), max(1, min(int(n_steps), 1024)),
                                               C.byref(na)))
        return int(na.value)


def _repair_pending(energy: StealthyEnergy, waves: Waves, basis: np.ndarray, settings: MDSettings,
                    st: MDState) -> MDState:
    """Recompute rho_prev rows marked pending (NaN) in running lanes (run.py:506-518)."""
    pend = np.isnan(st.rho_prev[:, :, 0]) & (st.status == STATUS_RUNNING)[:, None]
    lanes = np.nonzero(pend.any(axis=1))[0]
    if lanes.size == 0:
        return st
    inv = _basis_inverse(energy, basis)
    _, _, re, im = _forces_batch(energy, inv, st.r[lanes])
    rho_u = _unconstrained_rho(energy, waves, basis, inv, st.r[lanes])
    rho_prev = st.rho_prev.copy()
    for q, i in enumerate(lanes):
        rho = _lane_rho(st.r[i], re[q], im[q], waves, settings.scheme, rho_u[q])
        rho_prev[i] = np.where(pend[i][:, None], rho, rho_prev[i])
    return dataclasses.replace(st, rho_prev=rho_prev)


def _check_batch(energy: StealthyEnergy, waves: Waves, settings: MDSettings, state: MDState) -> None:
    check_settings(settings)
    if np.ndim(state.step) != 1:
        raise ValueError(f"state must be batched (leading lane axis), got step shape {np.shape(state.step)}")
    if state.step.shape[0] < 1:
        raise ValueError("need at least one lane")
    _check_lane(lane(state, 0), waves, settings)
    _check_waves(energy, waves)
    if state.r.shape[2] != energy.dim:
        raise ValueError(f"positions have dimension {state.r.shape[2]}, the energy {energy.dim}")


def _write(path: Path, state: MDState, done: int, n_steps: int) -> None:
    meta: Dict[str, Any] = {"source": CHECKPOINT_SOURCE, "steps_done": done, "n_steps": n_steps}
    save_state(Path(path), state, meta)


def simulate(
    energy: StealthyEnergy,
    waves: Waves,
    basis: np.ndarray,
    settings: MDSettings,
    state: MDState,
    n_steps: int,
    checkpoint: Optional[Path] = None,
    checkpoint_every_s: float = 1800.0,
) -> MDResult:
    """Run the batched lanes for n_steps on the device (run.py:599-654).

    Chunks of ``settings.chunk_steps`` steps; stops early when no lane is
    running.  With ``checkpoint`` the state is written there atomically
    (overwriting, never reading) whenever ``checkpoint_every_s`` seconds
    have passed after a chunk, and at the end.  ``step`` is absolute;
    snapshots and the trace persist across calls.  Raises ValueError on
    invalid settings, an unbatched state, buffers that do not match the
    settings, constrained waves that are not the energy's first n_vectors,
    n_steps < 0 or checkpoint_every_s <= 0.
    """
    _check_batch(energy, waves, settings, state)
    if int(n_steps) != n_steps or n_steps < 0 or not checkpoint_every_s > 0.0:
        raise ValueError(f"need n_steps >= 0 and checkpoint_every_s > 0, got {n_steps}, {checkpoint_every_s}")
    n_steps = int(n_steps)
    done, last = 0, time.monotonic()
    dev: Optional[MDSession] = None
    try:
        while done < n_steps and np.any(state.status == STATUS_RUNNING):
            n_this = min(settings.chunk_steps, n_steps - done)
            if dev is None:
                state = _repair_pending(energy, waves, basis, settings, state)
                b, n, d = state.r.shape
                dev = MDSession(energy, waves, basis, settings, b, n, d)
                dev.put(state)
            n_active = dev.run(n_this)
            done += n_this
            due = time.monotonic() - last >= checkpoint_every_s
            if n_active == 0 or (checkpoint is not None and done < n_steps and due):
                state = dev.get(state)
                if checkpoint is not None and done < n_steps and due:
                    _write(checkpoint, state, done, n_steps)
                    last = time.monotonic()
                if n_active == 0:
                    break
        if dev is not None:
            state = dev.get(state)
    finally:
        if dev is not None:
            dev.close()
    if checkpoint is not None:
        _write(checkpoint, state, done, n_steps)
    trace = state.trace if settings.record_rho_every > 0 else None
    return MDResult(state=state, snapshots=state.snap, trace=trace)


def md_chunk(energy: StealthyEnergy, waves: Waves, basis: np.ndarray, settings: MDSettings,
             state: MDState, n_steps: int) -> MDState:
    """Advance one lane by min(settings.chunk_steps, n_steps) steps (run.py:537-551).

    Pending rho_prev rows are recomputed first; a lane whose status is not
    STATUS_RUNNING (or turns STATUS_DRIFT in a step) keeps every leaf.
    """
    if state.batched:
        raise ValueError("md_chunk advances one lane (simulate runs batches)")
    n_do = min(int(settings.chunk_steps), int(n_steps))
    return lane(simulate(energy, waves, basis, settings, stack([state]), max(n_do, 0)).state, 0)


# ----- dt selection -----
def select_dt(
    energy: StealthyEnergy,
    waves: Waves,
    basis: np.ndarray,
    state: MDState,
    ladder: Sequence[float],
    n_steps: int,
    drift_rel: float,
) -> float:
    """Largest dt in ``ladder`` whose NVE drift |E_last - E_first| / <KE> over n_steps
    stays below ``drift_rel`` in every lane (run.py:674-702).

    Each trial restarts the lanes (r, v, temperature, key of the batched
    ``state``) with ``init_state`` under velocity Verlet (velocities as
    given, no drift abort, no recording); <KE> is the mean over the steps.
    Non-finite energies fail.  Raises ValueError on an empty ladder, a
    non-positive dt, n_steps < 1, drift_rel <= 0, an unbatched state, or when
    no dt meets the criterion.
    """
    dts = sorted({float(x) for x in ladder}, reverse=True)
    if not dts or not all(math.isfinite(x) and x > 0.0 for x in dts):
        raise ValueError(f"ladder must hold finite positive steps, got {ladder}")
    if n_steps < 1 or not drift_rel > 0.0 or np.ndim(state.step) != 1:
        raise ValueError("need n_steps >= 1, drift_rel > 0 and a batched state")
    for dt in dts:
        st = MDSettings(dt=dt, scheme="verlet", remove_com=False, chunk_steps=min(int(n_steps), 10_000),
                        drift_tol=math.inf)
        lanes = init_state(energy, waves, basis, state.r, state.v, state.temperature, state.key, st)
        out = simulate(energy, waves, basis, st, lanes, int(n_steps)).state
        e_last = out.phi + 0.5 * np.sum(out.v * out.v, axis=(1, 2))
        ke = out.sum_kinetic / np.maximum(out.n_sum, 1)
        with np.errstate(invalid="ignore", divide="ignore"):
            rel = np.abs(e_last - lanes.e0) / ke
        if np.all(rel < drift_rel):
            return dt
    raise ValueError(f"no dt in {tuple(ladder)} keeps the drift below {drift_rel}")


__all__ = [
    "MAX_RECORD_STEPS",
    "MDSession",
    "begin_segment",
    "init_state",
    "make_force_fn",
    "make_waves",
    "md_chunk",
    "select_dt",
    "simulate",
]
