"""Device MD (schemes, thermostats, tracked waves, records) against the oracle.

Mirrors tests/unit/test_md_run.py and test_md_thermostats.py of the
reference: every scheme over several chunks with tracked waves (constrained
and every-other-step ones), snapshots and the rho trace; lanes at different
temperatures, one stopped by the drift guard; begin_segment with a pending
rho_prev; checkpoints; md_chunk; select_dt.  The device and the oracle draw
the same Philox numbers, so trajectories, keys and counters agree lane by lane.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np
import pytest

from oracle import oracle as O
from xla_cco import _native
from xla_cco import kvectors as K
from xla_cco.box import cubic_box
from xla_cco.dynamics import (
    MDSession,
    MDSettings,
    begin_segment,
    init_state,
    lane,
    load_state,
    make_waves,
    md_chunk,
    prng_key,
    read_meta,
    select_dt,
    simulate,
)
from xla_cco.dynamics.state import leaves
from xla_cco.energy import make_stealthy_energy
from xla_cco.minimize.result import STATUS_DRIFT, STATUS_RUNNING

pytestmark = pytest.mark.gpu


def _system(n: int = 32, d: int = 2, chi: float = 0.2, track: float = 1.4):
    box = cubic_box(n, d)
    k_c = K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, k_c)
    e = make_stealthy_energy(box, ks)
    waves = make_waves(box.basis, k_c, track * k_c)
    assert waves.kvec.shape[0] > waves.n_constrained
    return e, box.basis, waves, O.make_value_grad(ks.n, ks.weights, box.volume)


def _cfg(st: MDSettings) -> O.MDConfig:
    return O.MDConfig(dt=st.dt, scheme=st.scheme, andersen_p=st.andersen_p, andersen_mode=st.andersen_mode,
                      gamma=st.gamma, wind_every=st.wind_every, record_steps=tuple(st.record_steps),
                      record_rho_every=st.record_rho_every, max_rho_records=st.max_rho_records,
                      phase_floor=st.phase_floor, drift_tol=st.drift_tol)


def _compare(dev, ref, i: int, tol: float = 1e-9) -> None:
    """Lane i of a device state against an oracle lane dict."""
    assert tuple(int(x) for x in dev.key[i]) == ref["key"]
    for name in ("step", "n_sum", "status"):
        assert int(getattr(dev, name)[i]) == ref[name], name
    assert np.array_equal(dev.n_valid[i], ref["n_valid"])
    assert int(dev.snap.count[i]) == ref["snap_count"] and int(dev.trace.count[i]) == ref["trace_count"]
    assert np.array_equal(dev.snap.step[i], ref["snap_step"])

    def close(a, b, what, scale=1.0):
        err = np.max(np.abs(np.asarray(a) - np.asarray(b))) if np.size(a) else 0.0
        assert err <= tol * scale, f"{what}: {err}"

    close(dev.r[i], ref["r"], "r")
    close(dev.v[i], ref["v"], "v")
    close(dev.f[i], ref["f"], "f", 10.0)
    n = dev.r.shape[1]
    close(dev.phi[i], ref["phi"], "phi", max(1.0, abs(ref["phi"])))
    close(dev.time[i], ref["time"], "time")
    close(dev.reset_acc[i], ref["reset_acc"], "reset_acc")
    close(dev.e0[i], ref["e0"], "e0", max(1.0, abs(ref["e0"])))
    close(dev.rho_prev[i], ref["rho_prev"], "rho_prev", n)
    close(dev.wind_abs[i], ref["wind_abs"], "wind_abs", 100.0)
    close(dev.wind_net[i], ref["wind_net"], "wind_net", 100.0)
    close(dev.sum_s[i], ref["sum_s"], "sum_s", n * max(1, ref["n_sum"]))
    for name in ("sum_kinetic", "sum_phi", "sum_phi2", "max_drift"):
        close(getattr(dev, name)[i], ref[name], name, max(1.0, abs(ref[name])))
    close(dev.snap.r[i], ref["snap_r"], "snap.r")
    close(dev.snap.rho[i], ref["snap_rho"], "snap.rho", n)
    close(dev.snap.wind_net[i], ref["snap_wind_net"], "snap.wind_net", 100.0)
    close(dev.snap.phi[i], ref["snap_phi"], "snap.phi", 10.0)
    close(dev.snap.kinetic[i], ref["snap_kinetic"], "snap.kinetic", 10.0)
    close(dev.snap.time[i], ref["snap_time"], "snap.time")
    close(dev.trace.rho[i], ref["trace_rho"], "trace.rho", n)
    close(dev.trace.time[i], ref["trace_time"], "trace.time")


SCHEMES = [
    dict(scheme="verlet"),
    dict(scheme="andersen", andersen_p=0.3, andersen_mode="count"),
    dict(scheme="andersen", andersen_p=0.4, andersen_mode="bernoulli"),
    dict(scheme="baoab", gamma=0.8),
]


@pytest.mark.parametrize("kw", SCHEMES, ids=lambda kw: f"{kw['scheme']}-{kw.get('andersen_mode', '')}")
def test_schemes_match_oracle(gpu_ctx, kw) -> None:
    e, basis, waves, vg = _system()
    st = MDSettings(dt=0.02, chunk_steps=7, wind_every=2, record_steps=(3, 10, 25), record_rho_every=4,
                    max_rho_records=5, drift_tol=0.5, **kw)
    rng = np.random.default_rng(0)
    B, n, d = 3, 32, 2
    r0 = rng.uniform(size=(B, n, d)) @ basis
    temps = np.array([0.02, 0.05, 1e-14])  # the last lane trips the Verlet drift guard
    v0 = rng.normal(size=(B, n, d)) * np.sqrt(np.array([0.02, 0.05, 0.05]))[:, None, None]
    keys = np.stack([prng_key(100 + i) for i in range(B)])
    state = init_state(e, waves, basis, r0, v0, temps, keys, st)
    out = simulate(e, waves, basis, st, state, 30).state
    cfg = _cfg(st)
    nc = waves.n_constrained
    for i in range(B):
        ref = O.md_run(vg, basis, O.md_lane(vg, basis, r0[i], v0[i], temps[i], tuple(int(x) for x in keys[i]),
                                            waves.kvec, cfg), waves.kvec, nc, cfg, 30)
        _compare(out, ref, i)
    if st.scheme == "verlet":
        assert out.status[2] == STATUS_DRIFT and out.step[2] < 30
    assert np.all(out.status[:2] == STATUS_RUNNING) and np.all(out.step[:2] == 30)


def test_segment_with_pending_rho(gpu_ctx) -> None:
    """begin_segment(reset_winding) marks rho_prev pending; the next run recomputes it first."""
    e, basis, waves, vg = _system(n=24)
    st = MDSettings(dt=0.02, scheme="baoab", gamma=0.5, chunk_steps=5, record_rho_every=3, max_rho_records=8)
    rng = np.random.default_rng(1)
    r0 = rng.uniform(size=(2, 24, 2)) @ basis
    v0 = 0.2 * rng.normal(size=(2, 24, 2))
    keys = np.stack([prng_key(7), prng_key(8)])
    first = simulate(e, waves, basis, st, init_state(e, waves, basis, r0, v0, 0.04, keys, st), 9).state
    seg = begin_segment(first, st, reset_sums=True, reset_winding=True)
    assert np.all(np.isnan(seg.rho_prev))
    assert np.array_equal(simulate(e, waves, basis, st, seg, 0).state.rho_prev, seg.rho_prev, equal_nan=True)
    out = simulate(e, waves, basis, st, seg, 8).state
    cfg = _cfg(st)
    for i in range(2):
        ref = O.md_run(vg, basis, O.md_lane(vg, basis, r0[i], v0[i], 0.04, tuple(int(x) for x in keys[i]),
                                            waves.kvec, cfg), waves.kvec, waves.n_constrained, cfg, 9)
        # begin_segment on the oracle lane: COM momentum out, e0 anew, sums and winding zeroed,
        # rho_prev recomputed from the current positions (COM-corrected: baoab)
        v = ref["v"] - ref["v"].mean(axis=0)
        ref.update(v=v, e0=ref["phi"] + 0.5 * float(np.sum(v * v)), max_drift=0.0, sum_kinetic=0.0, sum_phi=0.0,
                   sum_phi2=0.0, sum_s=np.zeros_like(ref["sum_s"]), n_sum=0, wind_abs=np.zeros_like(ref["wind_abs"]),
                   wind_net=np.zeros_like(ref["wind_net"]), n_valid=np.zeros_like(ref["n_valid"]),
                   rho_prev=O.com_rotate(O.tracked_rho(ref["r"], waves.kvec), waves.kvec, ref["r"].mean(axis=0)))
        ref = O.md_run(vg, basis, ref, waves.kvec, waves.n_constrained, cfg, 8)
        _compare(out, ref, i)


def test_checkpoint_and_md_chunk(gpu_ctx, tmp_path) -> None:
    e, basis, waves, _ = _system(n=20)
    st = MDSettings(dt=0.02, scheme="andersen", andersen_p=0.5, chunk_steps=4, record_steps=(2, 9))
    rng = np.random.default_rng(2)
    r0 = rng.uniform(size=(2, 20, 2)) @ basis
    v0 = 0.1 * rng.normal(size=(2, 20, 2))
    keys = np.stack([prng_key(1), prng_key(2)])
    state = init_state(e, waves, basis, r0, v0, 0.01, keys, st)
    path = tmp_path / "md.npz"
    res = simulate(e, waves, basis, st, state, 10, checkpoint=path, checkpoint_every_s=1e-9)
    assert read_meta(path) == {"source": "xla_cco.dynamics.run.simulate", "steps_done": 10, "n_steps": 10}
    back, _ = load_state(path, res.state)
    for a, b in zip(leaves(back), leaves(res.state)):
        assert np.array_equal(a, b)
    assert np.array_equal(back.snap.r, res.state.snap.r) and res.trace is None
    # resuming from the checkpoint equals running on
    more = simulate(e, waves, basis, st, back, 5).state
    cont = simulate(e, waves, basis, st, res.state, 5).state
    assert np.array_equal(more.r, cont.r) and np.array_equal(more.key, cont.key)
    # md_chunk: one lane, at most chunk_steps steps, equal to its lane of the batch
    one = md_chunk(e, waves, basis, st, lane(state, 1), 100)
    four = simulate(e, waves, basis, st, state, 4).state
    assert int(one.step) == 4
    assert np.array_equal(one.r, four.r[1]) and np.array_equal(one.key, four.key[1])
    init_one = init_state(e, waves, basis, r0[1], v0[1], 0.01, keys[1], st)
    assert np.allclose(init_one.rho_prev, state.rho_prev[1], rtol=0, atol=1e-10)
    with pytest.raises(ValueError):
        simulate(e, waves, basis, st, init_one, 3)  # unbatched


def test_waves_must_match_the_energy(gpu_ctx) -> None:
    e, basis, waves, _ = _system(n=20)
    nc = waves.n_constrained
    order = np.r_[np.arange(nc)[::-1], np.arange(nc, waves.kvec.shape[0])]  # constrained rows reversed
    bad = dataclasses.replace(waves, n_vectors=waves.n_vectors[order], kvec=waves.kvec[order])
    st = MDSettings(dt=0.02)
    r0 = np.random.default_rng(3).uniform(size=(1, 20, 2)) @ basis
    with pytest.raises(ValueError):
        init_state(e, bad, basis, r0, np.zeros_like(r0), 0.01, np.stack([prng_key(0)]), st)


@pytest.mark.parametrize("case", ["lattice-d3-2K", "off-lattice"])
def test_unconstrained_waves_match_oracle(gpu_ctx, case) -> None:
    """Waves on the reciprocal lattice take the factorised rho pass (here d = 3 with
    every wave |k| < 2K, the reference CLI's default, over several particle splits);
    waves off it are summed particle by particle.  Both equal the oracle's direct sums."""
    if case == "lattice-d3-2K":
        e, basis, waves, vg = _system(n=300, d=3, chi=0.1, track=2.0)
        n, d, steps = 300, 3, 6
        st = MDSettings(dt=0.02, scheme="andersen", andersen_p=0.5, wind_every=1, record_steps=(4,),
                        record_rho_every=2, max_rho_records=4, drift_tol=math.inf)
    else:
        e, basis, waves, vg = _system(n=24)
        nc = waves.n_constrained
        kvec = waves.kvec.copy()
        kvec[nc:] *= 1.01  # off the lattice
        waves = dataclasses.replace(waves, kvec=kvec)
        n, d, steps = 24, 2, 9
        st = MDSettings(dt=0.02, scheme="baoab", gamma=0.5, wind_every=3, record_rho_every=3, max_rho_records=4)
    rng = np.random.default_rng(6)
    r0 = rng.uniform(size=(2, n, d)) @ basis
    v0 = 0.05 * rng.normal(size=(2, n, d))
    keys = np.stack([prng_key(11), prng_key(12)])
    out = simulate(e, waves, basis, st, init_state(e, waves, basis, r0, v0, 0.0025, keys, st), steps).state
    cfg = _cfg(st)
    for i in range(2):
        ref = O.md_run(vg, basis, O.md_lane(vg, basis, r0[i], v0[i], 0.0025, tuple(int(x) for x in keys[i]),
                                            waves.kvec, cfg), waves.kvec, waves.n_constrained, cfg, steps)
        _compare(out, ref, i)
    assert np.all(out.n_valid[:, waves.n_constrained:] > 0)
    # which path ran: the tracked pass is profiling stage 9
    ses = MDSession(e, waves, basis, st, 2, n, d)
    try:
        ses.put(init_state(e, waves, basis, r0, v0, 0.0025, keys, st))
        _native.check(_native.lib().cco_relax_set_profiling(ses.handle, 1))
        ses.run(3)
        ms = (C.c_double * _native.CCO_PROFILE_STAGES)()
        launches = (C.c_int64 * _native.CCO_PROFILE_STAGES)()
        _native.check(_native.lib().cco_relax_profile(ses.handle, ms, launches))
        assert launches[9] == (3 if case == "lattice-d3-2K" else 0)
    finally:
        ses.close()


def test_tracked_pass_at_full_shape(gpu_ctx) -> None:
    """32 lanes x N = 10000 x d = 3, chi = 0.45, every wave |k| < 2K (108099 waves): the
    tracked pass runs with fewer particle splits than the energy; its densities after two
    steps equal direct sums over the returned positions (sampled waves, two lanes)."""
    d, n, b = 3, 10000, 32
    box = cubic_box(n, d)
    k_c = K.k_outer_from_ka(K.ka_from_chi(0.45, 0.05, d), 0.05, d)
    e = make_stealthy_energy(box, K.enumerate_kvectors(box.basis, 0.0, k_c))
    waves = make_waves(box.basis, k_c, 2.0 * k_c)
    rng = np.random.default_rng(9)
    r0 = rng.uniform(size=(b, n, d)) @ box.basis
    v0 = rng.normal(scale=1e-2, size=(b, n, d))
    keys = np.stack([prng_key(i) for i in range(b)])
    st = MDSettings(dt=0.01, drift_tol=math.inf, wind_every=1)
    out = simulate(e, waves, box.basis, st, init_state(e, waves, box.basis, r0, v0, 1e-4, keys, st), 2).state
    nc = waves.n_constrained
    pick = np.r_[nc, rng.choice(np.arange(nc + 1, waves.kvec.shape[0] - 1), 40, replace=False),
                 waves.kvec.shape[0] - 1]
    for i in (0, b - 1):
        th = out.r[i] @ waves.kvec[pick].T
        want = np.stack([np.cos(th).sum(axis=0), np.sin(th).sum(axis=0)], axis=-1)
        assert np.max(np.abs(out.rho_prev[i, pick] - want)) <= 1e-9 * n
        assert np.all(out.n_valid[i, pick] == 2)


@pytest.mark.parametrize("shape", ["d3-n300", "d3-n10000"])
def test_tracked_pass_split_precision(gpu_ctx, shape) -> None:
    """A split-precision MD session sums the tracked waves (every |k| < 2K) on the MFMA
    kernels as well: after a few steps their densities equal direct sums over the
    returned positions within the split-precision bound 2e-5 sqrt(N), while the FP64
    session's equal them to round-off (so the two sessions took different paths)."""
    if shape == "d3-n300":
        e, basis, waves, _ = _system(n=300, d=3, chi=0.1, track=2.0)
        n, d, b, steps = 300, 3, 2, 4
    else:
        d, n, b, steps = 3, 10000, 32, 2
        box = cubic_box(n, d)
        k_c = K.k_outer_from_ka(K.ka_from_chi(0.45, 0.05, d), 0.05, d)
        e = make_stealthy_energy(box, K.enumerate_kvectors(box.basis, 0.0, k_c))
        basis = box.basis
        waves = make_waves(basis, k_c, 2.0 * k_c)
    rng = np.random.default_rng(9)
    r0 = rng.uniform(size=(b, n, d)) @ basis
    v0 = rng.normal(scale=1e-2, size=(b, n, d))
    keys = np.stack([prng_key(i) for i in range(b)])
    nc, w = waves.n_constrained, waves.kvec.shape[0]
    pick = np.r_[nc, rng.choice(np.arange(nc + 1, w - 1), min(40, w - nc - 2), replace=False), w - 1]
    err = {}
    for prec in ("fp64", "mixed"):
        st = MDSettings(dt=0.01, drift_tol=math.inf, wind_every=1, precision=prec)
        out = simulate(e, waves, basis, st, init_state(e, waves, basis, r0, v0, 1e-4, keys, st), steps).state
        worst = 0.0
        for i in (0, b - 1):
            th = out.r[i] @ waves.kvec[pick].T
            want = np.stack([np.cos(th).sum(axis=0), np.sin(th).sum(axis=0)], axis=-1)
            worst = max(worst, float(np.max(np.abs(out.rho_prev[i, pick] - want))))
            assert np.all(out.n_valid[i, pick] >= 1)
        err[prec] = worst
    print(f"{shape}: tracked rho vs direct sums, FP64 {err['fp64']:.2e}, split {err['mixed']:.2e}", flush=True)
    assert err["fp64"] <= 1e-9 * n
    assert err["mixed"] <= 2e-5 * math.sqrt(n)
    assert err["mixed"] > 100.0 * err["fp64"]  # the split-precision kernels summed them


def test_session_call_order_and_counters(gpu_ctx) -> None:
    """cco_md_run / cco_md_get before a state is put are STATE errors; counters that
    would index past the snapshot or trace rows are rejected before any copy."""
    e, basis, waves, _ = _system(n=20)
    st = MDSettings(dt=0.02, record_steps=(2,), record_rho_every=2, max_rho_records=3, drift_tol=math.inf)
    r0 = np.random.default_rng(5).uniform(size=(2, 20, 2)) @ basis
    state = init_state(e, waves, basis, r0, np.zeros_like(r0), 0.01, np.stack([prng_key(0), prng_key(1)]), st)
    ses = MDSession(e, waves, basis, st, 2, 20, 2)
    try:
        with pytest.raises(RuntimeError, match="cco_md_put must come before"):
            ses.run(3)
        with pytest.raises(RuntimeError, match="cco_md_put must come before"):
            ses.get(state)
        bad = [dataclasses.replace(state, trace=dataclasses.replace(state.trace, count=np.array([0, -1]))),
               dataclasses.replace(state, trace=dataclasses.replace(state.trace, count=np.array([4, 0]))),
               dataclasses.replace(state, snap=dataclasses.replace(state.snap, count=np.array([0, 2]))),
               dataclasses.replace(state, step=np.array([-1, 0]))]
        for s in bad:
            with pytest.raises(ValueError, match="lane"):
                ses.put(s)
        with pytest.raises(RuntimeError):
            ses.run(1)  # a rejected put leaves no state to run
        ses.put(state)
        n_active = ses.run(4)
        back = ses.get(state)
        ref = simulate(e, waves, basis, st, state, 4).state
        assert np.array_equal(back.r, ref.r) and np.array_equal(back.trace.rho, ref.trace.rho)
        assert np.array_equal(back.status, ref.status) and n_active == 2
        assert np.array_equal(back.step, [4, 4]) and np.array_equal(back.trace.count, [2, 2])
    finally:
        ses.close()


def test_select_dt_matches_oracle_drift(gpu_ctx) -> None:
    e, basis, waves, vg = _system(n=24, chi=0.1)
    rng = np.random.default_rng(4)
    r0 = rng.uniform(size=(2, 24, 2)) @ basis
    v0 = 0.05 * rng.normal(size=(2, 24, 2))
    keys = np.stack([prng_key(0), prng_key(1)])
    state = init_state(e, waves, basis, r0, v0, 0.0025, keys, MDSettings())
    ladder, n_steps = (0.4, 0.1, 0.02), 50

    def rel_drift(dt: float) -> float:
        worst = 0.0
        for i in range(2):
            ref = O.verlet_nve(vg, basis, r0[i], v0[i], dt, n_steps)
            e_last = ref["phi"] + 0.5 * float(np.sum(ref["v"] ** 2))
            worst = max(worst, abs(e_last - ref["e0"]) / (ref["sum_kinetic"] / n_steps))
        return worst

    drifts = {dt: rel_drift(dt) for dt in ladder}
    limit = math.sqrt(drifts[0.1] * drifts[0.02]) if drifts[0.1] > drifts[0.02] else 2.0 * drifts[0.02]
    want = max(dt for dt in ladder if drifts[dt] < limit)
    assert select_dt(e, waves, basis, state, ladder, n_steps, limit) == want
