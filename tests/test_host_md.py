"""Host-side MD layer against the oracle and the reference's rules (CPU only).

The single-lane host forms (integrators, observables, thermal draws), the
state containers, begin_segment, make_waves and the checkpoint files; the
device runs of the same rules are checked in test_gpu_md.py.
"""

from __future__ import annotations

import dataclasses
import json
import math

import numpy as np
import pytest

from oracle import oracle as O
from systems import small_kset
from xla_cco.dynamics import checkpoint as ck
from xla_cco.dynamics import integrators as I
from xla_cco.dynamics import observables as obs
from xla_cco.dynamics import rng as R
from xla_cco.dynamics import state as S
from xla_cco.dynamics import thermal as T
from xla_cco.kvectors import enumerate_kvectors


def _np_force(nv, w, vol, basis):
    vg = O.make_value_grad(nv, w, vol)
    inv = np.linalg.inv(basis)

    def force_fn(r):
        phi, g = vg(r @ inv)
        re, im = O.density(r @ inv, nv)
        return phi, -g @ inv.T, re, im

    return vg, force_fn


def _setup(seed=0, n=16):
    nv, w, vol, basis = small_kset(n, 2, 2.0 * math.pi / 4.0 * 2.6)
    vg, force_fn = _np_force(nv, w, vol, basis)
    rng = np.random.default_rng(seed)
    r = rng.uniform(size=(n, 2)) @ basis
    v = 0.1 * rng.standard_normal((n, 2))
    return nv, basis, vg, force_fn, r, v


def test_baoab_step_equals_oracle() -> None:
    nv, basis, vg, force_fn, r, v = _setup(1)
    f = force_fn(r)[1]
    key = R.prng_key(77)
    k2, r2, v2, out = I.baoab_step(key, r, v, f, force_fn, 0.05, 0.7, 0.3)
    kvec = np.asarray(nv, np.float64) @ (2 * math.pi * np.linalg.inv(basis).T)
    cfg = O.MDConfig(dt=0.05, scheme="baoab", gamma=0.7)
    lane = O.md_lane(vg, basis, r, v, 0.3, tuple(int(x) for x in key), kvec, cfg)
    ref = O.md_run(vg, basis, lane, kvec, kvec.shape[0], cfg, 1)
    assert tuple(int(x) for x in k2) == ref["key"]
    assert np.max(np.abs(r2 - ref["r"])) < 1e-13 and np.max(np.abs(v2 - ref["v"])) < 1e-13


@pytest.mark.parametrize("mode,p", [("count", 0.5), ("bernoulli", 0.6)])
def test_andersen_reset_equals_oracle(mode: str, p: float) -> None:
    nv, basis, vg, force_fn, r, v = _setup(2)
    kvec = np.asarray(nv, np.float64) @ (2 * math.pi * np.linalg.inv(basis).T)
    cfg = O.MDConfig(dt=0.05, scheme="andersen", andersen_p=p, andersen_mode=mode)
    key = R.prng_key(5)
    lane = O.md_lane(vg, basis, r, v, 0.3, tuple(int(x) for x in key), kvec, cfg)
    f = lane["f"]
    acc, fired = 0.0, 0
    for _ in range(6):
        r, vh, out = I.verlet_step(r, v, f, force_fn, 0.05)
        key, v, acc, n_reset = I.andersen_reset(key, vh, acc, 0.3, p, mode)
        f = out[1]
        fired += n_reset
        lane = O.md_run(vg, basis, lane, kvec, kvec.shape[0], cfg, 1)
        assert tuple(int(x) for x in key) == lane["key"]
        assert np.max(np.abs(v - lane["v"])) < 1e-13
        assert acc == pytest.approx(lane["reset_acc"], abs=1e-15)
    assert fired >= 1


def test_andersen_checks() -> None:
    v = np.zeros((3, 2))
    with pytest.raises(ValueError):
        I.andersen_reset(R.prng_key(1), v, 0.0, 1.0, 1.5, "count")
    with pytest.raises(ValueError):
        I.andersen_reset(R.prng_key(1), v, 0.0, 1.0, 0.5, "sometimes")
    with pytest.raises(TypeError):
        I.andersen_reset(np.array([1, 2]), v, 0.0, 1.0, 0.5, "count")


def test_observables_against_oracle() -> None:
    rng = np.random.default_rng(4)
    r = rng.uniform(0, 5, size=(20, 3))
    kvec = rng.normal(size=(7, 3))
    assert np.max(np.abs(obs.wave_densities(r, kvec) - O.tracked_rho(r, kvec))) < 1e-12
    assert np.max(np.abs(obs.wave_densities(r, kvec, chunk_size=3) - O.tracked_rho(r, kvec))) < 1e-12
    rho = O.tracked_rho(r, kvec)
    cm = obs.center_of_mass(r)
    assert np.max(np.abs(obs.com_correct(rho, kvec, cm) - O.com_rotate(rho, kvec, cm))) < 1e-12
    # a uniform translation leaves the COM-corrected densities unchanged
    a = np.array([0.3, -1.1, 2.0])
    moved = obs.com_correct(obs.wave_densities(r + a, kvec), kvec, obs.center_of_mass(r + a))
    assert np.max(np.abs(moved - obs.com_correct(rho, kvec, cm))) < 1e-11


def test_winding_update_closed_form() -> None:
    rho_old = np.array([[2.0, 0.0], [0.0, 1.0], [1e-9, 0.0]])
    ang = np.array([0.3, -0.5, 0.2])
    rot = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    rho_new = np.stack([rho_old[:, 0] * rot[:, 0] - rho_old[:, 1] * rot[:, 1],
                        rho_old[:, 0] * rot[:, 1] + rho_old[:, 1] * rot[:, 0]], axis=1)
    active = np.array([True, True, True])
    rp, wa, wn, nv = obs.winding_update(rho_old, rho_new, np.zeros(3), np.zeros(3), np.zeros(3, np.int64),
                                        active, 4, 1e-12)
    assert np.allclose(wn[:2], ang[:2], atol=1e-14) and np.allclose(wa[:2], np.abs(ang[:2]), atol=1e-14)
    assert wn[2] == 0.0 and list(nv) == [1, 1, 0]  # |rho|^2 / N below the floor
    assert np.array_equal(rp, rho_new)
    rp, _, wn, nv = obs.winding_update(rho_old, rho_new, np.zeros(3), np.zeros(3), np.zeros(3, np.int64),
                                       np.array([True, False, False]), 4, 1e-12)
    assert np.array_equal(rp[1:], rho_old[1:]) and list(nv) == [1, 0, 0]


def test_maxwell_boltzmann_draws() -> None:
    v = T.maxwell_boltzmann(R.prng_key(9), 4000, 3, 0.7, remove_com=False)
    # NumPy's vectorised cos / sin / log1p may differ from libm's by an ulp
    assert np.max(np.abs(v.reshape(-1) - math.sqrt(0.7) * O.rng_normals((0, 9), 12000))) < 1e-14
    assert abs(v.var() - 0.7) < 0.03
    w = T.maxwell_boltzmann(R.prng_key(9), 50, 2, 0.7, remove_com=True)
    assert np.max(np.abs(w.sum(axis=0))) < 1e-12
    assert T.kinetic_temperature(w, 2 * 49) == pytest.approx(2 * T.kinetic_energy(w) / 98)
    with pytest.raises(ValueError):
        T.maxwell_boltzmann(R.prng_key(1), 0, 2, 1.0, True)
    with pytest.raises(TypeError):
        T.maxwell_boltzmann(np.array([1, 2], dtype=np.int64), 3, 2, 1.0, True)
    with pytest.raises(ValueError):
        T.kinetic_temperature(w, 0)


def test_make_waves_order_and_cutoffs() -> None:
    from xla_cco.dynamics.run import make_waves

    basis = np.diag([4.0, 5.0])
    wv = make_waves(basis, 3.0, 4.5)
    inner = enumerate_kvectors(basis, 0.0, 3.0)
    outer = enumerate_kvectors(basis, 3.0, 4.5)
    assert wv.n_constrained == inner.n.shape[0]
    assert np.array_equal(wv.n_vectors[: wv.n_constrained], inner.n)
    assert np.array_equal(wv.n_vectors[wv.n_constrained:], outer.n)
    kn = np.linalg.norm(wv.kvec, axis=1)
    assert np.all(kn[: wv.n_constrained] < 3.0) and np.all((kn[wv.n_constrained:] > 3.0) & (kn[wv.n_constrained:] < 4.5))
    assert np.allclose(wv.kvec, wv.n_vectors @ (2 * math.pi * np.linalg.inv(basis).T))
    assert list(wv.constrained) == [True] * wv.n_constrained + [False] * (len(kn) - wv.n_constrained)
    with pytest.raises(ValueError):
        make_waves(basis, 3.0, 2.0)
    with pytest.raises(ValueError):
        make_waves(basis, 0.1, 0.2)


def _hand_state(b=3, n=5, d=2, w=4, nr=2, q=3, seed=0) -> S.MDState:
    rng = np.random.default_rng(seed)
    f = lambda *shape: rng.standard_normal((b,) + shape)  # noqa: E731
    return S.MDState(
        r=f(n, d), v=f(n, d), f=f(n, d), phi=f(), temperature=np.abs(f()), key=rng.integers(0, 2**32, (b, 2)).astype(np.uint32),
        step=np.arange(b, dtype=np.int64), time=f(), reset_acc=np.abs(f()), rho_prev=f(w, 2), wind_abs=f(w),
        wind_net=f(w), n_valid=np.ones((b, w), np.int64), e0=f(), max_drift=np.abs(f()), sum_kinetic=f(),
        sum_phi=f(), sum_phi2=f(), sum_s=f(w), n_sum=np.full(b, 3, np.int64),
        status=np.array([0, 7, 0][:b], np.int64),
        snap=S.Snapshots(step=np.full((b, nr), -1, np.int64), time=f(nr), wind_abs=f(nr, w), wind_net=f(nr, w),
                         rho=f(nr, w, 2), r=f(nr, n, d), phi=f(nr), kinetic=f(nr), count=np.zeros(b, np.int64)),
        trace=S.RhoTrace(rho=f(q, w, 2), time=f(q), count=np.zeros(b, np.int64)),
    )


def test_begin_segment_rules() -> None:
    from xla_cco.dynamics.run import begin_segment

    st = _hand_state()
    out = begin_segment(st, S.MDSettings(), reset_sums=True, reset_winding=True)
    for i in (0, 2):  # running lanes
        assert np.max(np.abs(out.v[i].sum(axis=0))) < 1e-12
        assert out.e0[i] == pytest.approx(out.phi[i] + 0.5 * np.sum(out.v[i] ** 2))
        assert out.max_drift[i] == 0.0 and out.n_sum[i] == 0 and np.all(out.sum_s[i] == 0.0)
        assert np.all(np.isnan(out.rho_prev[i])) and np.all(out.n_valid[i] == 0)
        assert np.array_equal(out.r[i], st.r[i]) and np.array_equal(out.key[i], st.key[i])
    # a stopped lane is returned unchanged
    for a, b in zip(S.leaves(st), S.leaves(out)):
        assert np.array_equal(np.asarray(a)[1], np.asarray(b)[1], equal_nan=True)
        assert np.asarray(a).dtype == np.asarray(b).dtype
    keep = begin_segment(st, S.MDSettings(remove_com=False), reset_sums=False, reset_winding=False)
    assert np.array_equal(keep.v, st.v) and np.array_equal(keep.sum_s, st.sum_s)
    one = begin_segment(S.lane(st, 0), S.MDSettings(), True, False)
    assert np.allclose(one.v, out.v[0]) and np.array_equal(one.rho_prev, st.rho_prev[0])


def test_stack_lane_round_trip() -> None:
    st = _hand_state()
    again = S.stack([S.lane(st, i) for i in range(3)])
    for a, b in zip(S.leaves(st), S.leaves(again)):
        assert np.array_equal(a, b)


def test_settings_checks() -> None:
    S.check_settings(S.MDSettings())
    bad = [dict(scheme="nose"), dict(dt=0.0), dict(andersen_p=1.5), dict(chunk_steps=0), dict(record_steps=(3, 3)),
           dict(record_steps=(0,)), dict(record_rho_every=2, max_rho_records=0), dict(drift_tol=0.0),
           dict(record_steps=tuple(range(1, 258))), dict(precision="fp16")]
    for kw in bad:
        with pytest.raises(ValueError):
            S.check_settings(dataclasses.replace(S.MDSettings(), **kw))


def test_checkpoint_round_trip(tmp_path) -> None:
    st = _hand_state(seed=3)
    p = tmp_path / "ck.npz"
    ck.save_state(p, st, {"steps_done": 5})
    assert not ck.tmp_path_of(p).exists()
    back, meta = ck.load_state(p, _hand_state(seed=9))
    assert meta == {"steps_done": 5} and ck.read_meta(p) == meta
    for a, b in zip(S.leaves(st), S.leaves(back)):
        assert np.array_equal(a, b) and np.asarray(a).dtype == np.asarray(b).dtype
    with np.load(p) as z:  # the documented layout
        assert int(z["__n_leaves__"]) == len(list(S.leaves(st)))
        assert json.loads(str(z["__meta__"][()])) == meta
        assert np.array_equal(z["leaf_000"], st.r) and np.array_equal(z["leaf_005"], st.key)
    with pytest.raises(ck.CheckpointError):
        ck.load_state(p, _hand_state(n=6))  # shape mismatch
    (tmp_path / "bad.npz").write_bytes(b"not a zip at all")
    with pytest.raises(ValueError):
        ck.load_state(tmp_path / "bad.npz", st)
    with pytest.raises(ValueError):
        ck.read_meta(tmp_path / "bad.npz")
    with pytest.raises(TypeError):
        ck.save_state(p, st, {"x": object()})
    with pytest.raises(TypeError):
        ck.save_state(p, "state", {})
    back, _ = ck.load_state(p, st)  # a failed save kept the previous file
    assert np.array_equal(back.r, st.r)


def test_lattice_rows_recover_n() -> None:
    """The tracked-wave density pass takes waves whose n = B k / 2 pi is integral."""
    from xla_cco.dynamics.run import _lattice_rows, make_waves

    for d, basis in ((2, np.array([[3.0, 0.0], [1.2, 2.5]])), (3, 4.0 * np.eye(3))):
        waves = make_waves(basis, 2.5, 5.0)
        n = _lattice_rows(basis, waves.kvec)
        assert n is not None and np.array_equal(n, waves.n_vectors)
        off = waves.kvec.copy()
        off[-1] *= 1.0 + 1e-7
        assert _lattice_rows(basis, off) is None
        zero = waves.kvec.copy()
        zero[0] = 0.0
        assert _lattice_rows(basis, zero) is None
