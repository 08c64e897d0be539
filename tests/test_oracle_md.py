"""Pins of the oracle's MD step with thermostats and tracked waves (CPU only).

Closed forms and special cases that a dropped term, a wrong sign or index,
or an off-by-one in the cadence fails: Verlet against ``verlet_nve``; BAOAB
at gamma = 0 and Andersen at p = 0 reduce to Verlet; the count mode fires
exactly every 1/p steps; ballistic motion winds each wave by wrap(k . u dt)
per step and not at all once COM-corrected; BAOAB's exact harmonic
configurational variance; snapshots / trace rows against shorter runs; a
drift stop freezes the lane.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from systems import small_kset


def _wrap(x: np.ndarray) -> np.ndarray:
    return (x + math.pi) % (2.0 * math.pi) - math.pi


def _free(s):
    return 0.0, np.zeros_like(s)


def _stealthy(n: int = 16, d: int = 2, seed: int = 0):
    side = math.sqrt(n) if d == 2 else float(n)
    nv, w, vol, basis = small_kset(n, d, 2.0 * math.pi / side * 2.6)
    vg = O.make_value_grad(nv, w, vol)
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(size=(n, d)) @ basis
    v0 = 0.05 * rng.standard_normal((n, d))
    kvec = np.asarray(nv, dtype=np.float64) @ (2.0 * math.pi * np.linalg.inv(basis).T)
    return vg, basis, r0, v0, kvec, nv.shape[0]


def test_verlet_equals_verlet_nve() -> None:
    vg, basis, r0, v0, kvec, nc = _stealthy()
    cfg = O.MDConfig(dt=0.02)
    lane = O.md_lane(vg, basis, r0, v0, 0.01, (1, 2), kvec, cfg)
    out = O.md_run(vg, basis, lane, kvec, nc, cfg, 25)
    ref = O.verlet_nve(vg, basis, r0, v0, 0.02, 25, temperature=0.01)
    assert np.array_equal(out["r"], ref["r"]) and np.array_equal(out["v"], ref["v"])
    assert out["sum_kinetic"] == ref["sum_kinetic"] and out["sum_phi2"] == ref["sum_phi2"]
    assert out["step"] == 25 and out["time"] == pytest.approx(0.5, rel=1e-14)
    assert out["key"] == (1, 2)  # verlet draws nothing


@pytest.mark.parametrize("scheme", ["baoab", "andersen"])
def test_thermostats_reduce_to_verlet(scheme: str) -> None:
    vg, basis, r0, v0, kvec, nc = _stealthy(seed=1)
    cfg_v = O.MDConfig(dt=0.02)
    cfg_t = O.MDConfig(dt=0.02, scheme=scheme, gamma=0.0, andersen_p=0.0)
    a = O.md_run(vg, basis, O.md_lane(vg, basis, r0, v0, 0.01, (3, 4), kvec, cfg_v), kvec, nc, cfg_v, 20)
    b = O.md_run(vg, basis, O.md_lane(vg, basis, r0, v0, 0.01, (3, 4), kvec, cfg_t), kvec, nc, cfg_t, 20)
    assert np.max(np.abs(a["r"] - b["r"])) < 1e-12
    assert np.max(np.abs(a["v"] - b["v"])) < 1e-12
    assert b["key"] != (3, 4)  # the key still advances


def test_andersen_count_fires_every_one_over_p() -> None:
    vg, basis, r0, v0, kvec, nc = _stealthy(seed=2)
    cfg = O.MDConfig(dt=0.02, scheme="andersen", andersen_p=0.25)
    ref = O.MDConfig(dt=0.02, scheme="andersen", andersen_p=0.0)
    lane = O.md_lane(vg, basis, r0, v0, 0.5, (5, 6), kvec, cfg)
    lane_ref = O.md_lane(vg, basis, r0, v0, 0.5, (5, 6), kvec, ref)
    fired = []
    for step in range(1, 13):
        nxt = O.md_run(vg, basis, lane, kvec, nc, cfg, 1)
        # the same step without a reset: one row differs exactly when the count fires
        plain = O.md_run(vg, basis, dict(lane_ref, **{k: lane[k] for k in ("r", "v", "f", "phi", "key")}),
                         kvec, nc, ref, 1)
        rows = np.nonzero(np.any(nxt["v"] != plain["v"], axis=1))[0]
        assert len(rows) in (0, 1)
        if len(rows):
            fired.append(step)
            assert rows[0] == O.rng_randint(O.rng_split(lane["key"], 4)[1], r0.shape[0])
        lane = nxt
    assert fired == [4, 8, 12]
    assert lane["reset_acc"] == pytest.approx(0.0, abs=1e-12)


def test_ballistic_winding_closed_form() -> None:
    """No forces, one common velocity u: rho_k(t) = rho_k(0) exp(i k.u t)."""
    rng = np.random.default_rng(3)
    n, d, side = 12, 2, 4.0
    basis = side * np.eye(2)
    r0 = rng.uniform(0, side, size=(n, d))
    u = np.array([0.31, -0.17])
    v0 = np.tile(u, (n, 1))
    nv = np.array([[1, 0], [0, 1], [1, 1], [2, -1], [1, 2]], dtype=np.float64)
    kvec = nv @ (2.0 * math.pi * np.linalg.inv(basis).T)
    dt, steps = 0.05, 12
    for wind_every in (1, 3):
        cfg = O.MDConfig(dt=dt, wind_every=wind_every, phase_floor=0.0)
        lane = O.md_lane(_free, basis, r0, v0, 1.0, (0, 0), kvec, cfg)
        out = O.md_run(_free, basis, lane, kvec, 2, cfg, steps)
        per = kvec @ u * dt
        want = np.where(np.arange(5) < 2, steps * _wrap(per), (steps // wind_every) * _wrap(wind_every * per))
        assert np.max(np.abs(out["wind_net"] - want)) < 1e-11
        assert np.max(np.abs(out["wind_abs"] - np.abs(want))) < 1e-11
        assert list(out["n_valid"]) == [steps, steps] + [steps // wind_every] * 3
        s0 = (lane["rho_prev"] ** 2).sum(axis=1) / n
        assert np.max(np.abs(out["sum_s"] - steps * s0)) < 1e-10 * steps * max(1.0, s0.max())
    # COM-corrected (BAOAB without friction): translation removed, no winding
    cfg = O.MDConfig(dt=dt, scheme="baoab", gamma=0.0, phase_floor=0.0)
    out = O.md_run(_free, basis, O.md_lane(_free, basis, r0, v0, 1.0, (0, 0), kvec, cfg), kvec, 2, cfg, steps)
    assert np.max(np.abs(out["wind_net"])) < 1e-11 and np.max(out["wind_abs"]) < 1e-10


def test_phase_floor_excludes_silent_waves() -> None:
    """A lattice wave with rho_k = 0 never counts toward the winding."""
    n, side = 4, 2.0
    basis = side * np.eye(1)
    r0 = np.array([[0.0], [0.5], [1.0], [1.5]])  # rho = 0 at k = 2 pi / side
    kvec = np.array([[2.0 * math.pi / side], [4.0 * math.pi / side * 2.0]])  # second: |rho| = 4
    cfg = O.MDConfig(dt=0.1, phase_floor=1e-12)
    lane = O.md_lane(_free, basis, r0, np.full((n, 1), 0.2), 1.0, (0, 0), kvec, cfg)
    out = O.md_run(_free, basis, lane, kvec, 2, cfg, 5)
    assert out["n_valid"][0] == 0 and out["wind_abs"][0] == 0.0
    assert out["n_valid"][1] == 5


def test_baoab_harmonic_configurational_variance() -> None:
    """BAOAB samples the exact <x^2> = T / kappa of a harmonic well at any stable step."""
    kappa, T, dt = 1.0, 0.4, 1.2  # omega dt = 1.2: Verlet's <x^2> is T / (kappa (1 - 0.36)), 56 % high

    def well(s):
        return float(0.5 * kappa * np.sum(s * s)), kappa * s

    cfg = O.MDConfig(dt=dt, scheme="baoab", gamma=1.0)
    basis = np.eye(1)
    kvec = np.array([[1.0]])
    lane = O.md_lane(well, basis, np.zeros((2, 1)), np.zeros((2, 1)), T, (7, 9), kvec, cfg)
    xs = []
    for _ in range(500):
        lane = O.md_run(well, basis, lane, kvec, 1, cfg, 20)
        xs.append(lane["r"][:, 0].copy())
    x = np.concatenate(xs)
    # samples 24 time units (24 / gamma) apart: relative error ~ sqrt(2 / 1000) = 4.5 %
    assert np.mean(x * x) == pytest.approx(T / kappa, rel=0.15)


def test_records_equal_shorter_runs() -> None:
    vg, basis, r0, v0, kvec, nc = _stealthy(seed=4)
    extra = np.array([[3.0, 1.0], [2.0, 2.0]]) @ (2.0 * math.pi * np.linalg.inv(basis).T)
    kv = np.concatenate([kvec, extra])
    cfg = O.MDConfig(dt=0.02, scheme="andersen", andersen_p=0.3, andersen_mode="bernoulli", wind_every=2,
                     record_steps=(3, 7), record_rho_every=2, max_rho_records=4)
    lane = O.md_lane(vg, basis, r0, v0, 0.05, (11, 12), kv, cfg)
    out = O.md_run(vg, basis, lane, kv, nc, cfg, 10)
    assert list(out["snap_step"]) == [3, 7] and out["snap_count"] == 2
    assert out["trace_count"] == 4
    assert np.allclose(out["trace_time"], [0.04, 0.08, 0.12, 0.16], rtol=1e-14)
    for j, s in enumerate((3, 7)):
        short = O.md_run(vg, basis, lane, kv, nc, cfg, s)
        assert np.array_equal(out["snap_r"][j], short["r"])
        assert np.array_equal(out["snap_rho"][j], short["rho_prev"])
        assert np.array_equal(out["snap_wind_net"][j], short["wind_net"])
        assert out["snap_phi"][j] == short["phi"]
    short = O.md_run(vg, basis, lane, kv, nc, cfg, 6)
    assert np.array_equal(out["trace_rho"][2], short["rho_prev"])


def test_drift_stop_freezes() -> None:
    vg, basis, r0, v0, kvec, nc = _stealthy(seed=5)
    cfg = O.MDConfig(dt=2.0, drift_tol=1e-3)
    lane = O.md_lane(vg, basis, r0, 3.0 * v0, 1e-4, (0, 0), kvec, cfg)
    out = O.md_run(vg, basis, lane, kvec, nc, cfg, 30)
    assert out["status"] == O.STATUS_DRIFT and 1 <= out["step"] < 30
    again = O.md_run(vg, basis, out, kvec, nc, cfg, 5)
    assert again["step"] == out["step"] and np.array_equal(again["r"], out["r"])
