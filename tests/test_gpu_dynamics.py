"""Device gradient flow (DOPRI5) and NVE Verlet vs the oracle.

Mirrors tests/unit/test_flow.py and tests/unit/test_md_integrators.py of the
reference: closed-form linear flow, monotone energy, ground state, zero
gradient, recording, batch == single; Verlet against the NumPy statement,
time reversal on the device, drift guard.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from xla_cco import kvectors as K
from xla_cco.box import cubic_box
from xla_cco.dynamics import MDSettings, Waves, init_state, make_force_fn, prng_key, simulate, verlet_step
from xla_cco.energy import AbstractEnergy, make_stealthy_energy
from xla_cco.minimize.flow import FlowSettings, gradient_flow, gradient_flow_batch
from xla_cco.minimize.result import STATUS_BUDGET, STATUS_DRIFT, STATUS_ENERGY_TOL, STATUS_GRAD_TOL

pytestmark = pytest.mark.gpu


def _stealthy(n: int, d: int, shells: int):
    box = cubic_box(n, d)
    side = float(box.basis[0, 0])
    ks = K.enumerate_kvectors(box.basis, 0.0, 2.0 * np.pi / side * (shells + 0.5))
    return make_stealthy_energy(box, ks), box.basis, O.make_value_grad(ks.n, ks.weights, box.volume)


class _Quadratic(AbstractEnergy):
    def __init__(self, a: np.ndarray, c: float) -> None:
        self.a, self.c = a, c

    def __call__(self, s):
        return float(0.5 * self.c * np.sum((s - self.a) ** 2))

    def value_and_grad(self, s):
        d = s - self.a
        return float(0.5 * self.c * np.sum(d * d)), self.c * d


@pytest.mark.parametrize("side,c,t_end", [(0.5, 0.2, 0.1), (1.7, 3.0, 1.3)])
def test_flow_linear_closed_form(gpu_ctx, side: float, c: float, t_end: float) -> None:
    a = np.array([[0.3, 0.6]])
    s0 = np.array([[0.9, 0.1]])
    res = gradient_flow(_Quadratic(a, c), s0, np.eye(2) * side,
                        FlowSettings(rtol=1e-10, atol=1e-14, t_max=t_end, max_steps=100_000))
    assert int(res.status) == STATUS_BUDGET
    assert float(res.t) == pytest.approx(t_end, rel=1e-12)
    expected = a + (s0 - a) * np.exp(-c * float(res.t) / side**2)
    assert np.max(np.abs(res.s - expected)) < 1e-9


def test_flow_matches_oracle_and_monotone(gpu_ctx) -> None:
    e, basis, vg = _stealthy(16, 1, 5)
    s0 = np.random.default_rng(1).uniform(size=(16, 1))
    fs = FlowSettings(energy_tol=1e-20, record_every=1, max_records=50_000)
    res = gradient_flow(e, s0, basis, fs)
    ref = O.gradient_flow(vg, s0, basis, O.FlowSettings(energy_tol=1e-20, record_every=1, max_records=50_000))
    n = int(res.trajectory.count)
    ee = res.trajectory.energy[:n]
    assert n > 10 and np.all(np.diff(res.trajectory.t[:n]) > 0.0)
    assert np.all(np.diff(ee) <= 1e-12 * ee[:-1] + 1e-30)
    assert int(res.status) == STATUS_ENERGY_TOL and float(res.energy) < 1e-20
    # the first 20 accepted steps follow the oracle
    k = min(20, n, len(ref["traj"]))
    assert np.allclose(res.trajectory.t[:k], [r[1] for r in ref["traj"][:k]], rtol=1e-9)
    assert np.allclose(res.trajectory.energy[:k], [r[3] for r in ref["traj"][:k]], rtol=1e-8)
    assert abs(int(res.n_steps) - ref["n_steps"]) <= max(3, ref["n_steps"] // 20)


def test_flow_ground_state_zero_gradient_and_recording(gpu_ctx) -> None:
    e, basis, _ = _stealthy(32, 1, 4)
    s0 = np.random.default_rng(2).uniform(size=(32, 1))
    res = gradient_flow(e, s0, basis, FlowSettings(energy_tol=1e-20))
    assert int(res.status) == STATUS_ENERGY_TOL and float(res.energy) < 1e-20
    e1, b1, _ = _stealthy(1, 1, 3)
    z = gradient_flow(e1, np.zeros((1, 1)), b1, FlowSettings())
    assert int(z.status) == STATUS_GRAD_TOL and int(z.n_steps) == 0
    assert np.array_equal(z.s, np.zeros((1, 1)))
    e3, b3, _ = _stealthy(16, 1, 3)
    s3 = np.random.default_rng(3).uniform(size=(16, 1))
    r3 = gradient_flow(e3, s3, b3, FlowSettings(energy_tol=1e-20, record_every=3, max_records=4000))
    tr = r3.trajectory
    n = int(tr.count)
    assert 2 <= n <= 4000
    assert np.array_equal(tr.s[0], s3) and np.array_equal(tr.s[n - 1], r3.s)
    assert float(tr.t[n - 1]) == float(r3.t) and float(tr.t[0]) == 0.0


def test_flow_batch_equals_single_bitwise(gpu_ctx) -> None:
    e, basis, _ = _stealthy(8, 1, 2)
    s0 = np.random.default_rng(4).uniform(size=(3, 8, 1))
    fs = FlowSettings(energy_tol=1e-20)
    batched = gradient_flow_batch(e, s0, basis, fs)
    for i in range(3):
        one = gradient_flow(e, s0[i], basis, fs)
        assert np.array_equal(batched.s[i], one.s)
        assert int(batched.n_steps[i]) == int(one.n_steps)


def test_flow_d3_stealthy_reaches_ground_state(gpu_ctx) -> None:
    box = cubic_box(100, 3)
    ks = K.enumerate_kvectors(box.basis, 0.0, K.k_outer_from_ka(K.ka_from_chi(0.05, 0.05, 3), 0.05, 3))
    e = make_stealthy_energy(box, ks)
    s0 = np.random.default_rng(0).uniform(size=(4, 100, 3))
    res = gradient_flow_batch(e, s0, box.basis, FlowSettings(energy_tol=1e-20))
    assert np.all(res.status == STATUS_ENERGY_TOL) and np.all(res.energy < 1e-20)


def test_flow_rejects_bad_input(gpu_ctx) -> None:
    e, basis, _ = _stealthy(4, 1, 1)
    with pytest.raises(TypeError):
        gradient_flow(e, np.zeros((4, 1), np.float32), basis, FlowSettings())
    with pytest.raises(ValueError):
        gradient_flow(e, np.zeros((4, 1)), basis, FlowSettings(rtol=0.0))
    with pytest.raises(NotImplementedError):  # ROS2 needs the GPU Hessian of a StealthyEnergy
        gradient_flow(_Quadratic(np.zeros((1, 2)), 1.0), np.ones((1, 2)), np.eye(2),
                      FlowSettings(method="rosenbrock"))
    with pytest.raises(ValueError, match="rosenbrock"):
        gradient_flow(e, np.zeros((4, 1)), basis, FlowSettings(method="rosenbrock", parametrization="arclength"))


# ---------------- ROS2 (flow.py:367-398) ----------------


@pytest.mark.parametrize("n,d,shells,t_frac", [(16, 1, 4, 0.5), (24, 2, 2, 1.0)])
def test_rosenbrock_matches_oracle(gpu_ctx, n: int, d: int, shells: int, t_frac: float) -> None:
    e, basis, vg = _stealthy(n, d, shells)
    ks_n, ks_w = e.n_vectors, e.weights
    vol = float(e.volume)
    s0 = np.random.default_rng(60 + n).uniform(size=(n, d))
    full = O.gradient_flow(vg, s0, basis, O.FlowSettings(energy_tol=1e-12))
    st = dict(method="rosenbrock", rtol=1e-3, atol=1e-10, t_max=t_frac * float(full["t"]), max_steps=5000)
    ref = O.gradient_flow(vg, s0, basis, O.FlowSettings(**st), hess=lambda s: O.hessian(s, ks_n, ks_w, vol))
    res = gradient_flow(e, s0, basis, FlowSettings(**st))
    assert int(res.status) == ref["status"]
    assert int(res.n_steps) == ref["n_steps"] and int(res.n_evals) == ref["n_evals"]
    assert np.max(np.abs(res.s - ref["s"])) < 1e-9
    assert float(res.t) == pytest.approx(ref["t"], rel=1e-12)
    assert float(res.length) == pytest.approx(ref["length"], rel=1e-9)
    # the same flow as a tight DOPRI5 at the same time
    dop = gradient_flow(e, s0, basis, FlowSettings(rtol=1e-10, t_max=float(res.t)))
    assert np.max(np.abs(res.s - dop.s)) < 1e-5


def test_rosenbrock_batch_equals_single_and_records(gpu_ctx) -> None:
    e, basis, _ = _stealthy(20, 2, 2)
    s0 = np.random.default_rng(61).uniform(size=(3, 20, 2))
    fs = FlowSettings(method="rosenbrock", rtol=0.1, atol=1e-4, energy_tol=1e-6, max_steps=3000,
                      record_every=5, max_records=40)
    batched = gradient_flow_batch(e, s0, basis, fs)
    assert np.all(batched.status == STATUS_ENERGY_TOL) and np.all(batched.energy < 1e-6)
    for i in range(3):
        one = gradient_flow(e, s0[i], basis, fs)
        assert np.array_equal(batched.s[i], one.s)
        assert int(batched.n_steps[i]) == int(one.n_steps)
        k = int(one.trajectory.count)
        assert k >= 2 and np.all(np.diff(one.trajectory.t[:k]) > 0)
        assert np.all(np.diff(one.trajectory.energy[:k]) <= 1e-12)


# ---------------- velocity Verlet ----------------


def _md_system(n: int = 64, d: int = 2, chi: float = 0.2):
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d))
    e = make_stealthy_energy(box, ks)
    return e, box.basis, O.make_value_grad(ks.n, ks.weights, box.volume)


def _energy_waves(e, basis) -> Waves:
    """The energy's own k-set (its padding rows dropped) as the tracked waves, all constrained."""
    nv = np.asarray(e.n_vectors, dtype=np.float64)
    nv = nv[: int(np.count_nonzero(np.any(nv != 0.0, axis=1)))]
    return Waves(n_vectors=nv, kvec=nv @ (2.0 * np.pi * np.linalg.inv(basis).T),
                 constrained=np.ones(nv.shape[0], bool), n_constrained=nv.shape[0])


def _start(e, basis, r0, v0, temperature, st):
    keys = np.stack([prng_key(i) for i in range(r0.shape[0])])
    return init_state(e, _energy_waves(e, basis), basis, r0, v0, temperature, keys, st)


def _run(e, basis, st, state, n):
    return simulate(e, _energy_waves(e, basis), basis, st, state, n).state


def test_md_matches_oracle(gpu_ctx) -> None:
    e, basis, vg = _md_system()
    rng = np.random.default_rng(0)
    r0 = rng.uniform(size=(3, 64, 2)) @ basis
    v0 = 0.01 * rng.normal(size=(3, 64, 2))
    st = MDSettings(dt=0.01, remove_com=False, drift_tol=math.inf)
    state = _start(e, basis, r0, v0, 1e-4, st)
    out = _run(e, basis, st, state, 60)
    for i in range(3):
        ref = O.verlet_nve(vg, basis, r0[i], v0[i], 0.01, 60)
        assert int(out.step[i]) == 60 and int(out.status[i]) == 0
        assert np.max(np.abs(out.r[i] - ref["r"])) < 1e-11
        assert np.max(np.abs(out.v[i] - ref["v"])) < 1e-11
        assert out.phi[i] == pytest.approx(ref["phi"], rel=1e-9)
        assert out.e0[i] == pytest.approx(ref["e0"], rel=1e-12)
        assert out.sum_phi[i] == pytest.approx(ref["sum_phi"], rel=1e-9)
        assert out.sum_kinetic[i] == pytest.approx(ref["sum_kinetic"], rel=1e-9)


def test_md_time_reversal_on_device(gpu_ctx) -> None:
    e, basis, _ = _md_system(40, 3, 0.1)
    rng = np.random.default_rng(1)
    r0 = rng.uniform(size=(2, 40, 3)) @ basis
    v0 = 0.02 * rng.normal(size=(2, 40, 3))
    st = MDSettings(dt=0.01, remove_com=False, drift_tol=math.inf)
    fwd = _run(e, basis, st, _start(e, basis, r0, v0, 1.0, st), 150)
    back = _run(e, basis, st, _start(e, basis, fwd.r, -fwd.v, 1.0, st), 150)
    assert np.max(np.abs(back.r - r0)) < 1e-10


def test_md_drift_guard_freezes_lane(gpu_ctx) -> None:
    e, basis, _ = _md_system()
    rng = np.random.default_rng(2)
    r0 = rng.uniform(size=(2, 64, 2)) @ basis
    v0 = 0.01 * rng.normal(size=(2, 64, 2))
    # huge dt on random (high-energy) states: |E - e0| explodes at once
    st = MDSettings(dt=0.5, remove_com=False, drift_tol=1e-3)
    out = _run(e, basis, st, _start(e, basis, r0, v0, 1e-6, st), 40)
    assert np.all(out.status == STATUS_DRIFT)
    assert np.all(out.step < 40)
    frozen = out.r.copy()
    again = _run(e, basis, st, out, 5)
    assert np.array_equal(again.r, frozen)


def test_md_com_removal_and_energy_conservation(gpu_ctx) -> None:
    e, basis, vg = _md_system(64, 2, 0.1)
    rng = np.random.default_rng(3)
    r0 = rng.uniform(size=(1, 64, 2)) @ basis
    v0 = 1e-3 * rng.normal(size=(1, 64, 2)) + 0.5
    st = MDSettings(dt=0.005, remove_com=True, drift_tol=math.inf)
    from xla_cco.dynamics import begin_segment

    state = begin_segment(_start(e, basis, r0, v0, 1e-6, st), st, reset_sums=True, reset_winding=False)
    out = _run(e, basis, st, state, 200)
    assert np.abs(out.v[0].sum(axis=0)).max() < 1e-12
    assert out.max_drift[0] < 5e-3 * abs(out.e0[0])


def test_verlet_step_with_gpu_force_fn(gpu_ctx) -> None:
    e, basis, vg = _md_system()
    fn = make_force_fn(e, basis)
    rng = np.random.default_rng(4)
    r = rng.uniform(size=(64, 2)) @ basis
    v = 0.01 * rng.normal(size=(64, 2))
    phi, f, _, _ = fn(r)
    r1, v1, out = verlet_step(r, v, f, fn, 0.01)
    ref = O.verlet_nve(vg, basis, r, v, 0.01, 1)
    assert np.max(np.abs(r1 - ref["r"])) < 1e-13 and np.max(np.abs(v1 - ref["v"])) < 1e-12


def test_harmonic_sample_matches_oracle_hessian(gpu_ctx) -> None:
    """thermal.py:118-153 with the GPU Hessian equals the same draw from the oracle's Hessian."""
    from oracle import oracle as O
    from xla_cco import kvectors as K
    from xla_cco.box import cubic_box
    from xla_cco.dynamics.thermal import harmonic_sample
    from xla_cco.energy import AbstractEnergy, make_stealthy_energy

    n, d = 24, 2
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, 4.2)
    e = make_stealthy_energy(box, ks)

    class OracleHessian(AbstractEnergy):
        def __call__(self, s):  # type: ignore[no-untyped-def]
            return O.energy_grad(s, ks.n, ks.weights, box.volume)[0]

        def hessian(self, s):  # type: ignore[no-untyped-def]
            return O.hessian(s, ks.n, ks.weights, box.volume)

    s = np.random.default_rng(50).uniform(size=(n, d))
    a = harmonic_sample(np.random.default_rng(7), e, s, box.basis, 1e-3)
    b = harmonic_sample(np.random.default_rng(7), OracleHessian(), s, box.basis, 1e-3)
    assert np.allclose(a, b, rtol=0, atol=1e-9)
    assert not np.allclose(a, s @ box.basis, atol=1e-6)
