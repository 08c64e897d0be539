"""Pins of the oracle's gradient flow and velocity Verlet (CPU only).

Flow: the closed-form linear flow of a quadratic in a cubic cell
(s(t) = a + (s0 - a) exp(-c t / L^2), test_flow.py:70-93), monotone energy,
a stealthy ground state; ROS2 (test_flow.py:243-295): the same closed form, agreement
with DOPRI5 at equal times on a stealthy energy, far fewer steps than DOPRI5 on a stiff
quadratic, where a planted sign error in its Jacobian never converges.
Verlet: the exact discrete solution of the
harmonic oscillator (cos theta = 1 - (w h)^2 / 2), second-order global error,
time reversal.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from systems import small_kset


def _quadratic_vg(a: np.ndarray, c: float):
    def vg(s):
        d = s - a
        return float(0.5 * c * np.sum(d * d)), c * d

    return vg


@pytest.mark.parametrize("side,c,t_end", [(0.5, 0.2, 0.1), (1.7, 3.0, 1.3), (3.0, 5.0, 2.0)])
def test_flow_linear_closed_form(side: float, c: float, t_end: float) -> None:
    a = np.array([[0.3, 0.6]])
    s0 = np.array([[0.9, 0.1]])
    res = O.gradient_flow(_quadratic_vg(a, c), s0, side * np.eye(2),
                          O.FlowSettings(rtol=1e-10, atol=1e-14, t_max=t_end, max_steps=100_000))
    assert res["status"] == O.STATUS_BUDGET
    assert res["t"] == pytest.approx(t_end, rel=1e-12)
    expected = a + (s0 - a) * np.exp(-c * res["t"] / side**2)
    assert np.max(np.abs(res["s"] - expected)) < 1e-9


def test_flow_monotone_and_ground_state() -> None:
    side = 16.0
    nv, w, vol, basis = small_kset(16, 1, 2.0 * np.pi / side * 5.5)
    vg = O.make_value_grad(nv, w, vol)
    s0 = np.random.default_rng(1).uniform(size=(16, 1))
    res = O.gradient_flow(vg, s0, basis, O.FlowSettings(energy_tol=1e-20, record_every=1, max_records=50_000))
    e = np.array([r[3] for r in res["traj"]])
    t = np.array([r[1] for r in res["traj"]])
    assert len(e) > 10 and np.all(np.diff(t) > 0.0)
    assert np.all(np.diff(e) <= 1e-12 * e[:-1] + 1e-30)
    assert res["status"] == O.STATUS_ENERGY_TOL and res["energy"] < 1e-20


def test_flow_arclength_same_path() -> None:
    """The arclength parametrisation follows the same curve: equal end point at the same length."""
    a = np.array([[0.3, 0.6]])
    s0 = np.array([[0.9, 0.1]])
    vg = _quadratic_vg(a, 1.0)
    t_run = O.gradient_flow(vg, s0, np.eye(2), O.FlowSettings(rtol=1e-10, t_max=1.0))
    L = t_run["length"]
    exact_len = np.linalg.norm(s0 - a) * (1.0 - math.exp(-1.0))
    assert L == pytest.approx(exact_len, rel=1e-8)


def test_verlet_harmonic_exact_discrete_solution() -> None:
    """F = -w^2 r in a unit cell: the Verlet map is x_n = A cos(n th) + B sin(n th)."""
    w2 = 2.5
    h = 0.05

    def vg(s):  # Phi = w^2/2 |s|^2 in a unit cell: F = -w^2 r
        return float(0.5 * w2 * np.sum(s * s)), w2 * s

    r0 = np.array([[0.7, -0.2]])
    v0 = np.array([[0.1, 0.3]])
    n = 400
    out = O.verlet_nve(vg, np.eye(2), r0, v0, h, n)
    th = math.acos(1.0 - w2 * h * h / 2.0)
    x1 = r0 + h * v0 - 0.5 * h * h * w2 * r0
    A = r0
    B = (x1 - r0 * math.cos(th)) / math.sin(th)
    expect = A * math.cos(n * th) + B * math.sin(n * th)
    assert np.max(np.abs(out["r"] - expect)) < 1e-12


def test_verlet_second_order_and_time_reversal() -> None:
    nv, w, vol, basis = small_kset(8, 2, 4.0)
    vg = O.make_value_grad(nv, w, vol)
    rng = np.random.default_rng(0)
    r0 = rng.uniform(size=(8, 2)) @ basis
    v0 = 0.05 * rng.normal(size=(8, 2))
    T = 0.4
    errs = []
    ref = O.verlet_nve(vg, basis, r0, v0, 0.0025, int(T / 0.0025))["r"]
    for h in (0.02, 0.01):
        errs.append(np.max(np.abs(O.verlet_nve(vg, basis, r0, v0, h, int(round(T / h)))["r"] - ref)))
    assert 3.0 < errs[0] / errs[1] < 5.0
    fwd = O.verlet_nve(vg, basis, r0, v0, 0.01, 100)
    back = O.verlet_nve(vg, basis, fwd["r"], -fwd["v"], 0.01, 100)
    assert np.max(np.abs(back["r"] - r0)) < 1e-11



# ---------------- ROS2 (flow.py:367-398) ----------------


def _quadratic_hess(c: np.ndarray, n: int):
    """Hessian of 1/2 sum_ij c_j (s_ij - a_ij)^2: diag(c) per particle."""
    return lambda s: np.diag(np.tile(np.asarray(c, dtype=np.float64), n))


def _quadratic_vg_axes(a: np.ndarray, c: np.ndarray):
    def vg(s):
        d = s - a
        return float(0.5 * np.sum(c * d * d)), c * d

    return vg


def test_rosenbrock_linear_flow_closed_form() -> None:
    a = np.array([[0.3, 0.6]])
    s0 = np.array([[0.9, 0.1]])
    c = np.array([1.5, 1.5])
    st = O.FlowSettings(method="rosenbrock", rtol=1e-3, atol=1e-10, t_max=1.3)
    res = O.gradient_flow(_quadratic_vg_axes(a, c), s0, np.eye(2) * 1.2, st, hess=_quadratic_hess(c, 1))
    expected = a + (s0 - a) * np.exp(-1.5 * 1.3 / 1.2**2)
    assert res["t"] == pytest.approx(1.3, rel=1e-12)
    assert np.max(np.abs(res["s"] - expected)) < 1e-5


@pytest.mark.parametrize("frac", [0.1, 0.5, 0.9])
def test_rosenbrock_agrees_with_dopri(frac: float) -> None:
    side = 16.0
    nv, w, vol, basis = small_kset(16, 1, 2.0 * np.pi / side * 4.5)
    vg = O.make_value_grad(nv, w, vol)
    s0 = np.random.default_rng(6).uniform(size=(16, 1))
    full = O.gradient_flow(vg, s0, basis, O.FlowSettings(energy_tol=1e-12))
    t_stop = frac * float(full["t"])
    ref = O.gradient_flow(vg, s0, basis, O.FlowSettings(rtol=1e-10, t_max=t_stop))
    ros = O.gradient_flow(vg, s0, basis, O.FlowSettings(method="rosenbrock", rtol=1e-3, atol=1e-10, t_max=t_stop),
                          hess=lambda s: O.hessian(s, nv, w, vol))
    assert np.max(np.abs(ros["s"] - ref["s"])) < 1e-5


def test_rosenbrock_needs_fewer_steps_when_stiff() -> None:
    a = np.array([[0.3, 0.6]])
    s0 = np.array([[0.4, 0.9]])
    c = np.array([1e4, 1.0])
    vg = _quadratic_vg_axes(a, c)
    dop = O.gradient_flow(vg, s0, np.eye(2), O.FlowSettings(energy_tol=1e-20, max_steps=400_000))
    ros = O.gradient_flow(vg, s0, np.eye(2), O.FlowSettings(method="rosenbrock", rtol=1e-3, atol=1e-8,
                                                            energy_tol=1e-20, max_steps=400_000),
                          hess=_quadratic_hess(c, 1))
    assert dop["status"] == ros["status"] == O.STATUS_ENERGY_TOL
    assert ros["n_steps"] < dop["n_steps"] / 5
    # a Jacobian of the wrong sign (planted in the Hessian callable) loses the stiff mode
    bad = O.gradient_flow(vg, s0, np.eye(2), O.FlowSettings(method="rosenbrock", rtol=1e-3, atol=1e-8,
                                                            energy_tol=1e-20, max_steps=20_000),
                          hess=lambda s: -_quadratic_hess(c, 1)(s))
    assert bad["status"] != O.STATUS_ENERGY_TOL and bad["energy"] > 1e-6
