"""Pins of the oracle's minimizers (descent, two-loop, line search, L-BFGS).

CPU only.  The descent pin is an independent transcription of the C++ loop
(descent.py's own test restates StructureOptimization.cpp:699-724); the
two-loop recursion is pinned by the dense BFGS inverse-Hessian recursion;
the interpolation helpers by scipy and by polynomials with known minima;
L-BFGS by the reference's outcome-level tests.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from systems import chi_kset, small_kset


def _vg(n: int, d: int, k_outer: float):
    nv, w, vol, _ = small_kset(n, d, k_outer)
    return O.make_value_grad(nv, w, vol)


# ---------------- descent ----------------


def _cpp_descent(vg, s0: np.ndarray, grow: float, iters: int, step0: float):
    """Transcription of the C++ loop: recompute f1 and g at the top of every iteration."""
    x = s0.copy()
    step = step0
    xs, acc, steps = [], [], []
    for _ in range(iters):
        f1, g = vg(x)
        norm = float(np.sqrt(np.sum(g * g)))
        g = g / norm
        xn = x - g * step
        f2, _ = vg(xn)
        df = f1 - f2
        df2 = step * norm
        if abs(df - df2) / df2 > 0.01:
            step *= 0.5
            acc.append(False)
        else:
            x = xn
            step *= grow
            acc.append(True)
        xs.append(x.copy())
        steps.append(step)
    return xs, acc, steps


@pytest.mark.parametrize("grow", [1.0, 2.0])
def test_descent_matches_transcribed_cpp(grow: float) -> None:
    vg = _vg(16, 2, 4.0)
    s0 = np.random.default_rng(3).uniform(size=(16, 2))
    iters = 100
    xs, acc, steps = _cpp_descent(vg, s0, grow, iters, 3e-2)
    st = O.DescentSettings(step_init=3e-2, grow=grow, max_evals=2 * iters, record_every=1,
                           max_records=iters + 1)
    res = O.descent(vg, s0, st)
    tr = res.traj
    assert tr is not None and tr["count"] == iters + 1
    assert np.array_equal(tr["step"][1:], np.asarray(steps))
    dev = max(
        float(np.max(np.abs((tr["s"][m + 1] - xs[m]) - np.round(tr["s"][m + 1] - xs[m]))))
        for m in range(iters)
    )
    assert dev <= 1e-14
    assert any(acc) and not all(acc)


def test_descent_budget_and_odd_budget() -> None:
    vg = _vg(12, 2, 4.0)
    s0 = np.random.default_rng(1).uniform(size=(12, 2))
    r = O.descent(vg, s0, O.DescentSettings(max_evals=14))
    assert (r.n_iters, r.n_evals, r.status) == (7, 14, O.STATUS_BUDGET)
    r = O.descent(vg, s0, O.DescentSettings(max_evals=11))
    assert (r.n_iters, r.n_evals, r.status) == (5, 10, O.STATUS_BUDGET)
    r = O.descent(vg, s0, O.DescentSettings(max_evals=1))
    assert (r.n_iters, r.status) == (0, O.STATUS_BUDGET)


def test_descent_step_floor_after_40() -> None:
    """1e-3 * 0.5^40 < 1e-15 <= 1e-3 * 0.5^39 (test_descent.py:153-161)."""
    vg = _vg(12, 2, 4.0)
    s0 = np.random.default_rng(1).uniform(size=(12, 2))
    r = O.descent(vg, s0, O.DescentSettings(rel_tol=1e-30))
    assert r.status == O.STATUS_STEP_FLOOR and r.n_iters == 40
    assert np.array_equal(r.s, O.wrap_fractional(s0))


def test_descent_zero_gradient_is_grad_tol() -> None:
    def vg(s):
        return 1.0, np.zeros_like(s)

    r = O.descent(vg, np.full((4, 2), 0.5), O.DescentSettings())
    assert r.status == O.STATUS_GRAD_TOL and r.n_iters == 0


# ---------------- two-loop recursion vs dense BFGS ----------------


def test_two_loop_equals_dense_bfgs_inverse() -> None:
    rng = np.random.default_rng(7)
    n = 12
    A = rng.normal(size=(n, n))
    A = A @ A.T + n * np.eye(n)
    m = 5
    S = rng.normal(size=(m, n))
    Y = S @ A  # curvature pairs of a quadratic: y = A s
    rho = 1.0 / np.einsum("ij,ij->i", S, Y)
    gamma = float(S[-1] @ Y[-1] / (Y[-1] @ Y[-1]))
    g = rng.normal(size=n)
    # ring with memory 8, newest pair at slot (count - 1) % 8 for count = 5
    mem = 8
    Sr = np.zeros((mem, n))
    Yr = np.zeros((mem, n))
    rr = np.zeros(mem)
    for i in range(m):
        Sr[i], Yr[i], rr[i] = S[i], Y[i], rho[i]
    q = O.lbfgs_precondition(g, Sr, Yr, rr, gamma, m % mem)
    H = gamma * np.eye(n)
    for i in range(m):
        V = np.eye(n) - rho[i] * np.outer(Y[i], S[i])
        H = V.T @ H @ V + rho[i] * np.outer(S[i], S[i])
    assert np.allclose(q, H @ g, rtol=1e-10, atol=1e-12)
    # a transposed update (a plausible slip) does not match
    Hb = gamma * np.eye(n)
    for i in range(m):
        V = np.eye(n) - rho[i] * np.outer(Y[i], S[i])
        Hb = V @ Hb @ V.T + rho[i] * np.outer(S[i], S[i])
    assert not np.allclose(q, Hb @ g, rtol=1e-6)


def test_two_loop_ring_wraps() -> None:
    """With count > memory the oldest pairs are overwritten in ring order."""
    rng = np.random.default_rng(8)
    n, mem = 6, 3
    A = np.diag(np.arange(1.0, n + 1))
    pairs = [rng.normal(size=n) for _ in range(5)]
    Sr, Yr, rr = np.zeros((mem, n)), np.zeros((mem, n)), np.zeros(mem)
    for c, s in enumerate(pairs):
        y = A @ s
        Sr[c % mem], Yr[c % mem], rr[c % mem] = s, y, 1.0 / (s @ y)
    g = rng.normal(size=n)
    q = O.lbfgs_precondition(g, Sr, Yr, rr, 1.0, 5 % mem)
    H = np.eye(n)
    for s in pairs[-3:]:
        y = A @ s
        r = 1.0 / (s @ y)
        V = np.eye(n) - r * np.outer(y, s)
        H = V.T @ H @ V + r * np.outer(s, s)
    assert np.allclose(q, H @ g, rtol=1e-10)


# ---------------- interpolation helpers ----------------


def test_cubicmin_quadmin_known_minima_and_scipy() -> None:
    from scipy.optimize import _linesearch as sl

    # p(x) = (x - 1)(x - 3)^2 has a local minimum at 3
    p = lambda x: (x - 1) * (x - 3) ** 2  # noqa: E731
    dp = lambda x: (x - 3) ** 2 + 2 * (x - 1) * (x - 3)  # noqa: E731
    a, b, c = 2.0, 4.0, 5.0
    assert O.cubicmin(a, p(a), dp(a), b, p(b), c, p(c)) == pytest.approx(3.0, rel=1e-12)
    q = lambda x: 2.0 * (x - 0.7) ** 2 + 1.0  # noqa: E731
    dq = lambda x: 4.0 * (x - 0.7)  # noqa: E731
    assert O.quadmin(0.0, q(0.0), dq(0.0), 2.0, q(2.0)) == pytest.approx(0.7, rel=1e-13)
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, fa, fpa, b, fb, c, fc = rng.normal(size=7)
        ours = O.cubicmin(a, fa, fpa, b, fb, c, fc)
        ref = sl._cubicmin(a, fa, fpa, b, fb, c, fc)
        if ref is None:
            assert not math.isfinite(ours)
        else:
            assert ours == pytest.approx(ref, rel=1e-10)
        assert O.quadmin(a, fa, fpa, b, fb) == pytest.approx(sl._quadmin(a, fa, fpa, b, fb), rel=1e-10)


# ---------------- line search ----------------


def test_zoom_linesearch_satisfies_wolfe_on_quadratic() -> None:
    rng = np.random.default_rng(2)
    n = 8
    A = rng.normal(size=(n, n))
    A = A @ A.T + np.eye(n)
    c = rng.normal(size=n)

    def fg(x):
        d = x - c
        return 0.5 * d @ A @ d, A @ d

    x = np.zeros(n)
    f0, g0 = fg(x)
    for scale in (1e-3, 1.0, 30.0):
        u = -scale * g0
        out = O.zoom_linesearch(fg, x, u, f0, g0, 30)
        a = out.stepsize
        assert out.decrease_error == 0.0
        slope0 = float(g0 @ u)
        sa = float(out.grad @ u)
        armijo = out.value <= f0 + 1e-4 * a * slope0
        approx = (sa >= (2e-4 - 1.0) * slope0) and out.value <= f0 + 1e-6 * abs(f0)
        assert armijo or approx
        assert abs(sa) <= 0.9 * abs(slope0) + 1e-15
        assert np.array_equal(out.x, x + a * u)


def test_zoom_linesearch_nan_everywhere_fails() -> None:
    def fg(x):  # NaN value with a finite descent slope everywhere (the Cliff energy)
        return float("nan"), np.ones_like(x)

    out = O.zoom_linesearch(fg, np.zeros(3), -np.ones(3), 1.0, np.ones(3), 30)
    assert out.num_steps == 30 and not math.isfinite(out.value)


# ---------------- L-BFGS outcomes (test_lbfgs.py:105-271) ----------------


def _quadratic(n: int = 6, seed: int = 0):
    rng = np.random.default_rng(seed)
    m = rng.normal(size=(2 * n, 2 * n))
    a = m @ m.T / (2 * n) + np.eye(2 * n)
    c = rng.uniform(0.3, 0.7, size=(n, 2))

    def vg(s):
        d = (s - c).reshape(-1)
        return float(0.5 * d @ a @ d), (a @ d).reshape(s.shape)

    return vg, c


def test_lbfgs_spd_quadratic() -> None:
    vg, c = _quadratic()
    s0 = np.random.default_rng(5).uniform(0.3, 0.7, size=(6, 2))
    r = O.lbfgs(vg, s0, O.LBFGSSettings())
    assert r.energy < 1e-18
    assert np.max(np.abs(r.s - c)) < 1e-8


def test_lbfgs_passes_equal_chained_single_passes() -> None:
    nv, w, vol, _ = chi_kset(16, 2, 0.25)
    vg = O.make_value_grad(nv, w, vol)
    s0 = np.random.default_rng(0).uniform(size=(16, 2))
    cfg = O.LBFGSSettings(max_evals_per_pass=12)
    full = O.lbfgs(vg, s0, cfg)
    cur = s0
    n_evals, n_iters = 1, 0
    for r in cfg.rescales:
        one = O.lbfgs(vg, cur, O.LBFGSSettings(rescales=(r,), max_evals_per_pass=12))
        cur = one.s
        n_evals += one.n_evals - 1
        n_iters += one.n_iters
    assert full.n_iters == n_iters and full.n_evals == n_evals
    assert np.max(np.abs(full.s - cur)) < 1e-8
    assert 1 + 3 * 12 < full.n_evals <= 1 + 3 * (12 + 30)


def test_lbfgs_eval_count_exact() -> None:
    vg0, _ = _quadratic()
    calls = [0]

    def vg(s):
        calls[0] += 1
        return vg0(s)

    s0 = np.random.default_rng(5).uniform(0.3, 0.7, size=(6, 2))
    r = O.lbfgs(vg, s0, O.LBFGSSettings())
    assert calls[0] == r.n_evals > 3


def test_lbfgs_reaches_ground_state_d2_n64() -> None:
    nv, w, vol, _ = chi_kset(64, 2, 0.25)
    vg = O.make_value_grad(nv, w, vol)
    s0 = np.random.default_rng(0).uniform(size=(64, 2))
    r = O.lbfgs(vg, s0, O.LBFGSSettings())
    assert r.energy < 1e-20
    assert vg(r.s)[0] < 1e-20


def test_lbfgs_statuses() -> None:
    vg, c = _quadratic()
    flat = O.lbfgs(vg, c.copy(), O.LBFGSSettings(rescales=(1.0,), stagnation_patience=1))
    assert flat.status == O.STATUS_STAGNATED
    s0 = np.random.default_rng(5).uniform(0.3, 0.7, size=(6, 2))

    def cliff(s):
        v, g = vg(s)
        return (v if np.array_equal(s, s0) else float("nan")), g

    bad = O.lbfgs(cliff, s0, O.LBFGSSettings(rescales=(1.0,)))
    assert bad.status == O.STATUS_LINESEARCH_FAIL
    assert np.array_equal(bad.s, s0)
    short = O.lbfgs(vg, s0, O.LBFGSSettings(rescales=(1.0,), max_evals_per_pass=3))
    assert short.status == O.STATUS_BUDGET
    e = O.lbfgs(vg, s0, O.LBFGSSettings(energy_tol=1e-6))
    assert e.status == O.STATUS_ENERGY_TOL and e.energy < 1e-6
    gt = O.lbfgs(vg, s0, O.LBFGSSettings(grad_tol=1e-3))
    assert gt.status == O.STATUS_GRAD_TOL and gt.grad_norm < 1e-3
