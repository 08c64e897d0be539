"""NumPy restatement of the stealthy hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module, and only as the checker.

Each function cites the reference file:line (paths relative to the
sdawley1/xla-cco tree) whose algorithm it restates.

Parity pins (what checks each function; tests/test_oracle_*.py):

* ``halfspace_kvectors`` / ``ka_from_chi`` / ``k_outer_from_ka``: the
  reference's known answers (Ka=0.5963271031078479, K=2.6093, M=13 at d=3
  N=100; M=28/61/89/101/128 at chi 0.1/0.2/0.3/0.35/0.45), the C++ K bits
  ``0x1.24329cee12d94p+0`` and the 50 n-vectors of the C++ MD fixture, and the
  closed-form count of a cubic lattice shell.
* ``energy_grad`` / ``energy_loops``: the C++ energies of the 15 stored MD
  configurations (tests/golden/md_cxx_*.npz, relative 1e-12), the
  single-particle closed form Phi = sum(w)/V, a lattice below its first shell
  (Phi ~ 0), a central finite difference of the energy, translation
  invariance (sum of gradients ~ 0) and the C++ Cartesian force relation.
* ``hessian``: central finite differences of ``energy_grad`` (the FD pin of
  the gradient carries over), symmetry, and the translation null space (rows
  sum to zero over the particles of each axis, as Phi is translation
  invariant), which a dropped diagonal term or a transposed block fails.
* ``descent``: an independent transcription of the C++ loop (the one that
  recomputes f1 and g every iteration, test_descent.py:52-89), the 40-step
  floor (1e-3 * 0.5^40 < 1e-15) and the budget rules (max_evals // 2).
* ``lbfgs_precondition``: the dense BFGS inverse-Hessian recursion
  H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T from H0 = gamma I
  (a closed form the two-loop recursion must equal).
* ``cubicmin`` / ``quadmin``: exact minimisers of polynomials built with known
  roots, and scipy.optimize's ``_cubicmin``/``_quadmin``.
* ``zoom_linesearch``: the accepted step satisfies the strong-Wolfe or the
  approximate-decrease condition it claims; a convex quadratic's exact
  minimiser along the line is bracketed.
* ``md_lane`` / ``md_run`` (schemes, tracked waves, winding, records):
  Verlet equals ``verlet_nve``; BAOAB at gamma = 0 and Andersen at p = 0 equal
  Verlet; the count mode fires exactly every 1/p steps; ballistic motion
  (no forces, one common velocity u) winds every wave by wrap(k . u dt) per
  step (wind_every steps at once for the untracked-every-step rows) and by
  zero once COM-corrected, with S constant; BAOAB samples <x^2> = T / k of a
  harmonic well at a large step (exact for BAOAB); snapshots and trace rows
  equal the states of shorter runs; a drift stop freezes the lane.
* ``lbfgs``: the reference's outcome-level tests (SPD quadratic to 1e-18 and
  within 1e-8 of the minimum, d=2 N=64 chi=0.25 below 1e-20, three passes equal
  three chained single passes, exact evaluation count, stagnation and
  line-search-failure statuses).  The per-iterate arithmetic inside optax
  0.2.8 (not vendored, not installed) is **parity unpinned**: this oracle
  restates its published algorithm (Nocedal & Wright Alg. 3.5/3.6 with
  Hager-Zhang approximate decrease) and the HIP path is checked against this
  restatement.
"""

from __future__ import annotations

import dataclasses
import itertools
import math
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

TWO_PI = 2.0 * math.pi

STATUS_RUNNING = 0
STATUS_ENERGY_TOL = 1
STATUS_STEP_FLOOR = 2
STATUS_BUDGET = 3
STATUS_STAGNATED = 4
STATUS_LINESEARCH_FAIL = 5
STATUS_GRAD_TOL = 6
STATUS_DRIFT = 7

ValueGrad = Callable[[np.ndarray], Tuple[float, np.ndarray]]

# ============================================================
# K-SET (kvectors.py:60-169, tests/oracle.py:106-150)
# ============================================================

_CPP_UNIT_BALL = {1: 2.0, 2: math.pi, 3: 4.1887902047863909846, 4: 4.9348022005446793094}


def unit_ball_volume(dim: int) -> float:
    """C++ HyperSphere_C literals (kvectors.py:131-149)."""
    if dim in _CPP_UNIT_BALL:
        return _CPP_UNIT_BALL[dim]
    return float(math.pi ** (dim / 2.0) / math.gamma(dim / 2.0 + 1.0))


def k_outer_from_ka(ka: float, phi: float, dim: int) -> float:
    """kvectors.py:152-156."""
    return float(ka / (phi / unit_ball_volume(dim)) ** (1.0 / dim))


def ka_from_chi(chi: float, phi: float, dim: int) -> float:
    """kvectors.py:159-163."""
    arg = 2.0 ** (dim + 1) * dim * chi * math.gamma(dim / 2.0 + 1.0) ** 2 * phi
    return float(arg ** (1.0 / dim))


def modulus2(k: np.ndarray) -> float:
    """|k|^2 summed in index order (tests/oracle.py:106-111)."""
    acc = 0.0
    for c in range(k.shape[0]):
        acc += float(k[c]) * float(k[c])
    return acc


def halfspace_kvectors(
    basis: np.ndarray, k_inner: float, k_outer: float
) -> Tuple[np.ndarray, np.ndarray]:
    """Brute-force integer scan (tests/oracle.py:114-150), sorted like the KSet.

    The Cartesian k is summed in index order (kvectors.py:82-85) so the strict
    shell test sees the same float64 value as the reference; rows are then
    ordered by (k_sq, n lexicographic) (kvectors.py:91-93).
    """
    dim = basis.shape[0]
    recip = TWO_PI * np.linalg.inv(basis).T
    bounds = [
        int(math.floor(k_outer * float(np.linalg.norm(basis[i])) / TWO_PI)) + 1
        for i in range(dim)
    ]
    rows: List[Tuple[float, Tuple[int, ...]]] = []
    for n in itertools.product(*[range(-b, b + 1) for b in bounds]):
        nz = [c for c in n if c != 0]
        if not nz or nz[0] < 0:
            continue
        k = np.zeros(dim)
        for j in range(dim):
            acc = 0.0
            for i in range(dim):
                acc = acc + float(n[i]) * float(recip[i, j])
            k[j] = acc
        q = modulus2(k)
        if k_inner * k_inner < q < k_outer * k_outer:
            rows.append((q, tuple(n)))
    rows.sort()
    n_out = np.array([r[1] for r in rows], dtype=np.int64).reshape(-1, dim)
    return n_out, np.array([r[0] for r in rows], dtype=np.float64)


# ============================================================
# ENERGY AND FORCES (energy.py:45-50, 186-225, 301-324)
# ============================================================


def energy_loops(s: np.ndarray, n: np.ndarray, w: np.ndarray, volume: float) -> float:
    """Pure-Python double loop (tests/oracle.py:38-68); tiny inputs only."""
    total = 0.0
    for kk in range(n.shape[0]):
        rho = 0.0 + 0.0j
        for j in range(s.shape[0]):
            ph = TWO_PI * float(np.dot(n[kk], s[j]))
            rho += complex(math.cos(ph), math.sin(ph))
        total += float(w[kk]) * (rho.real**2 + rho.imag**2)
    return total / volume


def density(s: np.ndarray, n: np.ndarray, chunk: int = 512) -> Tuple[np.ndarray, np.ndarray]:
    """(Re rho, Im rho) per k row with rho_k = sum_j exp(2 pi i n_k.s_j).

    energy.py:45-50 (``_density_block``): theta = 2 pi (n @ s^T), row sums of
    cos and sin.  Chunked over k so the (c, N) block stays small.
    """
    n = np.asarray(n, dtype=np.float64)
    re = np.empty(n.shape[0])
    im = np.empty(n.shape[0])
    for lo in range(0, n.shape[0], chunk):
        th = TWO_PI * (n[lo : lo + chunk] @ s.T)
        re[lo : lo + chunk] = np.cos(th).sum(axis=1)
        im[lo : lo + chunk] = np.sin(th).sum(axis=1)
    return re, im


def energy_grad(
    s: np.ndarray, n: np.ndarray, w: np.ndarray, volume: float, chunk: int = 512
) -> Tuple[float, np.ndarray, np.ndarray, np.ndarray]:
    """(Phi, dPhi/ds, Re rho, Im rho): energy.py:206-225 with ``_gradient`` :301-324.

    dPhi/ds_j = (2 (2 pi) / V) sum_k w_k n_k (Im rho_k cos th_kj - Re rho_k sin th_kj).
    """
    n = np.asarray(n, dtype=np.float64)
    re, im = density(s, n, chunk)
    value = float(np.sum(w * (re * re + im * im)) / volume)
    acc = np.zeros_like(s)
    for lo in range(0, n.shape[0], chunk):
        nc = n[lo : lo + chunk]
        th = TWO_PI * (s @ nc.T)  # (N, c)
        coef = w[lo : lo + chunk] * (im[lo : lo + chunk] * np.cos(th) - re[lo : lo + chunk] * np.sin(th))
        acc = acc + coef @ nc
    return value, (2.0 * TWO_PI / volume) * acc, re, im


def hessian(s: np.ndarray, n: np.ndarray, w: np.ndarray, volume: float) -> np.ndarray:
    """Closed-form Hessian d^2 Phi / ds_ja ds_lb, flattened row-major to (N d, N d).

    energy.py:272-299: with theta_kj = 2 pi n_k.s_j,
    H[(j,a),(l,b)] = (2 (2 pi)^2 / V) sum_k w_k n_ka n_kb
                     [cos th_kj cos th_kl + sin th_kj sin th_kl
                      - delta_jl (Re rho_k cos th_kj + Im rho_k sin th_kj)].
    """
    s = np.asarray(s, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    N, d = s.shape
    th = TWO_PI * (s @ n.T)  # (N, M)
    c, sn = np.cos(th), np.sin(th)
    re, im = c.sum(axis=0), sn.sum(axis=0)
    pref = 2.0 * TWO_PI**2 / volume
    core = re * c + im * sn  # (N, M)
    h = np.zeros((N, d, N, d))
    for a in range(d):
        for b in range(d):
            wab = w * n[:, a] * n[:, b]
            pair = (c * wab) @ c.T + (sn * wab) @ sn.T
            h[:, a, :, b] = pref * (pair - np.diag(core @ wab))
    return h.reshape(N * d, N * d)


def cartesian_forces(
    s: np.ndarray, n: np.ndarray, w: np.ndarray, basis: np.ndarray, volume: float
) -> np.ndarray:
    """C++ forces F = -dPhi/dr (tests/oracle.py:71-98), loop over k."""
    recip = TWO_PI * np.linalg.inv(basis).T
    k = np.asarray(n, dtype=np.float64) @ recip
    r = s @ basis
    out = np.zeros_like(r)
    for kk in range(k.shape[0]):
        ph = r @ k[kk]
        rho = complex(np.cos(ph).sum(), np.sin(ph).sum())
        coef = rho.imag * np.cos(ph) - rho.real * np.sin(ph)
        out += -(2.0 * float(w[kk]) / volume) * coef[:, None] * k[kk][None, :]
    return out


def make_value_grad(n: np.ndarray, w: np.ndarray, volume: float) -> ValueGrad:
    """Closure (s) -> (Phi, dPhi/ds) over one k-set."""

    def vg(s: np.ndarray) -> Tuple[float, np.ndarray]:
        f, g, _, _ = energy_grad(s, n, w, volume)
        return f, g

    return vg


def wrap_fractional(s: np.ndarray) -> np.ndarray:
    """box.py:127-139."""
    return np.minimum(s - np.floor(s), np.nextafter(1.0, 0.0))


# ============================================================
# ADAPTIVE DESCENT (minimize/descent.py:52-232)
# ============================================================


@dataclasses.dataclass(frozen=True)
class DescentSettings:
    step_init: float = 1e-3
    step_min: float = 1e-15
    shrink: float = 0.5
    grow: float = 1.0
    rel_tol: float = 0.01
    max_evals: int = 10_000
    energy_tol: float = 0.0
    record_every: int = 0
    max_records: int = 0


def _descent_status(f: float, g: np.ndarray, step: float, it: int, st: DescentSettings) -> int:
    """Precedence energy tol > zero grad > step floor > budget (descent.py:92-104)."""
    code = STATUS_BUDGET if it >= st.max_evals // 2 else STATUS_RUNNING
    if step <= st.step_min:
        code = STATUS_STEP_FLOOR
    if np.all(g == 0.0):
        code = STATUS_GRAD_TOL
    if f < st.energy_tol:
        code = STATUS_ENERGY_TOL
    return code


@dataclasses.dataclass
class RelaxOut:
    s: np.ndarray
    energy: float
    grad_norm: float
    n_evals: int
    n_iters: int
    status: int
    traj: Optional[Dict[str, np.ndarray]] = None


def descent(vg: ValueGrad, s0: np.ndarray, st: DescentSettings) -> RelaxOut:
    """One configuration (descent.py:121-197): init, iterate to a status, finalize."""
    s = s0.copy()
    f, g = vg(s)
    step = float(st.step_init)
    it = 0
    n_evals = 0
    status = _descent_status(f, g, step, it, st)
    rec = _Recorder(st.record_every, st.max_records, s, f, step)
    while status == STATUS_RUNNING:
        norm = math.sqrt(float(np.sum(g * g)))
        gn = g / norm if norm > 0.0 else 0.0 * g
        xn = s - gn * step
        f2, g2 = vg(xn)
        df = f - f2
        df2 = step * norm
        with np.errstate(invalid="ignore", divide="ignore"):
            reject = bool(abs(df - df2) / df2 > st.rel_tol)  # NaN accepts
        if reject:
            step = step * st.shrink
        else:
            step = step * st.grow
            s, f, g = xn, f2, g2
        it += 1
        n_evals += 2
        status = _descent_status(f, g, step, it, st)
        rec.add(it, s, f, step)
    return RelaxOut(
        s=wrap_fractional(s),
        energy=f,
        grad_norm=math.sqrt(float(np.sum(g * g))),
        n_evals=n_evals,
        n_iters=it,
        status=status,
        traj=rec.result(),
    )


class _Recorder:
    """Trajectory rows every ``every`` iterations (minimize/_loop.py:189-248)."""

    def __init__(self, every: int, cap: int, s: np.ndarray, f: float, step: float):
        self.on = every > 0 and cap > 0
        self.every, self.cap = every, cap
        self.rows: List[Tuple[int, np.ndarray, float, float]] = []
        if self.on:
            self.rows.append((0, wrap_fractional(s), f, step))

    def add(self, it: int, s: np.ndarray, f: float, step: float) -> None:
        if self.on and it % self.every == 0 and len(self.rows) < self.cap:
            self.rows.append((it, wrap_fractional(s), f, step))

    def result(self) -> Optional[Dict[str, np.ndarray]]:
        if not self.on:
            return None
        return {
            "iteration": np.array([r[0] for r in self.rows], dtype=np.int64),
            "s": np.array([r[1] for r in self.rows]),
            "energy": np.array([r[2] for r in self.rows]),
            "step": np.array([r[3] for r in self.rows]),
            "count": len(self.rows),
        }


# ============================================================
# L-BFGS (minimize/lbfgs.py:55-350 + optax 0.2.8 lbfgs / zoom line search)
# ============================================================


@dataclasses.dataclass(frozen=True)
class LBFGSSettings:
    memory: int = 100
    rescales: Tuple[float, ...] = (1e4, 1.0, 1e-2)
    max_evals_per_pass: int = 500_000
    energy_tol: float = 0.0
    grad_tol: float = 0.0
    stagnation_patience: int = 50
    max_linesearch_steps: int = 30
    record_every: int = 0
    max_records: int = 0


def vdot(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.dot(a.ravel(), b.ravel()))


def lbfgs_precondition(
    g: np.ndarray,
    S: np.ndarray,
    Y: np.ndarray,
    rho: np.ndarray,
    gamma: float,
    memory_idx: int,
) -> np.ndarray:
    """Two-loop recursion over a ring of m pairs (optax ``_precondition_by_lbfgs``).

    Slots are visited from ``memory_idx`` (oldest) around the ring; the first
    loop runs newest -> oldest, the second oldest -> newest.  Unused slots have
    rho = 0 and zero vectors, so they contribute exactly nothing.
    """
    m = rho.shape[0]
    order = [(memory_idx + t) % m for t in range(m)]
    q = g.copy()
    alphas = np.zeros(m)
    for t in reversed(range(m)):
        i = order[t]
        a = rho[i] * vdot(S[i], q)
        q = q + (-a) * Y[i]
        alphas[t] = a
    q = gamma * q
    for t in range(m):
        i = order[t]
        b = rho[i] * vdot(Y[i], q)
        q = q + (alphas[t] - b) * S[i]
    return q


def cubicmin(a: float, fa: float, fpa: float, b: float, fb: float, c: float, fc: float) -> float:
    """Critical point of the cubic through (a,fa),(b,fb),(c,fc) with slope fpa at a.

    The same closed form as scipy.optimize's line-search helper; NaN when the
    radical is negative or the points are degenerate (the caller then rejects it).
    """
    with np.errstate(all="ignore"):
        C = fpa
        db = b - a
        dc = c - a
        denom = (db * dc) ** 2 * (db - dc)
        d1 = np.array([[dc**2, -(db**2)], [-(dc**3), db**3]])
        d2 = np.array([fb - fa - C * db, fc - fa - C * dc])
        A, B = (d1 @ d2) / denom
        radical = B * B - 3.0 * A * C
        return float(a + (-B + np.sqrt(radical)) / (3.0 * A))


def quadmin(a: float, fa: float, fpa: float, b: float, fb: float) -> float:
    """Minimiser of the parabola through (a,fa),(b,fb) with slope fpa at a."""
    with np.errstate(all="ignore"):
        D = fa
        C = fpa
        db = b - a
        B = (fb - D - C * db) / (db * db)
        return float(a - C / (2.0 * B))


@dataclasses.dataclass
class LineSearchOut:
    stepsize: float
    x: np.ndarray
    value: float
    grad: np.ndarray
    num_steps: int
    decrease_error: float


# optax.scale_by_zoom_linesearch defaults as restated for this path.
LS_SLOPE_RTOL = 1e-4
LS_CURV_RTOL = 0.9
LS_APPROX_DEC_RTOL = 1e-6
LS_INCREASE = 2.0
# The zoom phase stops once the bracket |high - low| is no wider than this.  optax's
# zoom search (ported from jaxopt) takes 1e-10 for 64-bit step sizes and 1e-5 for
# narrower ones; the positions here are float64.  Measured on the reference's own
# benchmark inputs (default_rng(0), d=3, N=100, 1000 lanes, max_evals_per_pass=200):
# total evaluations 451.6k / 461.8k / 461.0k / 457.5k / 561.5k at chi 0.05 / 0.1 /
# 0.2 / 0.3 / 0.45 against the reference's 455.4k / 461.7k / 458.8k / 457.5k / 557.0k
# (benchmarks/results_threshold_gpu.json); the old 1e-5 floor gave -6% to -22%.
LS_INTERVAL_MIN = 1e-10
LS_TOL = 0.0


def _decrease_error(a: float, va: float, sa: float, v0: float, s0: float) -> float:
    """Armijo error, relaxed by the Hager-Zhang approximate decrease; >= 0."""
    with np.errstate(all="ignore"):
        dec = va - (v0 + LS_SLOPE_RTOL * a * s0)
        approx = sa - (2.0 * LS_SLOPE_RTOL - 1.0) * s0
        delta = va - v0 - LS_APPROX_DEC_RTOL * abs(v0)
        approx = _nanmax(approx, delta)
        dec = _nanmin(approx, dec)
        return _nanmax(dec, 0.0)


def _curvature_error(sa: float, s0: float) -> float:
    return _nanmax(abs(sa) - LS_CURV_RTOL * abs(s0), 0.0)


def _nanmax(a: float, b: float) -> float:
    return float("nan") if (math.isnan(a) or math.isnan(b)) else max(a, b)


def _nanmin(a: float, b: float) -> float:
    return float("nan") if (math.isnan(a) or math.isnan(b)) else min(a, b)


def zoom_linesearch(
    fg: ValueGrad,
    x: np.ndarray,
    u: np.ndarray,
    f0: float,
    g0: np.ndarray,
    max_steps: int,
) -> LineSearchOut:
    """Strong-Wolfe zoom line search with initial guess 1 (optax 0.2.8 semantics).

    Bracketing phase (N&W Alg. 3.5) doubles the step from 1; the zoom phase
    (Alg. 3.6) picks cubic, then quadratic, then bisection trial points with
    the 0.2 / 0.1 interior safeguards.  One value-and-gradient per step.  On
    failure the last step with sufficient decrease is used when one exists
    (reporting decrease error 0), else the last trial point.
    """
    slope0 = vdot(u, g0)
    count = 0
    interval_found = done = failed = False
    prev = (0.0, f0, slope0)
    low = (0.0, f0, slope0)
    high = (0.0, f0, slope0)
    cref = (0.0, f0)
    safe: Tuple[float, float, np.ndarray, np.ndarray] = (0.0, f0, g0, x)
    last: Tuple[float, float, np.ndarray, np.ndarray, float] = (0.0, f0, g0, x, math.inf)
    while not (done or failed):
        if not interval_found:
            a = 1.0 if count == 0 else LS_INCREASE * prev[0]
            xa = x + a * u
            va, ga = fg(xa)
            sa = vdot(ga, u)
            dec = _decrease_error(a, va, sa, f0, slope0)
            err = _nanmax(dec, _curvature_error(sa, slope0))
            if dec <= LS_TOL:
                safe = (a, va, ga, xa)
            set_high = (dec > 0.0) or (va >= prev[1] and count > 0)
            set_low = (sa >= 0.0) and not set_high
            if set_low:
                low, high = (a, va, sa), prev
            else:
                low, high = prev, (a, va, sa)
            interval_found = set_high or set_low or err <= LS_TOL
            done = err <= LS_TOL
            cref = (low[0], low[1])
            prev = (a, va, sa)
            count += 1
            failed = count >= max_steps and not done
        else:
            delta = abs(high[0] - low[0])
            left, right = min(high[0], low[0]), max(high[0], low[0])
            mc = cubicmin(low[0], low[1], low[2], high[0], high[1], cref[0], cref[1])
            use_cubic = (mc > left + 0.2 * delta) and (mc < right - 0.2 * delta)
            mq = quadmin(low[0], low[1], low[2], high[0], high[1])
            use_quad = (not use_cubic) and (mq > left + 0.1 * delta) and (mq < right - 0.1 * delta)
            if use_cubic:
                a = mc
            elif use_quad:
                a = mq
            else:
                a = (low[0] + high[0]) / 2.0
            xa = x + a * u
            va, ga = fg(xa)
            sa = vdot(ga, u)
            dec = _decrease_error(a, va, sa, f0, slope0)
            err = _nanmax(dec, _curvature_error(sa, slope0))
            if dec <= LS_TOL and va < safe[1]:
                safe = (a, va, ga, xa)
            done = err <= LS_TOL
            hi_to_mid = (dec > 0.0) or (va >= low[1])
            hi_to_low = (sa * (high[0] - low[0]) >= 0.0) and not hi_to_mid
            old_low, old_high = low, high
            if hi_to_mid:
                high = (a, va, sa)
            elif hi_to_low:
                high = old_low
            if not hi_to_mid:
                low = (a, va, sa)
            cref = (old_high[0], old_high[1]) if (hi_to_mid or hi_to_low) else (old_low[0], old_low[1])
            prev = (a, va, sa)
            count += 1
            failed = (count >= max_steps or delta <= LS_INTERVAL_MIN) and not done
        last = (a, va, ga, xa, dec)
    if failed and safe[0] > 0.0:
        return LineSearchOut(safe[0], safe[3], safe[1], safe[2], count, 0.0)
    return LineSearchOut(last[0], last[3], last[1], last[2], count, last[4])


def lbfgs(vg: ValueGrad, s0: np.ndarray, st: LBFGSSettings) -> RelaxOut:
    """Three-pass L-BFGS on x = R s (minimize/lbfgs.py:163-313).

    Per iteration: memory update and two-loop (optax ``scale_by_lbfgs``),
    direction u = -H g, zoom line search on Phi(x / R), then the reference's
    pass logic: a failed search keeps the previous point, stagnation counter,
    per-pass budget, rescale advance with a fresh optimizer state.
    """
    rescales = [float(r) for r in st.rescales]
    n_pass = len(rescales)
    m = st.memory

    def phi_x(r: float) -> ValueGrad:
        def f(x: np.ndarray) -> Tuple[float, np.ndarray]:
            v, gs = vg(x / r)
            return v, gs / r

        return f

    s = s0.copy()
    r = rescales[0]
    f, gs = vg(s0)
    x = s0 * r
    gx = gs / r
    status = STATUS_ENERGY_TOL if f < st.energy_tol else STATUS_RUNNING
    n_evals, it, pass_idx, pass_evals, stag = 1, 0, 0, 0, 0
    last_step = 0.0
    rec = _Recorder(st.record_every, st.max_records, s0, f, 0.0)

    def fresh() -> Dict[str, object]:
        return {
            "count": 0,
            "px": np.zeros_like(s0),
            "pg": np.zeros_like(s0),
            "S": np.zeros((m,) + s0.shape),
            "Y": np.zeros((m,) + s0.shape),
            "rho": np.zeros(m),
        }

    opt = fresh()
    while status == STATUS_RUNNING:
        r = rescales[pass_idx]
        count = int(opt["count"])  # type: ignore[arg-type]
        S, Y, rho = opt["S"], opt["Y"], opt["rho"]
        if count > 0:
            ds = x - opt["px"]  # type: ignore[operator]
            dy = gx - opt["pg"]  # type: ignore[operator]
            sy = vdot(dy, ds)
            slot = (count - 1) % m
            S[slot] = ds  # type: ignore[index]
            Y[slot] = dy  # type: ignore[index]
            rho[slot] = 0.0 if sy == 0.0 else 1.0 / sy  # type: ignore[index]
            yy = vdot(dy, dy)
            gamma = sy / yy if yy > 0.0 else 1.0
        else:
            gnorm0 = math.sqrt(vdot(gx, gx))
            if math.isnan(gnorm0):
                gamma = float("nan")
            else:
                gamma = min(1.0, 1.0 / gnorm0) if gnorm0 > 0.0 else 1.0
        q = lbfgs_precondition(gx, S, Y, rho, gamma, count % m)  # type: ignore[arg-type]
        opt["px"], opt["pg"], opt["count"] = x, gx, count + 1
        u = -q
        ls = zoom_linesearch(phi_x(r), x, u, f, gx, st.max_linesearch_steps)
        fail = (ls.decrease_error > 0.0) or not math.isfinite(ls.value)
        evals = max(ls.num_steps, 1)
        f_old = f
        if not fail:
            x, s, f, gx = ls.x, ls.x / r, ls.value, ls.grad
        stag = 0 if f < f_old else stag + 1
        pass_evals += evals
        gnorm = math.sqrt(vdot(gx, gx)) * r
        glob = STATUS_GRAD_TOL if gnorm < st.grad_tol else STATUS_RUNNING
        if f < st.energy_tol:
            glob = STATUS_ENERGY_TOL
        local = STATUS_BUDGET if pass_evals >= st.max_evals_per_pass else STATUS_RUNNING
        if stag >= st.stagnation_patience:
            local = STATUS_STAGNATED
        if fail:
            local = STATUS_LINESEARCH_FAIL
        last = pass_idx >= n_pass - 1
        pass_end = local != STATUS_RUNNING and glob == STATUS_RUNNING
        advance = pass_end and not last
        status = glob if glob != STATUS_RUNNING else STATUS_RUNNING
        if pass_end and last:
            status = local
        if advance:
            r_next = rescales[pass_idx + 1]
            x = s * r_next
            gx = gx * r / r_next
            opt = fresh()
            pass_idx += 1
            pass_evals = 0
            stag = 0
        last_step = ls.stepsize
        it += 1
        n_evals += evals
        rec.add(it, s, f, last_step)
    r = rescales[pass_idx]
    return RelaxOut(
        s=wrap_fractional(s),
        energy=f,
        grad_norm=math.sqrt(vdot(gx, gx)) * r,
        n_evals=n_evals,
        n_iters=it,
        status=status,
        traj=rec.result(),
    )


def relax(
    vg: ValueGrad,
    s0: np.ndarray,
    settings: object,
) -> List[RelaxOut]:
    """Batched relax == per-configuration minimizer, in input order.

    minimize/__init__.py:102-151: lanes are independent; host compaction only
    changes which lanes are stepped, never their arithmetic.
    """
    if isinstance(settings, DescentSettings):
        return [descent(vg, s0[i], settings) for i in range(s0.shape[0])]
    if isinstance(settings, LBFGSSettings):
        return [lbfgs(vg, s0[i], settings) for i in range(s0.shape[0])]
    raise TypeError(f"unsupported settings type {type(settings).__name__}")


# ============================================================
# GRADIENT FLOW (minimize/flow.py:60-481, DOPRI5)
# ============================================================

_DP_A = (
    (),
    (1 / 5,),
    (3 / 40, 9 / 40),
    (44 / 45, -56 / 15, 32 / 9),
    (19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729),
    (9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656),
)
_DP_B5 = (35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0)
_DP_B4 = (5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40)
_DP_E = tuple(b5 - b4 for b5, b4 in zip(_DP_B5, _DP_B4))


@dataclasses.dataclass(frozen=True)
class FlowSettings:
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


_ROS_GAMMA = 1.0 + 1.0 / math.sqrt(2.0)


def gradient_flow(
    vg: ValueGrad,
    s0: np.ndarray,
    basis: np.ndarray,
    st: FlowSettings,
    hess: Optional[Callable[[np.ndarray], np.ndarray]] = None,
) -> Dict[str, object]:
    """Gradient flow dr/dt = -grad_r Phi (flow.py:212-481), one configuration.

    ``st.method`` "dopri5": Dormand-Prince 5(4) (flow.py:346-364).
    "rosenbrock": ROS2 (flow.py:367-398) with the exact Hessian ``hess(s)``
    (row-major (n d, n d)): W = I - gamma h J, J = dv/ds = -H (I (x) inv(B B^T)),
    (W) k1 = v(s), (W) k2 = v(s + h k1) - 2 k1, s+ = s + h (3/2 k1 + 1/2 k2),
    error W^-1 (h (k1 + k2) / 2), order 2, clocks (h, |chord|).  Step control
    (flow.py:401-451) is shared.  Returns a dict with s, energy, t, length,
    dt, n_steps, n_evals, status and, when recording, the trajectory rows
    (s, t, length, energy).
    """
    rosenbrock = st.method == "rosenbrock"
    if rosenbrock and hess is None:
        raise ValueError("method='rosenbrock' needs the Hessian")
    arclength = st.parametrization == "arclength"
    metric_inv = np.linalg.inv(basis @ basis.T)
    inv_t = np.linalg.inv(basis).T

    def velocity(s):
        f, g = vg(s)
        v = -g @ metric_inv
        n = math.sqrt(float(np.sum((g @ inv_t) ** 2)))
        if not arclength:
            return f, g, v, np.array([1.0, n]), n
        inv = 1.0 / n if n > 0.0 else 0.0
        return f, g, v * inv, np.array([inv, 1.0]), n

    def status(f, g, gn, t, dt, steps, tries):
        code = STATUS_STEP_FLOOR if dt < st.dt_min else STATUS_RUNNING
        if t >= st.t_max or steps >= st.max_steps or tries >= 4 * st.max_steps:
            code = STATUS_BUDGET
        if np.all(g == 0.0) or gn < st.grad_tol:
            code = STATUS_GRAD_TOL
        if f < st.energy_tol:
            code = STATUS_ENERGY_TOL
        return code

    def scaled_error(err_s, step):
        move = math.sqrt(float(np.mean((step @ basis) ** 2)))
        tol = st.atol + st.rtol * move
        return math.sqrt(float(np.mean((err_s @ basis) ** 2))) / tol

    s = s0.copy()
    ev = velocity(s)
    t = length = 0.0
    dt = min(st.dt_init, st.dt_max)
    steps = tries = 0
    n_evals = 1
    code = status(ev[0], ev[1], ev[4], t, dt, steps, tries)
    rows = [(s.copy(), 0.0, 0.0, ev[0])] if st.record_every > 0 else []
    while code == STATUS_RUNNING:
        h = dt if arclength else min(dt, st.t_max - t)
        if rosenbrock:
            n, d = s.shape
            H = hess(s).reshape(n, d, n, d)
            jac = -np.einsum("ibjc,ba->iajc", H, metric_inv).reshape(n * d, n * d)
            W = np.eye(n * d) - _ROS_GAMMA * h * jac
            k1 = np.linalg.solve(W, ev[2].reshape(-1)).reshape(n, d)
            ev1 = velocity(s + h * k1)
            k2 = np.linalg.solve(W, (ev1[2] - 2.0 * k1).reshape(-1)).reshape(n, d)
            s_new = s + h * (1.5 * k1 + 0.5 * k2)
            ev_new = velocity(s_new)
            chord = math.sqrt(float(np.sum(((s_new - s) @ basis) ** 2)))
            err_f = np.linalg.solve(W, (0.5 * h * (k1 + k2)).reshape(-1)).reshape(n, d)
            err = scaled_error(err_f, s_new - s)
            with np.errstate(all="ignore"):
                factor = float(np.clip(0.9 * err ** (-1.0 / 2), 0.2, 5.0))
            if not math.isfinite(factor):
                factor = 0.2
            accept = err <= 1.0
            dt = min(h * factor, st.dt_max)
            if accept:
                s, ev = s_new, ev_new
                t += h
                length += chord
                steps += 1
                if st.record_every > 0 and steps % st.record_every == 0 and len(rows) < st.max_records:
                    rows.append((s.copy(), t, length, ev[0]))
            tries += 1
            n_evals += 2
            code = status(ev[0], ev[1], ev[4], t, dt, steps, tries)
            continue
        evs = [ev]
        for i in range(1, 6):
            incr = sum(a * e[2] for a, e in zip(_DP_A[i], evs))
            evs.append(velocity(s + h * incr))
        s_new = s + h * sum(b * e[2] for b, e in zip(_DP_B5[:6], evs))
        tl_new = h * sum(b * e[3] for b, e in zip(_DP_B5[:6], evs))
        ev_new = velocity(s_new)
        err_s = h * sum(e * x[2] for e, x in zip(_DP_E, evs + [ev_new]))
        err_tl = h * sum(e * x[3] for e, x in zip(_DP_E, evs + [ev_new]))
        err = scaled_error(err_s, s_new - s)
        tol_tl = st.atol + st.rtol * np.abs(tl_new)
        with np.errstate(all="ignore"):
            err = _nanmax(err, float(np.max(np.abs(err_tl) / tol_tl)))
            factor = float(np.clip(0.9 * err ** (-1.0 / 5), 0.2, 5.0))
        if not math.isfinite(factor):
            factor = 0.2
        accept = err <= 1.0
        dt = min(h * factor, st.dt_max)
        if accept:
            s, ev = s_new, ev_new
            t += tl_new[0]
            length += tl_new[1]
            steps += 1
            if st.record_every > 0 and steps % st.record_every == 0 and len(rows) < st.max_records:
                rows.append((s.copy(), t, length, ev[0]))
        tries += 1
        n_evals += 6
        code = status(ev[0], ev[1], ev[4], t, dt, steps, tries)
    out: Dict[str, object] = {
        "s": s, "energy": ev[0], "t": t, "length": length, "dt": dt,
        "n_steps": steps, "n_evals": n_evals, "status": code,
    }
    if st.record_every > 0:
        last = rows[-1]
        if not (last[1] == t and last[2] == length):
            if len(rows) < st.max_records:
                rows.append((s.copy(), t, length, ev[0]))
            else:
                rows[-1] = (s.copy(), t, length, ev[0])
        out["traj"] = rows
    return out


# ============================================================
# VELOCITY VERLET (dynamics/integrators.py:83-96, run.py:219-442)
# ============================================================


def verlet_nve(
    vg: ValueGrad,
    basis: np.ndarray,
    r0: np.ndarray,
    v0: np.ndarray,
    dt: float,
    n_steps: int,
    temperature: float = 1.0,
    drift_tol: float = math.inf,
) -> Dict[str, object]:
    """NVE Verlet with forces at s = r inv(B), F = -g inv(B)^T; drift guard and sums."""
    inv = np.linalg.inv(basis)
    inv_t = inv.T
    n, d = r0.shape

    def force(r):
        f, g = vg(r @ inv)
        return f, -g @ inv_t

    r, v = r0.copy(), v0.copy()
    phi, f = force(r)
    e0 = phi + 0.5 * float(np.sum(v * v))
    half = 0.5 * dt
    status = STATUS_RUNNING
    max_drift = 0.0
    sums = [0.0, 0.0, 0.0]
    step = 0
    hist = []
    for _ in range(n_steps):
        if status != STATUS_RUNNING:
            break
        v_half = v + half * f
        r = r + dt * v_half
        phi, f = force(r)
        v = v_half + half * f
        ke = 0.5 * float(np.sum(v * v))
        drift = abs(phi + ke - e0)
        if math.isfinite(drift_tol):
            limit = drift_tol * 0.5 * d * (n - 1) * temperature
            if not drift <= limit:
                status = STATUS_DRIFT
        max_drift = _nanmax(max_drift, drift)
        sums[0] += ke
        sums[1] += phi
        sums[2] += phi * phi
        step += 1
        hist.append((phi, ke))
    return {
        "r": r, "v": v, "f": f, "phi": phi, "e0": e0, "max_drift": max_drift,
        "sum_kinetic": sums[0], "sum_phi": sums[1], "sum_phi2": sums[2], "step": step,
        "status": status, "hist": hist,
    }


# ============================================================
# MD WITH THERMOSTATS AND TRACKED WAVES (dynamics/run.py:259-537,
# integrators.py:99-176, observables.py:56-125)
# ============================================================


@dataclasses.dataclass
class MDConfig:
    """The MDSettings fields the step reads (dynamics/state.py:57-88)."""

    dt: float
    scheme: str = "verlet"
    andersen_p: float = 0.01
    andersen_mode: str = "count"
    gamma: float = 0.0
    wind_every: int = 1
    record_steps: Tuple[int, ...] = ()
    record_rho_every: int = 0
    max_rho_records: int = 0
    phase_floor: float = 1e-16
    drift_tol: float = math.inf


def tracked_rho(r: np.ndarray, kvec: np.ndarray) -> np.ndarray:
    """rho_k = sum_j exp(i k . r_j) as (w, 2) by the trig sum over particles (observables.py:56-78)."""
    out = np.zeros((kvec.shape[0], 2))
    for i in range(kvec.shape[0]):
        th = r @ kvec[i]
        out[i, 0] = np.cos(th).sum()
        out[i, 1] = np.sin(th).sum()
    return out


def com_rotate(rho: np.ndarray, kvec: np.ndarray, r_cm: np.ndarray) -> np.ndarray:
    """rho_k exp(-i k . R_cm) (observables.py:89-98)."""
    out = np.empty_like(rho)
    for i in range(rho.shape[0]):
        ph = float(kvec[i] @ r_cm)
        c, s_ = math.cos(ph), math.sin(ph)
        out[i, 0] = rho[i, 0] * c + rho[i, 1] * s_
        out[i, 1] = rho[i, 1] * c - rho[i, 0] * s_
    return out


def md_lane(vg: ValueGrad, basis: np.ndarray, r0: np.ndarray, v0: np.ndarray, temperature: float,
            key, kvec: np.ndarray, cfg: MDConfig) -> Dict[str, object]:
    """One lane at step 0 (init_state, run.py:300-352): forces, rho_prev, empty buffers."""
    inv = np.linalg.inv(basis)
    phi, g = vg(r0 @ inv)
    n, d = r0.shape
    w = kvec.shape[0]
    rho = tracked_rho(r0, kvec)
    if cfg.scheme != "verlet":
        rho = com_rotate(rho, kvec, r0.mean(axis=0))
    nr, q = len(cfg.record_steps), cfg.max_rho_records
    return {
        "r": r0.copy(), "v": v0.copy(), "f": -g @ inv.T, "phi": float(phi), "temperature": float(temperature),
        "key": (int(key[0]), int(key[1])), "step": 0, "time": 0.0, "reset_acc": 0.0, "rho_prev": rho,
        "wind_abs": np.zeros(w), "wind_net": np.zeros(w), "n_valid": np.zeros(w, np.int64),
        "e0": float(phi) + 0.5 * float(np.sum(v0 * v0)), "max_drift": 0.0, "sum_kinetic": 0.0,
        "sum_phi": 0.0, "sum_phi2": 0.0, "sum_s": np.zeros(w), "n_sum": 0, "status": STATUS_RUNNING,
        "snap_step": np.full(nr, -1, np.int64), "snap_time": np.zeros(nr), "snap_wind_abs": np.zeros((nr, w)),
        "snap_wind_net": np.zeros((nr, w)), "snap_rho": np.zeros((nr, w, 2)), "snap_r": np.zeros((nr, n, d)),
        "snap_phi": np.zeros(nr), "snap_kinetic": np.zeros(nr), "snap_count": 0,
        "trace_rho": np.zeros((q, w, 2)), "trace_time": np.zeros(q), "trace_count": 0,
    }


def md_run(vg: ValueGrad, basis: np.ndarray, lane: Dict[str, object], kvec: np.ndarray, nc: int,
           cfg: MDConfig, n_steps: int) -> Dict[str, object]:
    """n_steps steps of one lane (run.py:415-537); a lane that leaves STATUS_RUNNING stops.

    Per step: the scheme (verlet; andersen = verlet then a reset from
    (key', k_pick, k_draw, k_coin) = split(key, 4); baoab = B A O A B with
    (key', k_xi) = split(key, 2)), rho of the tracked waves (the first nc
    every step, the rest on steps with step % wind_every == 0, else their
    previous value), COM correction of the fresh rows unless verlet, winding
    where S = |rho|^2 / N exceeds phase_floor at both ends, the sums (sum_s
    over every row), the drift guard (verlet), then the snapshot (step in
    record_steps) and the trace row (step % record_rho_every == 0).
    """
    st = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in lane.items()}
    inv = np.linalg.inv(basis)
    n, d = st["r"].shape
    w = kvec.shape[0]
    dt, half = cfg.dt, 0.5 * cfg.dt
    for _ in range(n_steps):
        if st["status"] != STATUS_RUNNING:
            break
        key, T = st["key"], st["temperature"]
        r, v, f = st["r"], st["v"], st["f"]
        if cfg.scheme == "baoab":
            key, k_xi = rng_split(key, 2)
            c = math.exp(-cfg.gamma * dt)
            var = -math.expm1(-2.0 * cfg.gamma * dt)
            v = v + half * f
            r = r + half * v
            v = c * v + math.sqrt(var * T) * rng_normals(k_xi, n * d).reshape(n, d)
            r = r + half * v
            phi, g = vg(r @ inv)
            f = -g @ inv.T
            v = v + half * f
        else:
            vh = v + half * f
            r = r + dt * vh
            phi, g = vg(r @ inv)
            f = -g @ inv.T
            v = vh + half * f
            if cfg.scheme == "andersen":
                key, k_pick, k_draw, k_coin = rng_split(key, 4)
                if cfg.andersen_mode == "count":
                    st["reset_acc"] += cfg.andersen_p
                    fire = st["reset_acc"] >= 1.0
                    if fire:
                        st["reset_acc"] -= 1.0
                else:
                    fire = rng_uniform(k_coin) < cfg.andersen_p
                if fire:
                    v = v.copy()
                    v[rng_randint(k_pick, n)] = math.sqrt(T) * rng_normals(k_draw, d)
        step = st["step"] + 1
        wind = step % cfg.wind_every == 0
        fresh = tracked_rho(r, kvec)
        if cfg.scheme != "verlet":
            fresh = com_rotate(fresh, kvec, r.mean(axis=0))
        rho_prev = st["rho_prev"]
        for i in range(w):
            if not (i < nc or wind):
                continue
            o_re, o_im = rho_prev[i]
            n_re, n_im = fresh[i]
            delta = math.atan2(n_im * o_re - n_re * o_im, n_re * o_re + n_im * o_im)
            if (o_re * o_re + o_im * o_im) / n > cfg.phase_floor and (n_re * n_re + n_im * n_im) / n > cfg.phase_floor:
                st["wind_abs"][i] += abs(delta)
                st["wind_net"][i] += delta
                st["n_valid"][i] += 1
            rho_prev[i] = fresh[i]
        ke = 0.5 * float(np.sum(v * v))
        drift = abs(float(phi) + ke - st["e0"])
        if cfg.scheme == "verlet" and math.isfinite(cfg.drift_tol):
            if not drift <= cfg.drift_tol * 0.5 * d * (n - 1) * T:
                st["status"] = STATUS_DRIFT
        st.update(r=r, v=v, f=f, phi=float(phi), key=key, step=step, time=st["time"] + dt)
        st["max_drift"] = _nanmax(st["max_drift"], drift)
        st["sum_kinetic"] += ke
        st["sum_phi"] += float(phi)
        st["sum_phi2"] += float(phi) * float(phi)
        st["sum_s"] = st["sum_s"] + (rho_prev[:, 0] ** 2 + rho_prev[:, 1] ** 2) / n
        st["n_sum"] += 1
        if step in cfg.record_steps:
            j = cfg.record_steps.index(step)
            st["snap_step"][j] = step
            st["snap_time"][j] = st["time"]
            st["snap_wind_abs"][j] = st["wind_abs"]
            st["snap_wind_net"][j] = st["wind_net"]
            st["snap_rho"][j] = rho_prev
            st["snap_r"][j] = r
            st["snap_phi"][j] = float(phi)
            st["snap_kinetic"][j] = ke
            st["snap_count"] += 1
        every, cap = cfg.record_rho_every, cfg.max_rho_records
        if every > 0 and cap > 0 and step % every == 0 and st["trace_count"] < cap:
            st["trace_rho"][st["trace_count"]] = rho_prev
            st["trace_time"][st["trace_count"]] = st["time"]
            st["trace_count"] += 1
    return st


# ============================================================
# ANALYSIS (tests/oracle.py:158-209, analysis/structure.py:111-201)
# ============================================================


def structure_factor(s: np.ndarray, n_vectors: np.ndarray) -> np.ndarray:
    """S(n) = |rho(n)|^2 / N by the trig double loop (vectorised over particles)."""
    out = np.zeros(n_vectors.shape[0])
    for kk in range(n_vectors.shape[0]):
        ph = TWO_PI * (s @ np.asarray(n_vectors[kk], dtype=np.float64))
        out[kk] = (np.cos(ph).sum() ** 2 + np.sin(ph).sum() ** 2) / s.shape[0]
    return out


def min_image_distances(s: np.ndarray, basis: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """(pair distances i < j, nearest-neighbour distance per particle) over 3^d images."""
    n, dim = s.shape
    shifts = np.array(list(itertools.product((-1.0, 0.0, 1.0), repeat=dim)))
    ds = s[None, :, :] - s[:, None, :]
    ds = ds - np.round(ds)
    best = np.full((n, n), np.inf)
    for sh in shifts:
        r = (ds + sh) @ basis
        r2 = np.sum(r * r, axis=-1)
        if not np.any(sh):
            np.fill_diagonal(r2, np.inf)
        best = np.minimum(best, r2)
    dist = np.sqrt(best)
    iu = np.triu_indices(n, 1)
    return dist[iu], dist.min(axis=1)


def isotropic_kset(basis: np.ndarray, k_dir_max: float, k_max: float) -> Tuple[np.ndarray, np.ndarray]:
    """Integer vectors (M, d) and |k|^2 of the isotropic S(k) set (structure.py:111-144)."""
    if basis.shape[0] == 1:
        b = TWO_PI / basis[0, 0]
        n_prim, k_prim = np.ones((1, 1), dtype=np.int64), np.array([[b]])
    else:
        n_all, _ = halfspace_kvectors(basis, 0.0, k_dir_max)
        recip = TWO_PI * np.linalg.inv(basis).T
        prim = np.gcd.reduce(np.abs(n_all), axis=1) == 1
        n_prim = n_all[prim]
        k_prim = np.array([[sum(float(n[i]) * recip[i, j] for i in range(len(n))) for j in range(basis.shape[0])]
                           for n in n_prim])
    rows, ksq = [], []
    for n_row, k_row in zip(n_prim, k_prim):
        m = 1
        while True:
            kk = m * k_row
            q = modulus2(kk)
            if q > k_max * k_max:
                break
            rows.append(m * n_row)
            ksq.append(q)
            m += 1
    return np.array(rows, dtype=np.int64), np.array(ksq)


# ============================================================
# COUNTER-BASED DRAWS (the thermostats' random numbers)
# ============================================================
#
# Philox-4x32-10 (Salmon, Moraes, Dror and Shaw, SC 2011) in plain Python
# integers, and the split / normal / uniform / integer draws built on it.
# Pinned by the published known-answer vectors of Philox-4x32-10 and by the
# moments and Kolmogorov-Smirnov statistics of the draws.

_PH_M = (0xD2511F53, 0xCD9E8D57)
_PH_W = (0x9E3779B9, 0xBB67AE85)
TAG_SPLIT, TAG_NORMAL, TAG_UNIF = 0x53504C54, 0x4E524D4C, 0x554E4946


def philox4x32_10(ctr: Tuple[int, int, int, int], key: Tuple[int, int]) -> Tuple[int, int, int, int]:
    c0, c1, c2, c3 = (int(x) & 0xFFFFFFFF for x in ctr)
    k0, k1 = (int(x) & 0xFFFFFFFF for x in key)
    for _ in range(10):
        p0 = _PH_M[0] * c0
        p1 = _PH_M[1] * c2
        c0, c1, c2, c3 = (p1 >> 32) ^ c1 ^ k0, p1 & 0xFFFFFFFF, (p0 >> 32) ^ c3 ^ k1, p0 & 0xFFFFFFFF
        k0, k1 = (k0 + _PH_W[0]) & 0xFFFFFFFF, (k1 + _PH_W[1]) & 0xFFFFFFFF
    return c0, c1, c2, c3


def rng_split(key, num: int):
    words: List[int] = []
    i = 0
    while len(words) < 2 * num:
        words += list(philox4x32_10((i, 0, 0, TAG_SPLIT), key))
        i += 1
    return [(words[2 * j], words[2 * j + 1]) for j in range(num)]


def _u53(hi: int, lo: int) -> float:
    return ((hi >> 5) * 67108864 + (lo >> 6)) / 9007199254740992.0


def rng_normals(key, count: int) -> np.ndarray:
    out = []
    for i in range((count + 1) // 2):
        w = philox4x32_10((i, 0, 0, TAG_NORMAL), key)
        u1, u2 = _u53(w[0], w[1]), _u53(w[2], w[3])
        rho = math.sqrt(-2.0 * math.log1p(-u1))
        out += [rho * math.cos(2.0 * math.pi * u2), rho * math.sin(2.0 * math.pi * u2)]
    return np.array(out[:count])


def rng_uniform(key) -> float:
    w = philox4x32_10((0, 0, 0, TAG_UNIF), key)
    return _u53(w[0], w[1])


def rng_randint(key, n: int) -> int:
    w = philox4x32_10((0, 0, 0, TAG_UNIF), key)
    return min(int(math.floor(_u53(w[2], w[3]) * n)), n - 1)
