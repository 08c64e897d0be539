"""Small test systems shared by the CPU and GPU tests (tests/systems.py of the reference)."""

from __future__ import annotations

from typing import Tuple

import numpy as np

from xla_cco.box import cubic_box
from xla_cco.kvectors import enumerate_kvectors, k_outer_from_ka, ka_from_chi


def small_kset(n: int, d: int, k_outer: float) -> Tuple[np.ndarray, np.ndarray, float, np.ndarray]:
    """(n_vectors int64 (M, d), weights (M,), volume, basis) of a cubic box."""
    box = cubic_box(n, d)
    ks = enumerate_kvectors(box.basis, 0.0, k_outer)
    return ks.n, ks.weights, box.volume, box.basis


def chi_kset(n: int, d: int, chi: float, phi: float = 0.05):
    """k-set at constraint ratio chi (Ka from chi at packing phi)."""
    k = k_outer_from_ka(ka_from_chi(chi, phi, d), phi, d)
    return small_kset(n, d, k)


def square_lattice(side: int) -> np.ndarray:
    g = np.arange(side) / side
    return np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1).reshape(-1, 2)


# ---------------- the reference's published benchmark outcomes ----------------

_OUTCOMES = None


def benchmark_outcomes() -> dict:
    """tests/golden/benchmark_outcomes.json (made by make_benchmark_fixture.py)."""
    global _OUTCOMES
    if _OUTCOMES is None:
        import json
        from pathlib import Path

        path = Path(__file__).resolve().parent / "golden" / "benchmark_outcomes.json"
        _OUTCOMES = json.loads(path.read_text())
    return _OUTCOMES


def outcome(kind: str, d: int, n: int, chi: float, batch: int, path: str = "trig") -> dict:
    """The one published row for (kind, d, n, chi, batch, path)."""
    rows = [
        r for r in benchmark_outcomes()["rows"]
        if r["kind"] == kind and r["d"] == d and r["n"] == n and r["batch"] == batch
        and r["path"] == path and abs(r["chi"] - chi) < 1e-12
    ]
    assert rows, f"no published row for {(kind, d, n, chi, batch, path)}"
    return rows[0]


def benchmark_case(d: int, n: int, chi: float, batch: int):
    """Inputs of the reference's benchmark rows (benchmarks/bench_energy.py:74-92).

    Returns (n_vectors, weights, volume, basis, s0) with s0 =
    default_rng(0).random((batch, n, d)) on a cubic unit-density box, k-set at
    packing fraction 0.05.
    """
    nv, w, vol, basis = chi_kset(n, d, chi, phi=0.05)
    s0 = np.random.default_rng(0).random((batch, n, d))
    return nv, w, vol, basis, s0


# Relative bounds on the total L-BFGS evaluation count against the reference's trig
# row.  Below the round-off floor the line search runs on energies that are mostly
# rounding noise (about 90% of the evaluations at chi=0.05), so totals follow the
# rounding of the energy: the reference's own trig and table paths differ by
# 0.01-0.3% at chi <= 0.3 and by 2.4% (64 lanes) and 6.2% (1000 lanes) at chi = 0.45.
# At chi = 0.45 the count falls with the floor the relax reaches: the reference's
# trig path ends at mean Phi 1.09e-27 after 557k evaluations, its table path at
# 5.8e-28 after 522k; a more accurate energy ends lower still and sooner, hence the
# wider bound there (checked together with the floor itself, see floor_error).
LBFGS_EVAL_RTOL = {(1000, 0.05): 0.015, (1000, 0.1): 0.015, (1000, 0.2): 0.015,
                   (1000, 0.3): 0.015, (1000, 0.45): 0.12, (64, 0.05): 0.015, (64, 0.45): 0.12}


def energy_extended(s: np.ndarray, nv: np.ndarray, w: np.ndarray, volume: float) -> np.ndarray:
    """Phi of (B, N, d) configurations in 80-bit long double, phases reduced mod 1.

    A check of energies at the round-off floor (Phi ~ 1e-28), where float64
    evaluations of rho carry relative errors of order one.
    """
    n = nv.astype(np.longdouble)
    two_pi = np.longdouble("6.28318530717958647692528676655900577")
    out = np.empty(s.shape[0])
    for i in range(s.shape[0]):
        dot = s[i].astype(np.longdouble) @ n.T
        th = two_pi * (dot - np.rint(dot))
        re = np.cos(th).sum(axis=0)
        im = np.sin(th).sum(axis=0)
        out[i] = float((w * (re * re + im * im)).sum() / volume)
    return out


def check_lbfgs_outcome(chi: float, batch: int, n_evals: np.ndarray, energy: np.ndarray) -> str:
    """Assert the published L-BFGS outcome at (chi, batch) within LBFGS_EVAL_RTOL."""
    ref = outcome("relax_lbfgs", 3, 100, chi, batch)
    total = int(np.sum(n_evals))
    rel = total / ref["n_evals"] - 1.0
    msg = f"chi={chi} B={batch}: total n_evals {total} vs reference {ref['n_evals']} ({rel:+.3%})"
    assert abs(rel) <= LBFGS_EVAL_RTOL[(batch, chi)], msg
    if "energy_mean" in ref:
        # round-off level energies: the same order of magnitude as the reference's
        ratio = float(np.mean(energy)) / ref["energy_mean"]
        assert 0.2 < ratio < 5.0, f"{msg}; mean energy {np.mean(energy):.3e} vs {ref['energy_mean']:.3e}"
    return msg
