"""The HIP path against the outcomes the reference itself published.

Same inputs and bounds as tests/test_oracle_reference_outcomes.py (CPU), here
through the C ABI on the GPU: default_rng(0).random((B, N, d)) on the
reference's benchmark shapes (benchmarks/bench_energy.py:74-92,
benchmarks/bench_threshold.py:46-63), compared with
benchmarks/results_{cpu,gpu}.json and results_threshold_gpu.json as
extracted into tests/golden/benchmark_outcomes.json.

The L-BFGS totals are also held to the C oracle run with power-table phases
(oracle/cco_oracle.c, phases=1) on the same inputs: the HIP kernels form phases
in that rounding class, and below the round-off floor the evaluation count
follows the rounding of the energy (at chi = 0.45 the reference's trig form
stops about 10% later than either).  Totals over B lanes of per-lane counts
that differ at the floor: within 1% at B=1000 (measured at most 0.41%) and 3%
at B=64 (sqrt(1000/64) times the spread).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_oracle
from oracle import oracle as O
from systems import benchmark_case, check_lbfgs_outcome, energy_extended, outcome
from xla_cco.box import cubic_box
from xla_cco.energy import make_stealthy_energy
from xla_cco.minimize import relax
from xla_cco.minimize.descent import DescentSettings
from xla_cco.minimize.lbfgs import LBFGSSettings

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def _energy(d: int, n: int, chi: float, batch: int):
    nv, w, vol, basis, s0 = benchmark_case(d, n, chi, batch)
    box = cubic_box(n, d)
    from xla_cco.kvectors import enumerate_kvectors, k_outer_from_ka, ka_from_chi

    ks = enumerate_kvectors(box.basis, 0.0, k_outer_from_ka(ka_from_chi(chi, 0.05, d), 0.05, d))
    assert np.array_equal(ks.n, nv)
    return make_stealthy_energy(box, ks), s0


@pytest.mark.parametrize("chi", [0.05, 0.45])
def test_energy_mean_and_grad_max_d3(gpu_ctx, chi: float) -> None:
    """4000 lanes, d=3, N=100: mean Phi and max |grad| to 1e-12 of the published values."""
    e, s = _energy(3, 100, chi, 4000)
    phi, g = e.value_and_grad_batch(s)
    ref = outcome("value_and_grad", 3, 100, chi, 4000)
    print(f"chi={chi}: mean {phi.mean()!r} (ref {ref['energy_mean']!r}), "
          f"max|g| {np.abs(g).max()!r} (ref {ref['grad_absmax']!r})", flush=True)
    assert abs(phi.mean() / ref["energy_mean"] - 1.0) <= 1e-12
    assert abs(np.abs(g).max() / ref["grad_absmax"] - 1.0) <= 1e-12


def test_energy_d2_n10000(gpu_ctx) -> None:
    """One lane, d=2, N=10^4, chi=0.45: Phi and max |grad| to 1e-12."""
    e, s = _energy(2, 10_000, 0.45, 1)
    phi, g = e.value_and_grad(s[0])
    ref = outcome("value_and_grad", 2, 10_000, 0.45, 1)
    print(f"Phi {phi!r} (ref {ref['energy_mean']!r}), max|g| {np.abs(g).max()!r}", flush=True)
    assert abs(phi / ref["energy_mean"] - 1.0) <= 1e-12
    assert abs(np.abs(g).max() / ref["grad_absmax"] - 1.0) <= 1e-12


@pytest.mark.parametrize("chi,batch", [(0.05, 64), (0.45, 64), (0.05, 1000), (0.45, 1000)])
def test_descent_400_energy_mean(gpu_ctx, chi: float, batch: int) -> None:
    """relax with DescentSettings(max_evals=400): mean energy to 1e-11 relative."""
    e, s0 = _energy(3, 100, chi, batch)
    out = relax(e, s0, DescentSettings(max_evals=400))
    ref = outcome("descent_400", 3, 100, chi, batch)
    print(f"chi={chi} B={batch}: mean {out.energy.mean()!r} (ref {ref['energy_mean']!r})", flush=True)
    assert int(np.sum(out.n_evals)) == 400 * batch
    assert abs(out.energy.mean() / ref["energy_mean"] - 1.0) <= 1e-11


@pytest.mark.parametrize(
    "chi,batch",
    [(0.05, 64), (0.45, 64), (0.05, 1000), (0.1, 1000), (0.2, 1000), (0.3, 1000), (0.45, 1000)],
)
def test_lbfgs_eval_totals(gpu_ctx, chi: float, batch: int) -> None:
    """LBFGSSettings(max_evals_per_pass=200): total n_evals and mean energy."""
    e, s0 = _energy(3, 100, chi, batch)
    out = relax(e, s0, LBFGSSettings(max_evals_per_pass=200))
    print(check_lbfgs_outcome(chi, batch, out.n_evals, out.energy), flush=True)
    nv, w, vol, _, _ = benchmark_case(3, 100, chi, 1)
    _, _, ne_tab, _, _ = c_oracle.relax(s0, nv, w, vol, O.LBFGSSettings(max_evals_per_pass=200),
                                        threads=ORACLE_THREADS, phases=1)
    rel = int(out.n_evals.sum()) / int(ne_tab.sum()) - 1.0
    print(f"power-table oracle {int(ne_tab.sum())}: HIP {rel:+.3%}", flush=True)
    assert abs(rel) <= (0.01 if batch >= 1000 else 0.03)
    if chi == 0.45:
        # the floor: reported energies agree with an 80-bit evaluation of the
        # returned configurations, and it lies no higher than the reference's
        nv, w, vol, _, _ = benchmark_case(3, 100, chi, 1)
        lanes = slice(0, 64)
        true = energy_extended(out.s[lanes], nv, w, vol)
        err = np.median(np.abs(out.energy[lanes] - true) / true)
        ref = outcome("relax_lbfgs", 3, 100, chi, batch)
        print(f"floor: mean Phi {true.mean():.3e} (80-bit), median rel. error of the "
              f"reported energies {err:.3f}; reference mean {ref['energy_mean']:.3e}", flush=True)
        assert err <= 0.1
        assert true.mean() <= ref["energy_mean"]
