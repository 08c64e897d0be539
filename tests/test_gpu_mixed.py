"""Split-precision (FP16 hi/lo MFMA, FP32 accumulation) energy, forces and relax.

The precision row of the target design: the density and force contractions run
on the matrix cores while a lane's energy is above the switch, FP64 below.
These tests state the error of every result against the FP64 path and the
oracle, and check that a mixed relax ends with the FP64 path's success set.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from xla_cco import _native
from xla_cco import kvectors as K
from xla_cco.box import cubic_box
from xla_cco.energy import make_stealthy_energy
from xla_cco.minimize import LBFGSSettings, RelaxSession, relax

pytestmark = pytest.mark.gpu

# Error bounds of the split-precision evaluations (operands carry ~22
# significant bits, products and sums are FP32): relative energy error, force
# error relative to max |grad|, density error relative to sqrt(N).
MIXED_RTOL_PHI = 2e-5
MIXED_RTOL_GRAD = 2e-4
MIXED_RTOL_RHO = 2e-5


def _system(n: int, d: int, chi: float):
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d))
    return box, ks


def _pair(n: int, d: int, chi: float):
    box, ks = _system(n, d, chi)
    e64 = make_stealthy_energy(box, ks)
    emp = make_stealthy_energy(box, ks)
    assert emp.set_precision("mixed")
    return e64, emp, ks, box


@pytest.mark.parametrize(
    "n,d,chi",
    [(64, 1, 0.3), (256, 2, 0.25), (1000, 2, 0.45), (100, 3, 0.45), (700, 3, 0.3), (257, 3, 0.2)],
)
def test_mixed_energy_forces_density_errors(gpu_ctx, n: int, d: int, chi: float) -> None:
    e64, emp, ks, box = _pair(n, d, chi)
    s = np.random.default_rng(n + d).uniform(size=(3, n, d))
    p64, g64, re64, im64 = e64.value_grad_density_batch(s)
    pmp, gmp, remp, immp = emp.value_grad_density_batch(s)
    assert np.all(np.abs(pmp - p64) <= MIXED_RTOL_PHI * p64)
    for i in range(3):
        assert np.max(np.abs(gmp[i] - g64[i])) <= MIXED_RTOL_GRAD * np.max(np.abs(g64[i]))
    scale = np.sqrt(n)
    assert np.max(np.abs(remp - re64)) <= MIXED_RTOL_RHO * scale
    assert np.max(np.abs(immp - im64)) <= MIXED_RTOL_RHO * scale
    # and against the oracle directly (first configuration)
    vg = O.make_value_grad(ks.n, ks.weights, box.volume)
    phi_o, g_o = vg(s[0])
    assert abs(pmp[0] - phi_o) <= MIXED_RTOL_PHI * phi_o
    assert np.max(np.abs(gmp[0] - g_o)) <= MIXED_RTOL_GRAD * np.max(np.abs(g_o))


EOPT_MP_LANE = 2  # CCO_EOPT_MP_LANE


def _mp_lane(e, on: bool | None = None) -> int:
    """Set (on given) and read the energy's one-workgroup-per-lane option."""
    L = _native.lib()
    if on is not None:
        _native.check(L.cco_energy_set_option(e.handle(), EOPT_MP_LANE, 1 if on else 0))
    v = C.c_int64()
    _native.check(L.cco_energy_get_option(e.handle(), EOPT_MP_LANE, C.byref(v)))
    return int(v.value)


@pytest.mark.parametrize(
    "n,d,chi",
    [(64, 1, 0.3), (256, 2, 0.25), (1000, 2, 0.45), (1500, 2, 0.45), (77, 2, 0.45), (100, 3, 0.45),
     (257, 3, 0.2)],
)
def test_lane_kernel_errors_and_tiled_agreement(gpu_ctx, n: int, d: int, chi: float) -> None:
    """Small lanes evaluate in one workgroup each (rho, Phi, C and the forces in
    LDS); the tiled split-precision kernels stay selectable.  Both are held to
    the split-precision bounds against FP64, and the energy-only call returns
    the same Phi bits as the value-and-gradient call."""
    e64, emp, ks, box = _pair(n, d, chi)
    assert _mp_lane(emp) == 1
    s = np.random.default_rng(n * d + 5).uniform(size=(5, n, d))
    p64, g64, re64, im64 = e64.value_grad_density_batch(s)
    lane = emp.value_grad_density_batch(s)
    assert np.array_equal(emp.energy_batch(s), lane[0])
    assert _mp_lane(emp, False) == 0
    tiled = emp.value_grad_density_batch(s)
    scale = np.sqrt(n)
    for pmp, gmp, remp, immp in (lane, tiled):
        assert np.all(np.abs(pmp - p64) <= MIXED_RTOL_PHI * p64)
        for i in range(5):
            assert np.max(np.abs(gmp[i] - g64[i])) <= MIXED_RTOL_GRAD * np.max(np.abs(g64[i]))
        assert np.max(np.abs(remp - re64)) <= MIXED_RTOL_RHO * scale
        assert np.max(np.abs(immp - im64)) <= MIXED_RTOL_RHO * scale
    assert _mp_lane(emp, True) == 1
    again = emp.value_grad_density_batch(s)
    for a, b in zip(lane, again):
        assert np.array_equal(a, b)


EOPT_MP_LAYOUT = 3  # CCO_EOPT_MP_LAYOUT


@pytest.mark.parametrize("n,d,chi", [(700, 3, 0.3), (2000, 3, 0.2), (4000, 3, 0.45), (1500, 3, 0.1)])
def test_pair_layout_errors_and_column_layout(gpu_ctx, n: int, d: int, chi: float) -> None:
    """At d = 3 the tiled split-precision kernels run on the pair layout
    (columns (n1, n2), rows n3 + P3); the column layout stays selectable.  Both
    are held to the split-precision bounds against FP64, densities included, and
    the energy-only call returns the value-and-gradient call's Phi bits."""
    e64, emp, ks, box = _pair(n, d, chi)
    _mp_lane(emp, False)
    L = _native.lib()
    v = C.c_int64()
    _native.check(L.cco_energy_get_option(emp.handle(), EOPT_MP_LAYOUT, C.byref(v)))
    assert v.value == 1
    s = np.random.default_rng(n + 31).uniform(size=(4, n, d))
    p64, g64, re64, im64 = e64.value_grad_density_batch(s)
    pair = emp.value_grad_density_batch(s)
    assert np.array_equal(emp.energy_batch(s), pair[0])
    _native.check(L.cco_energy_set_option(emp.handle(), EOPT_MP_LAYOUT, 0))
    _native.check(L.cco_energy_get_option(emp.handle(), EOPT_MP_LAYOUT, C.byref(v)))
    assert v.value == 0
    col = emp.value_grad_density_batch(s)
    scale = np.sqrt(n)
    for pmp, gmp, remp, immp in (pair, col):
        assert np.all(np.abs(pmp - p64) <= MIXED_RTOL_PHI * p64)
        for i in range(4):
            assert np.max(np.abs(gmp[i] - g64[i])) <= MIXED_RTOL_GRAD * np.max(np.abs(g64[i]))
        assert np.max(np.abs(remp - re64)) <= MIXED_RTOL_RHO * scale
        assert np.max(np.abs(immp - im64)) <= MIXED_RTOL_RHO * scale


def test_mixed_error_is_absolute_near_ground_state(gpu_ctx) -> None:
    """The split-precision density error is absolute (|d rho_k| <~ 1e-6 sqrt(N) at
    any configuration), so close to a ground state the mixed path stops being
    accurate and lanes switch to FP64.  The forces stay within the bound that
    error implies, |d g_j| <= pref sum_k w_k |n_k| |d rho_k|: the per-lane
    power-of-two scale of C keeps the fp16 operands in range however small rho
    is (a wrong scale would overflow or flush to zero far beyond this bound)."""
    e64, emp, ks, box = _pair(64, 2, 0.2)
    s0 = np.random.default_rng(3).uniform(size=(1, 64, 2))
    res = relax(e64, s0, LBFGSSettings(energy_tol=1e-14))
    s = res.s
    p64, g64, re64, im64 = e64.value_grad_density_batch(s)
    pmp, gmp, remp, immp = emp.value_grad_density_batch(s)
    assert 0.0 < p64[0] < 1e-12
    m = ks.n.shape[0]
    drho = np.hypot(remp - re64, immp - im64)[0, :m]
    assert np.max(drho) <= 1e-6 * np.sqrt(64)
    w = ks.weights
    pref = 4.0 * np.pi / box.volume
    bound = pref * np.sum(w * np.abs(ks.n).max(axis=1) * drho) * np.sqrt(2.0)
    assert np.max(np.abs(gmp - g64)) <= 2.0 * bound + 1e-14
    rho = np.hypot(re64, im64)[0, :m]
    phi_bound = np.sum(w * (2.0 * rho * drho + drho**2)) / box.volume
    assert abs(pmp[0] - p64[0]) <= 2.0 * phi_bound + 1e-18


def test_mixed_relax_same_success_as_fp64(gpu_ctx) -> None:
    box, ks = _system(300, 2, 0.3)
    e = make_stealthy_energy(box, ks)
    s0 = np.random.default_rng(7).uniform(size=(12, 300, 2))
    tol = 9.99999968e-21
    r64 = relax(e, s0, LBFGSSettings(energy_tol=tol))
    rmp = relax(e, s0, LBFGSSettings(energy_tol=tol, precision="mixed"))
    assert np.array_equal(r64.energy < tol, rmp.energy < tol)
    assert np.all(rmp.energy < tol)
    # converged lanes were verified in FP64 (they switched below 1e-8)
    assert np.all(e.energy_batch(rmp.s) < 1e-19)


@pytest.mark.parametrize("n,d,chi,b", [(300, 2, 0.3, 12), (3000, 3, 0.2, 4)])
def test_mixed_rescue_turns_failing_lanes_fp64(gpu_ctx, n, d, chi, b) -> None:
    """With a switch energy no lane reaches in split precision (1e-30), every lane's
    line search fails or stagnates on split-precision energies near the floor; the lane
    then turns FP64, re-evaluates its point and goes on, so the relax still reaches
    E < tol on FP64's lanes (register two-loop at N d = 600, Gram form at 9000)."""
    box, ks = _system(n, d, chi)
    e = make_stealthy_energy(box, ks)
    s0 = np.random.default_rng(17).uniform(size=(b, n, d))
    tol = 9.99999968e-21
    r64 = relax(e, s0, LBFGSSettings(energy_tol=tol))
    sess = RelaxSession(e, b, n, d, LBFGSSettings(energy_tol=tol, precision="mixed", precision_switch=1e-30))
    sess.init(s0)
    assert sess.precision_state() == (True, b)
    sess.run(10_000_000, 64)
    out = sess.result()
    assert sess.precision_state() == (True, 0)  # every lane was rescued into FP64
    sess.close()
    print(f"N={n} d={d}: evaluations {int(r64.n_evals.sum())} FP64, {int(out.n_evals.sum())} split + rescue",
          flush=True)
    assert np.array_equal(r64.energy < tol, out.energy < tol)
    assert np.all(out.energy < tol)
    assert np.all(e.energy_batch(out.s) < 1e-19)


OPT_ADVANCE_VARIANT = 1  # CCO_OPT_ADVANCE_VARIANT
OPT_SPLIT_HISTORY = 5  # CCO_OPT_SPLIT_HISTORY


@pytest.mark.parametrize("n,d,chi,b", [(300, 2, 0.3, 12), (1001, 3, 0.2, 6), (1500, 3, 0.2, 4)])
def test_mixed_fp32_history(gpu_ctx, n, d, chi, b) -> None:
    """Split-precision lanes on the register two-loop with CCO_OPT_SPLIT_HISTORY = 1 keep
    FP32 (s, y) pairs and widen them to FP64 when they turn FP64: every lane converges as
    in FP64, with evaluation totals within 10% of the same relax on FP64 pairs (the
    default).  The three shapes take
    the 512-thread paired loads (N d = 600), the one-pair-per-element loads (3003) and the
    eight-element loop (4500)."""
    box, ks = _system(n, d, chi)
    e = make_stealthy_energy(box, ks)
    s0 = np.random.default_rng(23).uniform(size=(b, n, d))
    tol = 9.99999968e-21
    r64 = relax(e, s0, LBFGSSettings(energy_tol=tol))
    L = _native.lib()
    outs = {}
    for hist in (1, 0):
        sess = RelaxSession(e, b, n, d, LBFGSSettings(energy_tol=tol, precision="mixed",
                                                      split_history="fp32" if hist else "fp64"))
        v = C.c_int64(-1)
        _native.check(L.cco_relax_get_option(sess.handle, OPT_SPLIT_HISTORY, C.byref(v)))
        assert v.value == hist
        sess.init(s0)
        _native.check(L.cco_relax_get_option(sess.handle, OPT_ADVANCE_VARIANT, C.byref(v)))
        assert v.value == 2  # the register two-loop
        assert sess.precision_state() == (True, b)
        sess.run(10_000_000, 64)
        outs[hist] = sess.result()
        assert sess.precision_state() == (True, 0)  # every lane turned FP64
        assert L.cco_relax_set_option(sess.handle, OPT_SPLIT_HISTORY, 2) != 0
        sess.close()
    e32, e64 = int(outs[1].n_evals.sum()), int(outs[0].n_evals.sum())
    print(f"N={n} d={d}: evaluations {int(r64.n_evals.sum())} FP64, {e64} split + FP64 pairs, "
          f"{e32} split + FP32 pairs", flush=True)
    for out in outs.values():
        assert np.array_equal(r64.energy < tol, out.energy < tol)
        assert np.all(out.energy < tol)
        assert np.all(e.energy_batch(out.s) < 1e-19)
    assert abs(e32 / e64 - 1.0) <= 0.10


def test_mixed_fp32_history_shard_and_graph_invariant(gpu_ctx) -> None:
    """FP32 (s, y) pairs are per lane: 64 lanes split over 1 or 3 shards, with or without
    HIP-graph replays, give bit-identical positions, energies and evaluation counts."""
    n, d, b = 300, 2, 64
    box, ks = _system(n, d, 0.3)
    e = make_stealthy_energy(box, ks)
    s0 = np.random.default_rng(29).uniform(size=(b, n, d))
    L = _native.lib()
    outs = []
    for shards, graph in ((1, 0), (3, 0), (3, 8)):
        sess = RelaxSession(e, b, n, d, LBFGSSettings(energy_tol=9.99999968e-21, precision="mixed"))
        _native.check(L.cco_relax_set_option(sess.handle, OPT_SPLIT_HISTORY, 1))
        _native.check(L.cco_relax_set_option(sess.handle, 2, shards))
        _native.check(L.cco_relax_set_option(sess.handle, 3, graph))
        sess.init(s0)
        sess.run(10_000_000, 64)
        outs.append(sess.result())
        assert sess.precision_state() == (True, 0)
        sess.close()
    for out in outs[1:]:
        assert np.array_equal(out.s, outs[0].s)
        assert np.array_equal(out.energy, outs[0].energy)
        assert np.array_equal(out.n_evals, outs[0].n_evals)
    assert np.all(outs[0].energy < 9.99999968e-21)


def test_mixed_session_switches_every_lane(gpu_ctx) -> None:
    box, ks = _system(200, 3, 0.2)
    e = make_stealthy_energy(box, ks)
    s0 = np.random.default_rng(8).uniform(size=(4, 200, 3))
    sess = RelaxSession(e, 4, 200, 3, LBFGSSettings(energy_tol=1e-20, precision="mixed"))
    sess.init(s0)
    on, split = sess.precision_state()
    assert on and split == 4
    sess.run(10_000_000, 64)
    on, split = sess.precision_state()
    assert on and split == 0
    assert np.all(sess.result().energy < 1e-20)
    sess.close()
    # switch 0: lanes stay split for the whole run
    sess = RelaxSession(e, 4, 200, 3, LBFGSSettings(energy_tol=1e-6, precision="mixed", precision_switch=0.0))
    sess.init(s0)
    sess.run(10_000_000, 64)
    assert sess.precision_state() == (True, 4)
    sess.close()


def test_mixed_rejects_bad_settings(gpu_ctx) -> None:
    box, ks = _system(16, 2, 0.2)
    e = make_stealthy_energy(box, ks)
    with pytest.raises(ValueError):
        e.set_precision("fp32")
    with pytest.raises(ValueError):
        RelaxSession(e, 1, 16, 2, LBFGSSettings(precision="half"))
    with pytest.raises(ValueError):
        RelaxSession(e, 1, 16, 2, LBFGSSettings(precision="mixed", split_history="half"))
    sess = RelaxSession(e, 1, 16, 2, LBFGSSettings(precision="mixed"))
    v = C.c_int64(-1)
    _native.check(_native.lib().cco_relax_get_option(sess.handle, OPT_SPLIT_HISTORY, C.byref(v)))
    assert v.value == 0  # the default: FP64 pairs
    sess.close()


def test_mixed_md_energy_drift_vs_fp64(gpu_ctx) -> None:
    """NVE with split-precision forces from a relaxed state at T = 1e-4: the run
    stays stable and its energy drift is reported against the FP64 run."""
    import math

    from xla_cco.dynamics import MDSettings, Waves, init_state, prng_key, simulate

    box, ks = _system(400, 3, 0.3)
    e = make_stealthy_energy(box, ks)
    s0 = np.random.default_rng(21).uniform(size=(2, 400, 3))
    res = relax(e, s0, LBFGSSettings(energy_tol=1e-20))
    r0 = res.s @ box.basis
    v0 = np.random.default_rng(22).normal(scale=math.sqrt(1e-4), size=r0.shape)
    nv = np.asarray(ks.n, dtype=np.float64)
    waves = Waves(n_vectors=nv, kvec=nv @ (2.0 * np.pi * np.linalg.inv(box.basis).T),
                  constrained=np.ones(nv.shape[0], bool), n_constrained=nv.shape[0])
    keys = np.stack([prng_key(1), prng_key(2)])
    out = {}
    for prec in ("fp64", "mixed"):
        st = MDSettings(dt=0.01, remove_com=False, drift_tol=math.inf, precision=prec)
        lanes = init_state(e, waves, box.basis, r0, v0, 1e-4, keys, st)
        out[prec] = simulate(e, waves, box.basis, st, lanes, 400).state
    a, m = out["fp64"], out["mixed"]
    assert np.all(m.step == 400) and np.all(m.status == 0)
    kin = 0.5 * np.sum(v0**2, axis=(1, 2))
    # FP64 Verlet drift is O(dt^2); the split-precision forces add a bounded error
    assert np.all(a.max_drift < 1e-3 * kin)
    assert np.all(m.max_drift < 1e-2 * kin)
    # the trajectories stay close over 400 steps
    assert np.max(np.abs(m.r - a.r)) < 1e-3
