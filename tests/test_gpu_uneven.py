"""Uneven batches: configurations of different N, d and k-set relaxed in one session.

Each lane's arithmetic is the single-configuration path's (the same unit,
split and reduction order when the lane's rho units fit one workgroup), so the
uneven batch reproduces per-configuration relax() results bit for bit there,
and to the oracle tolerances otherwise.
"""

from __future__ import annotations

import numpy as np
import pytest

from xla_cco import kvectors as K
from xla_cco.box import cubic_box
from xla_cco.energy import make_stealthy_energy
from xla_cco.minimize import DescentSettings, LBFGSSettings, relax, relax_uneven

pytestmark = pytest.mark.gpu


def _energy(n: int, d: int, chi: float):
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d))
    return make_stealthy_energy(box, ks)


def _batch(specs, seed: int):
    rng = np.random.default_rng(seed)
    cache = {}
    es, s0 = [], []
    for n, d, chi in specs:
        key = (n, d, chi)
        if key not in cache:
            cache[key] = _energy(n, d, chi)
        es.append(cache[key])
        s0.append(rng.uniform(size=(n, d)))
    return es, s0


SMALL = [(37, 2, 0.2), (64, 3, 0.1), (100, 2, 0.45), (64, 3, 0.1), (150, 3, 0.2), (37, 2, 0.2),
         (200, 2, 0.3), (90, 3, 0.3)]


@pytest.mark.parametrize("method", ["lbfgs", "descent"])
def test_uneven_equals_single_configuration_bitwise(gpu_ctx, method: str) -> None:
    es, s0 = _batch(SMALL, 1)
    st = (LBFGSSettings(energy_tol=1e-20) if method == "lbfgs"
          else DescentSettings(max_evals=400, step_init=3e-2))
    out = relax_uneven(es, s0, st)
    assert len(out) == len(s0)
    for e, s, r in zip(es, s0, out):
        one = relax(e, s[None], st)
        assert np.array_equal(r.s, one.s[0])
        assert r.energy == one.energy[0]
        assert int(r.n_evals) == int(one.n_evals[0])
        assert int(r.status) == int(one.status[0])
    if method == "lbfgs":
        assert all(float(r.energy) < 1e-20 for r in out)


def test_uneven_large_lanes_match_to_tolerance(gpu_ctx) -> None:
    """Lanes with more rho units than one workgroup (d = 3, N = 1200, chi = 0.45)
    sum their particle pairs in another order than the single path."""
    specs = [(1200, 3, 0.45), (300, 2, 0.1), (1200, 3, 0.45)]
    es, s0 = _batch(specs, 2)
    st = LBFGSSettings(energy_tol=1e-20, rescales=(1.0,), max_evals_per_pass=40, record_every=1,
                       max_records=12)
    out = relax_uneven(es, s0, st)
    for e, s, r in zip(es, s0, out):
        one = relax(e, s[None], st)
        k = int(min(r.trajectory.count, one.trajectory.count[0]))
        assert k >= 10
        assert np.allclose(r.trajectory.energy[:k], one.trajectory.energy[0, :k], rtol=1e-8)
        assert r.trajectory.s.shape == (12, s.shape[0], s.shape[1])


def test_uneven_trajectories_and_order(gpu_ctx) -> None:
    es, s0 = _batch(SMALL[:4], 3)
    st = LBFGSSettings(energy_tol=1e-20, record_every=2, max_records=6)
    out = relax_uneven(es, s0, st)
    for s, r in zip(s0, out):
        assert r.s.shape == s.shape
        tr = r.trajectory
        assert tr is not None and tr.s.shape == (6, s.shape[0], s.shape[1])
        assert np.array_equal(tr.s[0], s - np.floor(s))  # the initial row is the wrapped start


def test_uneven_rejects_bad_input(gpu_ctx) -> None:
    es, s0 = _batch(SMALL[:2], 4)
    with pytest.raises(ValueError):
        relax_uneven(es, s0[:1], LBFGSSettings())
    with pytest.raises(ValueError):
        relax_uneven(es, [s0[0], s0[0]], LBFGSSettings())  # d = 2 positions for a d = 3 energy
    with pytest.raises(ValueError):
        relax_uneven(es, s0, LBFGSSettings(precision="half"))
    with pytest.raises(TypeError):
        relax_uneven(es, [s0[0].astype(np.float32), s0[1]], LBFGSSettings())


@pytest.mark.parametrize("split_history", [0, 1])
def test_uneven_split_precision_success_and_switch(gpu_ctx, split_history: int) -> None:
    """Split precision in an uneven batch: lanes whose k-set and N fit the
    one-workgroup-per-lane kernel start split and turn FP64 below the switch
    (every lane of this batch fits); the success set equals the FP64 batch's,
    and the converged configurations are checked in FP64.  With split_history 1
    the split lanes keep FP32 (s, y) pairs at their per-lane history offsets."""
    from xla_cco import _native
    from xla_cco.minimize import UnevenSession

    specs = SMALL + [(1000, 2, 0.45), (300, 3, 0.2), (757, 2, 0.3)]
    es, s0 = _batch(specs, 9)
    tol = 9.99999968e-21
    r64 = relax_uneven(es, s0, LBFGSSettings(energy_tol=tol))
    st = LBFGSSettings(energy_tol=tol, precision="mixed")
    sess = UnevenSession(es, [np.shape(s) for s in s0], st)
    try:
        _native.check(_native.lib().cco_relax_set_option(sess.handle, 5, split_history))  # CCO_OPT_SPLIT_HISTORY
        sess.init(s0)
        on, split = sess.precision_state()
        assert on and split == len(specs)
        sess.run(10_000_000, 64)
        on, split = sess.precision_state()
        assert on and split == 0
        rmp = sess.result()
    finally:
        sess.close()
    ok64 = np.array([float(r.energy) < tol for r in r64])
    okmp = np.array([float(r.energy) < tol for r in rmp])
    assert np.array_equal(ok64, okmp) and okmp.all()
    for e, r in zip(es, rmp):
        assert e.energy_batch(r.s[None])[0] < 1e-19


def test_uneven_split_precision_lanes_that_do_not_fit_stay_fp64(gpu_ctx) -> None:
    """A lane whose column box is too wide for the lane kernel (d=3, N=4000)
    runs FP64 from the start; the small lanes beside it run split."""
    from xla_cco.minimize import UnevenSession

    specs = [(100, 2, 0.3), (4000, 3, 0.3), (200, 3, 0.2)]
    es, s0 = _batch(specs, 10)
    st = LBFGSSettings(energy_tol=1e-6, precision="mixed", precision_switch=0.0)
    sess = UnevenSession(es, [np.shape(s) for s in s0], st)
    try:
        sess.init(s0)
        assert sess.precision_state() == (True, 2)
        sess.run(10_000_000, 64)
        out = sess.result()
    finally:
        sess.close()
    # the FP64 lane is the FP64 batch's lane bit for bit
    r64 = relax_uneven(es, s0, LBFGSSettings(energy_tol=1e-6))
    assert np.array_equal(out[1].s, r64[1].s)
    assert int(out[1].n_evals) == int(r64[1].n_evals)
    assert all(float(r.energy) < 1e-6 for r in out)
