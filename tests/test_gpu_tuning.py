"""Autotuner and its knobs: every knob leaves the results bit-identical."""

from __future__ import annotations

import json

import numpy as np
import pytest

from xla_cco import _native
from xla_cco import kvectors as K
from xla_cco.box import cubic_box
from xla_cco.energy import make_stealthy_energy
from xla_cco.minimize import DescentSettings, LBFGSSettings, relax
from xla_cco.minimize import tuning

pytestmark = pytest.mark.gpu


def _energy(n: int, d: int, chi: float):
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d))
    return make_stealthy_energy(box, ks)


@pytest.mark.parametrize("n,d,chi", [(300, 2, 0.45), (400, 3, 0.3), (2500, 2, 0.2)])
def test_rho_chunk_is_bit_neutral(gpu_ctx, n: int, d: int, chi: float) -> None:
    e = _energy(n, d, chi)
    s = np.random.default_rng(n).uniform(size=(5, n, d))
    chunks = tuning._rho_chunks(e)
    assert chunks and chunks == sorted(chunks, reverse=True)
    outs = []
    for c in chunks:
        _native.check(_native.lib().cco_energy_set_option(e.handle(), tuning.EOPT_RHO_CHUNK, c))
        outs.append(e.value_grad_density_batch(s))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("method", ["lbfgs", "descent"])
def test_autotuned_relax_is_bit_identical(gpu_ctx, tmp_path, monkeypatch, method: str) -> None:
    monkeypatch.setenv(tuning.CACHE_ENV, str(tmp_path / "tuning.json"))
    tuning.forget()
    e = _energy(200, 2, 0.3)
    s0 = np.random.default_rng(5).uniform(size=(40, 200, 2))
    st = (LBFGSSettings(energy_tol=1e-20) if method == "lbfgs"
          else DescentSettings(grow=2.0, max_evals=3000, energy_tol=1e-20))
    plain = relax(e, s0, st)
    tuned = relax(e, s0, st, autotune=True)
    for f in ("s", "energy", "grad_norm", "n_evals", "n_iters", "status"):
        assert np.array_equal(getattr(plain, f), getattr(tuned, f)), f
    key = tuning.workload_key(e, 40, 200, 2, st)
    saved = json.loads((tmp_path / "tuning.json").read_text())
    assert key in saved
    choice = tuning.TuneChoice(**saved[key])
    print(f"{key}: {choice}", flush=True)
    assert choice.rho_chunk in (16, 32, 64) and 1 <= choice.shards <= 4 and choice.graph in (0, 8)
    # a second process-level lookup reads the file instead of tuning again
    tuning.forget()
    assert tuning.cached_choice(key) == choice


@pytest.mark.parametrize("n,d,chi", [(300, 2, 0.45), (2500, 2, 0.2), (3000, 3, 0.1)])
def test_rho_split_regroups_sums_only(gpu_ctx, n: int, d: int, chi: float) -> None:
    """Longer rho splits change the density and energy at rounding level only; a
    request no longer than the default runs the default (bitwise)."""
    e = _energy(n, d, chi)
    L = _native.lib()
    s = np.random.default_rng(n + 1).uniform(size=(4, n, d))
    base = e.value_grad_density_batch(s)
    with pytest.raises(ValueError):
        _native.check(L.cco_energy_set_option(e.handle(), tuning.EOPT_RHO_SPLIT, 100))
    _native.check(L.cco_energy_set_option(e.handle(), tuning.EOPT_RHO_SPLIT, tuning.SPLIT_UNIT))
    same = e.value_grad_density_batch(s)
    for a, b in zip(base, same):
        assert np.array_equal(a, b)
    cands = tuning.split_candidates(n)
    assert cands[0] == 0 and len(cands) >= 2
    for sp in cands[1:]:
        _native.check(L.cco_energy_set_option(e.handle(), tuning.EOPT_RHO_SPLIT, sp))
        phi, grad, re, im = e.value_grad_density_batch(s)
        assert np.allclose(phi, base[0], rtol=1e-13, atol=0)
        scale = np.abs(base[1]).max()
        assert np.abs(grad - base[1]).max() <= 1e-12 * scale
        assert np.abs(re - base[2]).max() <= 1e-12 * n and np.abs(im - base[3]).max() <= 1e-12 * n
    _native.check(L.cco_energy_set_option(e.handle(), tuning.EOPT_RHO_SPLIT, 0))


def test_rounding_level_tuning_same_outcomes(gpu_ctx, tmp_path, monkeypatch) -> None:
    """tune_exact=False may pick another rho split and advance variant: every lane
    still converges, evaluation totals stay within 5%, and the energy's options
    are handed back unchanged."""
    monkeypatch.setenv(tuning.CACHE_ENV, str(tmp_path / "tuning.json"))
    tuning.forget()
    e = _energy(200, 2, 0.3)
    s0 = np.random.default_rng(6).uniform(size=(40, 200, 2))
    st = LBFGSSettings(energy_tol=1e-20)
    before = tuning.energy_options(e)
    plain = relax(e, s0, st)
    tuned = relax(e, s0, st, autotune=True, tune_exact=False)
    assert tuning.energy_options(e) == before
    assert np.array_equal(plain.status, tuned.status)
    assert np.all(tuned.energy < 1e-20) and np.all(plain.energy < 1e-20)
    tot_p, tot_t = int(plain.n_evals.sum()), int(tuned.n_evals.sum())
    assert abs(tot_t - tot_p) <= 0.05 * tot_p, (tot_p, tot_t)
    key = tuning.workload_key(e, 40, 200, 2, st, exact=False)
    choice = tuning.cached_choice(key)
    print(f"{key}: {choice}; n_evals {tot_p} -> {tot_t}", flush=True)
    assert choice is not None and choice.advance in (1, 2, 3)
    assert choice.rho_split in tuning.split_candidates(200)
    assert tuning.cached_choice(tuning.workload_key(e, 40, 200, 2, st)) is None
    if choice.advance == 2 and choice.rho_split == 0:  # the defaults at this size
        assert np.array_equal(plain.s, tuned.s)
