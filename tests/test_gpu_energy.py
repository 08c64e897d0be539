"""HIP energy / forces / density vs the oracle (energy.py:186-225, 301-358).

Tolerances: energy rtol 1e-12 and forces rtol 1e-10 relative to the largest
component (tests/tolerances.py:18-20 of the reference), with an absolute
floor for energies that cancel to ~0.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from systems import chi_kset, small_kset, square_lattice
from xla_cco import kvectors as K
from xla_cco.box import cubic_box
from xla_cco.energy import StealthyEnergy, collective_density, make_stealthy_energy

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden" / "md_cxx_phi.npz"

RTOL_E = 1e-12
RTOL_F = 1e-10


def _energy(n: int, d: int, k_outer: float, pad_to: int = 64):
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, k_outer)
    return make_stealthy_energy(box, ks, pad_to=pad_to), ks, box


def _check(e: StealthyEnergy, ks, vol: float, s: np.ndarray) -> None:
    f, g, re, im = e.value_grad_density(s)
    fo, go, reo, imo = O.energy_grad(s, ks.n, ks.weights, vol)
    m = ks.n.shape[0]
    n = s.shape[0]
    scale = np.sqrt(n * m * max(fo, 1e-300))
    assert abs(f - fo) <= RTOL_E * fo + 4.0 * 2.2e-16 * scale, (f, fo)
    gmax = max(np.abs(go).max(), 1e-300)
    assert np.max(np.abs(g - go)) <= RTOL_F * gmax + 1e-13, np.max(np.abs(g - go)) / gmax
    assert np.allclose(re[:m], reo, rtol=0, atol=1e-11 * np.sqrt(n))
    assert np.allclose(im[:m], imo, rtol=0, atol=1e-11 * np.sqrt(n))
    # padding rows (n = 0): rho = N exactly
    assert np.all(re[m:] == n) and np.all(im[m:] == 0.0)


@pytest.mark.parametrize(
    "n,d,k_outer",
    [(1, 3, 9.0), (50, 1, 9.0), (16, 2, 4.0), (64, 2, 5.2), (100, 3, 2.6093), (37, 3, 4.4), (300, 2, 6.5)],
)
def test_value_grad_density_matches_oracle(gpu_ctx, n: int, d: int, k_outer: float) -> None:
    e, ks, box = _energy(n, d, k_outer)
    rng = np.random.default_rng(n * 10 + d)
    for trial in range(2):
        s = rng.uniform(size=(n, d))
        if trial == 1:
            s = s + rng.integers(-3, 4, size=s.shape)  # unwrapped iterates
        _check(e, ks, box.volume, s)


@pytest.mark.parametrize("n,d,k_inner,k_outer", [(40, 1, 3.0, 9.0), (200, 2, 2.0, 5.0), (150, 3, 1.5, 3.2),
                                                  (500, 3, 2.2, 4.4)])
def test_kset_with_inner_cutoff_matches_oracle(gpu_ctx, n: int, d: int, k_inner: float, k_outer: float) -> None:
    """A k-set k_inner < |k| < k_outer (enumerate_kvectors' inner cutoff): the rho runs that
    hold no k-vector (near n1 = 0 inside the inner sphere) write zero sums without work."""
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, k_inner, k_outer)
    assert ks.n.shape[0] > 0 and np.all(np.linalg.norm(ks.n @ (2 * np.pi * np.linalg.inv(box.basis).T), axis=1) > k_inner)
    e = make_stealthy_energy(box, ks)
    rng = np.random.default_rng(n + d)
    _check(e, ks, box.volume, rng.uniform(size=(n, d)))
    s2 = rng.uniform(size=(3, n, d))
    phi, g, re, im = e.value_grad_density_batch(s2)
    for b in range(3):
        fo, go, reo, imo = O.energy_grad(s2[b], ks.n, ks.weights, box.volume)
        assert abs(phi[b] - fo) <= RTOL_E * fo + 1e-12
        assert np.allclose(re[b, : ks.n.shape[0]], reo, rtol=0, atol=1e-11 * np.sqrt(n))


@pytest.mark.parametrize("d,n,chi", [(2, 1000, 0.45), (3, 400, 0.3), (2, 1500, 0.2), (1, 800, 0.45), (2, 3000, 0.1)])
def test_bench_shapes_match_oracle(gpu_ctx, d: int, n: int, chi: float) -> None:
    """Shapes with several rho splits, column tiles and row segments; d = 1 puts all
    M = 359 rows in one column (the longest phase recurrences); N = 3000 takes
    256-particle splits."""
    nv, w, vol, basis = chi_kset(n, d, chi)
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d))
    e = make_stealthy_energy(box, ks)
    s = np.random.default_rng(7).uniform(size=(n, d))
    _check(e, ks, vol, s)


def test_cpp_golden_energies(gpu_ctx) -> None:
    """C++ energies of the 15 stored MD configurations at relative 1e-12."""
    z = np.load(GOLDEN)
    basis = z["basis"]
    ks = K.enumerate_kvectors(basis, 0.0, float(z["K"]))
    from xla_cco.box import Box

    e = make_stealthy_energy(Box(basis=basis), ks)
    phi = e.energy_batch(np.ascontiguousarray(z["s"]))
    rel = np.abs(phi - z["phi_cpp"]) / z["phi_cpp"]
    assert rel.max() <= 1e-12, rel.max()


def test_batched_equals_single_bitwise(gpu_ctx) -> None:
    e, ks, box = _energy(64, 2, 5.2)
    s = np.random.default_rng(3).uniform(size=(5, 64, 2))
    phi, g = e.value_and_grad_batch(s)
    for i in range(5):
        f1, g1 = e.value_and_grad(s[i])
        assert f1 == phi[i]
        assert np.array_equal(g1, g[i])


def test_call_and_value_and_grad_agree_bitwise(gpu_ctx) -> None:
    e, ks, box = _energy(40, 3, 4.0)
    s = np.random.default_rng(5).uniform(size=(40, 3))
    f, g = e.value_and_grad(s)
    f2, g2, _, _ = e.value_grad_density(s)
    assert f == e(s) == f2 and np.array_equal(g, g2)


def test_single_particle_and_lattice(gpu_ctx) -> None:
    e, ks, box = _energy(1, 3, 9.0)
    f = e(np.array([[0.3, 0.1, 0.77]]))
    assert f == pytest.approx(ks.weights.sum() / box.volume, rel=1e-14)
    e2, _, _ = _energy(16, 2, 5.0)
    assert e2(square_lattice(4)) < 1e-26


def test_translation_invariance_and_grad_sum(gpu_ctx) -> None:
    e, ks, box = _energy(120, 3, 3.6)
    s = np.random.default_rng(9).uniform(size=(120, 3))
    f0, g = e.value_and_grad(s)
    f1 = e(s + np.array([0.31, -0.2, 0.05]))
    assert f1 == pytest.approx(f0, rel=1e-12)
    assert np.abs(g.sum(axis=0)).max() <= 1e-9 * np.abs(g).max()


def test_arbitrary_rows_negative_first_component(gpu_ctx) -> None:
    """A k-set with n1 < 0 rows, duplicates and non-unit weights (StealthyEnergy fields)."""
    n = np.array([[-1, 2], [0, 1], [2, -1], [-2, 1], [1, 1], [1, 1]], dtype=np.float64)
    w = np.array([1.0, 0.5, 2.0, 1.0, 0.25, 0.75])
    e = StealthyEnergy(n, w, 9.0)
    s = np.random.default_rng(2).uniform(size=(11, 2))
    f, g, re, im = e.value_grad_density(s)
    fo, go, reo, imo = O.energy_grad(s, n.astype(np.int64), w, 9.0)
    assert f == pytest.approx(fo, rel=1e-12)
    assert np.allclose(g, go, rtol=0, atol=1e-10 * np.abs(go).max())
    assert np.allclose(re, reo, atol=1e-12) and np.allclose(im, imo, atol=1e-12)


def test_collective_density(gpu_ctx) -> None:
    nv, w, vol, _ = small_kset(30, 2, 5.0)
    s = np.random.default_rng(4).uniform(size=(30, 2))
    re, im = collective_density(s, nv.astype(np.float64))
    reo, imo = O.density(s, nv)
    assert np.allclose(re, reo, atol=1e-12) and np.allclose(im, imo, atol=1e-12)


def test_rejects_float32_and_bad_shapes(gpu_ctx) -> None:
    e, ks, box = _energy(8, 2, 4.0)
    with pytest.raises(TypeError):
        e.value_and_grad(np.zeros((8, 2), np.float32))
    with pytest.raises(ValueError):
        e.value_and_grad(np.zeros((8, 3)))
    with pytest.raises(ValueError):
        make_stealthy_energy(box, K.enumerate_kvectors(box.basis, 0.0, 0.1))


# ---------------- Hessian (energy.py:272-299) ----------------

RTOL_H = 1e-11  # of max |H|; the pair term is a K = 2M FP64 product


@pytest.mark.parametrize("n,d,k_outer", [(1, 2, 9.0), (12, 1, 9.0), (16, 2, 4.0), (30, 3, 3.4), (150, 2, 5.5)])
def test_hessian_matches_oracle(gpu_ctx, n: int, d: int, k_outer: float) -> None:
    e, ks, box = _energy(n, d, k_outer)
    s = np.random.default_rng(40 + n).uniform(size=(n, d))
    H = e.hessian(s)
    Ho = O.hessian(s, ks.n, ks.weights, box.volume)
    assert H.shape == (n * d, n * d)
    # the size of the summed terms: a vanishing Hessian (one particle) is judged against it
    terms = 2.0 * (2 * np.pi) ** 2 / box.volume * float(np.sum(ks.weights * np.sum(ks.n**2, axis=1))) * n
    scale = max(np.abs(Ho).max(), terms)
    assert np.abs(H - Ho).max() <= RTOL_H * scale, np.abs(H - Ho).max() / scale
    assert np.array_equal(H, H.T) or np.abs(H - H.T).max() <= 1e-13 * scale


def test_hessian_batch_equals_single_and_null_space(gpu_ctx) -> None:
    e, ks, box = _energy(40, 2, 4.6)
    s = np.random.default_rng(41).uniform(size=(3, 40, 2))
    Hb = e.hessian_batch(s)
    for i in range(3):
        assert np.array_equal(Hb[i], e.hessian(s[i]))
    scale = np.abs(Hb[0]).max()
    for b in range(2):
        t = np.zeros((40, 2))
        t[:, b] = 1.0
        assert np.abs(Hb[0] @ t.reshape(-1)).max() < 1e-10 * scale
    # the gradient's finite difference along one coordinate is the matching column
    h = 1e-6
    sp, sm = s[0].copy(), s[0].copy()
    sp[5, 1] += h
    sm[5, 1] -= h
    col = (e.value_and_grad(sp)[1] - e.value_and_grad(sm)[1]).reshape(-1) / (2 * h)
    assert np.allclose(Hb[0][:, 5 * 2 + 1], col, atol=1e-5 * scale)
    with pytest.raises(TypeError):
        e.hessian(s[0].astype(np.float32))
    with pytest.raises(ValueError):
        e.hessian(s[0][:, :1].copy())
