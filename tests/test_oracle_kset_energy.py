"""Pins of the oracle's k-set and energy against the reference's known answers.

CPU only: the oracle must be right before it judges the HIP path.
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from systems import chi_kset, small_kset
from xla_cco import kvectors as K
from xla_cco.box import cubic_box

GOLDEN = Path(__file__).resolve().parent / "golden" / "md_cxx_phi.npz"


# ---------------- k-set (kvectors.py:60-169; test_kvectors.py:46-142) ----------------


def test_ka_and_k_outer_known_answers() -> None:
    """ka_from_chi(0.05, 0.05, 3) and K at d=3 N=100 (test_kvectors.py:46-73)."""
    ka = K.ka_from_chi(0.05, 0.05, 3)
    assert ka == pytest.approx(0.5963271031078479, rel=1e-15)
    k = K.k_outer_from_ka(0.5963271031078479, 0.05, 3)
    assert k == pytest.approx(2.6093, abs=1e-4)
    assert (0.05 / K.unit_ball_volume(3)) ** (1 / 3) == pytest.approx(0.228539, abs=1e-6)
    assert O.ka_from_chi(0.05, 0.05, 3) == ka
    assert O.k_outer_from_ka(ka, 0.05, 3) == K.k_outer_from_ka(ka, 0.05, 3)


@pytest.mark.parametrize("chi,m", [(0.05, 13), (0.1, 28), (0.2, 61), (0.3, 89), (0.35, 101), (0.45, 128)])
def test_kset_counts_d3_n100(chi: float, m: int) -> None:
    """M at d=3 N=100 (benchmarks/RESULTS.md:96-100, 118-120)."""
    n, _, _, basis = chi_kset(100, 3, chi)
    assert n.shape[0] == m
    if chi == 0.05:
        assert K.chi_from_count(m, 100, 3) == pytest.approx(0.043771, abs=1e-6)


def test_kset_cpp_fixture_bits() -> None:
    """K bits and the 50 n-vectors of the C++ MD fixture (test_md_cxx_reference.py:165-177)."""
    z = np.load(GOLDEN)
    k = K.k_outer_from_ka(float(z["Ka"]), float(z["phi_fake"]), 2)
    assert float.hex(k) == str(z["K_hex"])
    ks = K.enumerate_kvectors(z["basis"], 0.0, float(z["K"]))
    assert ks.n.shape[0] == int(z["M"]) == 50
    assert sorted(map(tuple, ks.n.tolist())) == sorted(map(tuple, z["n_vectors"].tolist()))


@pytest.mark.parametrize("d,n,k_outer", [(1, 50, 9.0), (2, 64, 4.0), (2, 37, 3.3), (3, 27, 3.0)])
def test_enumerator_equals_bruteforce(d: int, n: int, k_outer: float) -> None:
    """Vectorised host enumerator == brute-force integer scan, rows and k_sq bits."""
    basis = cubic_box(n, d).basis
    ks = K.enumerate_kvectors(basis, 0.0, k_outer)
    nb, qb = O.halfspace_kvectors(basis, 0.0, k_outer)
    assert np.array_equal(ks.n, nb)
    assert np.array_equal(ks.k_sq, qb)


def test_kset_sheared_basis_and_inner_cutoff() -> None:
    basis = np.array([[3.0, 0.0], [1.2, 2.5]])
    ks = K.enumerate_kvectors(basis, 1.0, 6.0)
    nb, qb = O.halfspace_kvectors(basis, 1.0, 6.0)
    assert np.array_equal(ks.n, nb) and np.array_equal(ks.k_sq, qb)
    assert np.all(ks.k_sq > 1.0) and np.all(ks.k_sq < 36.0)


def test_cubic_shell_count_closed_form() -> None:
    """Unit-cell shells: count of integer points with 0 < |n|^2 < R^2 over two (half space)."""
    basis = 2.0 * math.pi * np.eye(3)  # k = n exactly
    ks = K.enumerate_kvectors(basis, 0.0, math.sqrt(5.5))
    full = sum(
        1
        for a in range(-3, 4)
        for b in range(-3, 4)
        for c in range(-3, 4)
        if 0 < a * a + b * b + c * c < 5.5
    )
    assert ks.n.shape[0] == full // 2


def test_kset_strict_shell_ulps() -> None:
    """A k exactly on the outer shell is excluded, one ulp outside included (test_kvectors.py:86-100)."""
    basis = 2.0 * math.pi * np.eye(2)
    k_exact = 2.0
    assert not any(np.all(r == [2, 0]) for r in K.enumerate_kvectors(basis, 0.0, k_exact).n)
    up = np.nextafter(k_exact, 3.0)
    assert any(np.all(r == [2, 0]) for r in K.enumerate_kvectors(basis, 0.0, up).n)


def test_kset_rejects_bad_input() -> None:
    with pytest.raises(ValueError):
        K.enumerate_kvectors(np.eye(2), 2.0, 1.0)
    with pytest.raises(ValueError):
        K.enumerate_kvectors(np.zeros((2, 2)), 0.0, 1.0)


# ---------------- energy (energy.py:45-50, 186-225, 301-324) ----------------


def test_oracle_energy_matches_cpp_golden() -> None:
    """C++ energies of 15 stored configurations, relative 1e-12 (reference gate 1.9e-13)."""
    z = np.load(GOLDEN)
    ks = K.enumerate_kvectors(z["basis"], 0.0, float(z["K"]))
    vol = float(abs(np.linalg.det(z["basis"])))
    for i in range(z["s"].shape[0]):
        f, _, _, _ = O.energy_grad(z["s"][i], ks.n, ks.weights, vol)
        assert abs(f - z["phi_cpp"][i]) <= 1e-12 * z["phi_cpp"][i]


def test_vectorised_energy_equals_loops() -> None:
    n, w, vol, _ = small_kset(12, 2, 4.0)
    s = np.random.default_rng(0).uniform(size=(12, 2))
    f, _, _, _ = O.energy_grad(s, n, w, vol)
    assert f == pytest.approx(O.energy_loops(s, n, w, vol), rel=1e-13)


def test_single_particle_closed_form() -> None:
    """One particle: |rho_k| = 1 so Phi = sum(w) / V (test_energy.py:113-119)."""
    n, w, vol, _ = small_kset(1, 3, 9.0)
    s = np.array([[0.3, 0.1, 0.77]])
    f, g, re, im = O.energy_grad(s, n, w, vol)
    assert f == pytest.approx(w.sum() / vol, rel=1e-14)
    assert np.allclose(re * re + im * im, 1.0)
    assert np.max(np.abs(g)) < 1e-12


def test_lattice_below_first_shell_is_zero() -> None:
    """A 4x4 lattice with K below 2 pi/a has rho_k = 0 (test_energy.py:122-133)."""
    side = 4
    g1 = np.arange(side) / side
    s = np.stack(np.meshgrid(g1, g1, indexing="ij"), axis=-1).reshape(-1, 2)
    n, w, vol, _ = small_kset(16, 2, 5.0)
    f, _, _, _ = O.energy_grad(s, n, w, vol)
    assert f < 1e-26


def test_gradient_matches_finite_difference() -> None:
    n, w, vol, _ = small_kset(10, 2, 4.0)
    s = np.random.default_rng(1).uniform(size=(10, 2))
    _, g, _, _ = O.energy_grad(s, n, w, vol)
    h = 1e-6
    fd = np.zeros_like(s)
    for j in range(s.shape[0]):
        for a in range(2):
            sp, sm = s.copy(), s.copy()
            sp[j, a] += h
            sm[j, a] -= h
            fd[j, a] = (O.energy_grad(sp, n, w, vol)[0] - O.energy_grad(sm, n, w, vol)[0]) / (2 * h)
    assert np.allclose(g, fd, atol=1e-6 * max(1.0, np.abs(g).max()))


def test_translation_invariance_and_grad_sum() -> None:
    n, w, vol, _ = small_kset(20, 3, 3.5)
    s = np.random.default_rng(2).uniform(size=(20, 3))
    f0, g, _, _ = O.energy_grad(s, n, w, vol)
    f1, _, _, _ = O.energy_grad(s + np.array([0.13, -0.4, 0.71]), n, w, vol)
    assert f1 == pytest.approx(f0, rel=1e-12)
    assert np.abs(g.sum(axis=0)).max() < 1e-10 * np.abs(g).max() * 20


def test_gradient_equals_cpp_cartesian_relation() -> None:
    """dPhi/ds = -F B^T with the C++ Cartesian forces (energy.py:198-199)."""
    n, w, vol, basis = small_kset(9, 2, 4.5)
    s = np.random.default_rng(3).uniform(size=(9, 2))
    _, g, _, _ = O.energy_grad(s, n, w, vol)
    F = O.cartesian_forces(s, n, w, basis, vol)
    assert np.allclose(g, -F @ basis.T, rtol=1e-10, atol=1e-12)


def test_a_planted_sign_error_is_caught() -> None:
    """The FD pin fails for a gradient with a flipped sin term (planted in the oracle copy)."""
    n, w, vol, _ = small_kset(8, 2, 4.0)
    s = np.random.default_rng(4).uniform(size=(8, 2))
    nf = n.astype(np.float64)
    re, im = O.density(s, nf)
    th = 2 * math.pi * (s @ nf.T)
    bad = (2 * 2 * math.pi / vol) * ((w * (im * np.cos(th) + re * np.sin(th))) @ nf)
    _, good, _, _ = O.energy_grad(s, n, w, vol)
    h = 1e-6
    sp, sm = s.copy(), s.copy()
    sp[0, 0] += h
    sm[0, 0] -= h
    fd = (O.energy_grad(sp, n, w, vol)[0] - O.energy_grad(sm, n, w, vol)[0]) / (2 * h)
    assert abs(good[0, 0] - fd) < 1e-6
    assert abs(bad[0, 0] - fd) > 1e-3


# ---------------- Hessian (energy.py:272-299) ----------------


def _fd_hessian(s, n, w, vol, h=1e-6):
    N, d = s.shape
    out = np.zeros((N * d, N * d))
    for j in range(N):
        for a in range(d):
            sp, sm = s.copy(), s.copy()
            sp[j, a] += h
            sm[j, a] -= h
            gp = O.energy_grad(sp, n, w, vol)[1].reshape(-1)
            gm = O.energy_grad(sm, n, w, vol)[1].reshape(-1)
            out[:, j * d + a] = (gp - gm) / (2 * h)
    return out


@pytest.mark.parametrize("N,d,k_outer", [(8, 2, 4.0), (6, 3, 3.2), (12, 1, 9.0)])
def test_hessian_matches_finite_difference_of_gradient(N: int, d: int, k_outer: float) -> None:
    n, w, vol, _ = small_kset(N, d, k_outer)
    s = np.random.default_rng(30 + N).uniform(size=(N, d))
    H = O.hessian(s, n, w, vol)
    fd = _fd_hessian(s, n, w, vol)
    assert np.allclose(H, fd, atol=2e-6 * max(1.0, np.abs(H).max()))


def test_hessian_symmetric_with_translation_null_space() -> None:
    n, w, vol, _ = small_kset(10, 3, 3.5)
    s = np.random.default_rng(31).uniform(size=(10, 3))
    H = O.hessian(s, n, w, vol)
    scale = np.abs(H).max()
    assert np.allclose(H, H.T, atol=1e-13 * scale)
    # Phi(s + c) = Phi(s): the Hessian maps a uniform translation along any axis to zero
    for b in range(3):
        t = np.zeros((10, 3))
        t[:, b] = 1.0
        assert np.abs(H @ t.reshape(-1)).max() < 1e-11 * scale


def test_hessian_planted_missing_diagonal_is_caught() -> None:
    """Dropping the delta_jl term (planted in a copy) fails the translation pin and the FD pin."""
    N, d = 7, 2
    n, w, vol, _ = small_kset(N, d, 4.0)
    s = np.random.default_rng(32).uniform(size=(N, d))
    nf = n.astype(np.float64)
    th = 2 * math.pi * (s @ nf.T)
    c, sn = np.cos(th), np.sin(th)
    pref = 2.0 * (2 * math.pi) ** 2 / vol
    bad = np.zeros((N, d, N, d))
    for a in range(d):
        for b in range(d):
            wab = w * nf[:, a] * nf[:, b]
            bad[:, a, :, b] = pref * ((c * wab) @ c.T + (sn * wab) @ sn.T)
    bad = bad.reshape(N * d, N * d)
    t = np.zeros((N, d))
    t[:, 0] = 1.0
    assert np.abs(bad @ t.reshape(-1)).max() > 1e-3 * np.abs(bad).max()
    assert not np.allclose(bad, _fd_hessian(s, n, w, vol), atol=1e-4)
