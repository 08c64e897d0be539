"""Host logic of the harmonic draw (dynamics/thermal.py:47-153) on energies with known Hessians."""

from __future__ import annotations

import numpy as np
import pytest

from xla_cco.dynamics.thermal import cartesian_hessian, harmonic_sample
from xla_cco.energy import AbstractEnergy


class _Quadratic(AbstractEnergy):
    """Phi(s) = 1/2 sum_i k_i s_i^2 over the flattened coordinates: Hessian diag(k)."""

    def __init__(self, k: np.ndarray) -> None:
        self.k = np.asarray(k, dtype=np.float64)

    def __call__(self, s: np.ndarray) -> float:
        x = np.asarray(s).reshape(-1)
        return float(0.5 * np.sum(self.k * x * x))

    def value_and_grad(self, s: np.ndarray):  # type: ignore[no-untyped-def]
        x = np.asarray(s).reshape(-1)
        return self(s), (self.k * x).reshape(np.shape(s))

    def hessian(self, s: np.ndarray) -> np.ndarray:
        return np.diag(self.k)


def test_cartesian_hessian_transforms_with_the_basis() -> None:
    k = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    e = _Quadratic(k)
    s = np.zeros((3, 2))
    basis = np.array([[2.0, 0.0], [0.0, 4.0]])
    H = cartesian_hessian(e, s, basis)
    # s = r inv(B): axis a scales by 1 / B_aa twice
    want = np.diag(k / np.tile([4.0, 16.0], 3))
    assert np.allclose(H, want, rtol=0, atol=1e-15)
    # a sheared basis: H_r = (I (x) inv(B)) H_s (I (x) inv(B))^T
    sh = np.array([[1.0, 0.0], [0.5, 1.0]])
    inv = np.linalg.inv(sh)
    T = np.kron(np.eye(3), inv)
    assert np.allclose(cartesian_hessian(e, s, sh), T @ np.diag(k) @ T.T, atol=1e-14)


def test_harmonic_sample_draw_and_mode_selection() -> None:
    k = np.array([0.0, 1e-12, 2.0, 8.0])  # the two soft modes are dropped
    e = _Quadratic(k)
    s = np.array([[0.1, 0.2], [0.3, 0.4]])
    basis = np.eye(2)
    assert np.array_equal(harmonic_sample(np.random.default_rng(0), e, s, basis, 0.0), s)
    out = harmonic_sample(np.random.default_rng(5), e, s, basis, 0.5)
    z = np.random.default_rng(5).standard_normal(2)
    disp = np.zeros(4)
    disp[2] = np.sqrt(0.5 / 2.0) * z[0]
    disp[3] = np.sqrt(0.5 / 8.0) * z[1]
    assert np.allclose(np.abs(out - s).reshape(-1), np.abs(disp), atol=1e-15)


def test_harmonic_sample_rejects_bad_input() -> None:
    e = _Quadratic(np.ones(4))
    s = np.zeros((2, 2))
    with pytest.raises(ValueError):
        harmonic_sample(np.random.default_rng(0), e, s, np.eye(2), -1.0)
    with pytest.raises(ValueError):
        harmonic_sample(np.random.default_rng(0), e, s, np.eye(2), float("nan"))
    with pytest.raises(TypeError):
        harmonic_sample(0, e, s, np.eye(2), 1.0)  # type: ignore[arg-type]
    with pytest.raises(ValueError):
        harmonic_sample(np.random.default_rng(0), e, s, np.eye(3), 1.0)
    with pytest.raises(ValueError):
        harmonic_sample(np.random.default_rng(0), _Quadratic(-np.ones(4)), s, np.eye(2), 1.0)
