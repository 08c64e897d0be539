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
