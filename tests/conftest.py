"""Shared test setup: import paths, the ``gpu`` marker, GPU fixtures.

``-m "not gpu"`` runs the oracle pins, host logic and the ABI export check on
any machine; ``-m gpu`` runs the HIP parity tests on an MI355X.
"""

from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "xla-cco_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP library on a visible GPU)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def gpu_ctx():
    """The HIP context of device 0; fails (not skips) when the library is missing."""
    from xla_cco import _native

    return _native.context(0)


@pytest.fixture
def rng() -> np.random.Generator:
    return np.random.default_rng(1234)
