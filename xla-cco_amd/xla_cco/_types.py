"""float64 guard shared by the host layer.

Every array that crosses into the HIP library is float64; anything else is
rejected with ``TypeError`` before any device work starts
(reference: src/xla_cco/_types.py:34-52).
"""

from __future__ import annotations

import numpy as np


def require_float64(x: object, name: str) -> None:
    """Raise ``TypeError`` unless ``x`` carries dtype float64.

    Mirrors ``require_float64`` of src/xla_cco/_types.py:34 (same message).
    """
    assert isinstance(name, str), "name must be a string"
    dtype = getattr(x, "dtype", None)
    if dtype is None or dtype != np.float64:
        raise TypeError(f"{name} must be float64, got dtype {dtype}")
