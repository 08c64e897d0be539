"""C++ ``WriteFunction`` table writer (src/xla_cco/io/tables.py:36-69).

Four tab-separated ``%1.10e`` columns per row with a trailing tab; missing
columns are 0.0; NaN prints as glibc does ("nan" / "-nan"); inf is refused.
"""

from __future__ import annotations

from pathlib import Path
from typing import Sequence

import numpy as np

N_FIELDS = 4


def write_function_table(path: Path, columns: Sequence[np.ndarray]) -> None:
    if not 1 <= len(columns) <= N_FIELDS:
        raise ValueError(f"need 1..{N_FIELDS} columns, got {len(columns)}")
    cols = [np.asarray(c, dtype=np.float64) for c in columns]
    if any(c.ndim != 1 for c in cols):
        raise ValueError("columns must be 1-D")
    rows = cols[0].size
    if any(c.size != rows for c in cols):
        raise ValueError("columns have unequal lengths")
    table = np.zeros((rows, N_FIELDS))
    for j, c in enumerate(cols):
        table[:, j] = c
    if np.isinf(table).any():
        raise ValueError("infinite value in table")

    def fmt(v: float) -> str:
        if v != v:
            return "-nan" if np.signbit(v) else "nan"
        return f"{v:1.10e}"

    lines = ["\t".join(fmt(v) for v in row) + "\t\n" for row in table.tolist()]
    Path(path).write_bytes("".join(lines).encode("ascii"))
