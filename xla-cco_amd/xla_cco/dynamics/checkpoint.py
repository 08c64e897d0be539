"""Atomic checkpoints of (batched) MD states.

src/xla_cco/dynamics/checkpoint.py:55-205, same file layout: an ``.npz``
(no pickles) with ``leaf_000``, ``leaf_001``, ... holding the MDState arrays
in field order (nested snapshot and trace fields in theirs), ``__n_leaves__``
(int64) and ``__meta__`` (0-d unicode, JSON).  ``save_state`` writes
``<path>.tmp``, flushes it to disk and renames it over ``path``, so a crash
never leaves a partial checkpoint.  ``load_state`` rebuilds a state with the
structure of a template (``like``: any MDState whose leaves have the wanted
shapes and dtypes); ``read_meta`` reads only the metadata.  Unreadable files
(truncated, garbage, missing entries) raise CheckpointError, a ValueError.
"""

from __future__ import annotations

import contextlib
import json
import os
import zipfile
from pathlib import Path
from typing import Any, Dict, List, Tuple

import numpy as np

from xla_cco.dynamics.state import MDState, leaves, with_leaves

META_KEY = "__meta__"
COUNT_KEY = "__n_leaves__"
LEAF_FORMAT = "leaf_{:03d}"
TMP_SUFFIX = ".tmp"


class CheckpointError(ValueError):
    """A checkpoint file that is unreadable or does not match its template."""


def tmp_path_of(path: Path) -> Path:
    """Path of the temporary file that ``save_state`` renames over ``path``."""
    return Path(str(path) + TMP_SUFFIX)


def save_state(path: Path, state: MDState, meta: Dict[str, Any]) -> None:
    """Write ``state`` and ``meta`` atomically to ``path`` (.npz).

    On any error the temporary file is removed and ``path`` keeps its
    previous content.  Raises TypeError when ``state`` is not an MDState or
    ``meta`` is not a JSON-serializable dict.
    """
    if not isinstance(state, MDState):
        raise TypeError(f"state must be an MDState, got {type(state).__name__}")
    if not isinstance(meta, dict):
        raise TypeError(f"meta must be a dict, got {type(meta).__name__}")
    meta_json = json.dumps(meta, sort_keys=True)  # TypeError if not serializable
    arrays: Dict[str, Any] = {LEAF_FORMAT.format(i): np.asarray(x) for i, x in enumerate(leaves(state))}
    arrays[META_KEY] = np.asarray(meta_json)
    arrays[COUNT_KEY] = np.int64(len(arrays) - 1)
    path = Path(path)
    tmp = tmp_path_of(path)
    try:
        with open(tmp, "wb") as fh:
            np.savez(fh, **arrays)
            fh.flush()
            os.fsync(fh.fileno())
        os.replace(tmp, path)
    except BaseException:
        with contextlib.suppress(OSError):
            tmp.unlink()
        raise


def _meta_of(z: Any, path: Path) -> Dict[str, Any]:
    files = set(z.files)
    if META_KEY not in files or COUNT_KEY not in files:
        raise CheckpointError(f"{path} is not an MD checkpoint (no {META_KEY})")
    meta = json.loads(str(z[META_KEY][()]))
    if not isinstance(meta, dict):
        raise CheckpointError(f"{path}: metadata is not a JSON object")
    return meta


def _leaves_of(z: Any, path: Path, like: MDState) -> List[np.ndarray]:
    specs = [(np.shape(x), np.asarray(x).dtype) for x in leaves(like)]
    n_file = int(z[COUNT_KEY])
    if n_file != len(specs):
        raise CheckpointError(f"{path} holds {n_file} leaves; the template has {len(specs)}")
    out = []
    for i, (shape, dtype) in enumerate(specs):
        a = np.asarray(z[LEAF_FORMAT.format(i)])
        if a.shape != shape or a.dtype != dtype:
            raise CheckpointError(f"{path} leaf {i}: stored {a.shape} {a.dtype}, template {shape} {dtype}")
        out.append(a)
    return out


def _corrupt(path: Path, exc: Exception) -> CheckpointError:
    return CheckpointError(
        f"{path} is not a readable MD checkpoint (truncated or corrupt: {type(exc).__name__}: {exc})"
    )


def read_meta(path: Path) -> Dict[str, Any]:
    """Only the metadata of a checkpoint; ValueError on unreadable files, OSError if unopenable."""
    try:
        with np.load(Path(path), allow_pickle=False) as z:
            return _meta_of(z, Path(path))
    except CheckpointError:
        raise
    except (zipfile.BadZipFile, EOFError, KeyError, ValueError) as exc:
        raise _corrupt(Path(path), exc) from exc


def load_state(path: Path, like: MDState) -> Tuple[MDState, Dict[str, Any]]:
    """(state, meta) read into the structure of ``like``.

    Raises CheckpointError (a ValueError) on truncated or corrupt files, a
    missing entry, or a leaf count, shape or dtype that differs from the
    template; TypeError when ``like`` is not an MDState; OSError when the
    file cannot be opened.
    """
    if not isinstance(like, MDState):
        raise TypeError(f"like must be an MDState, got {type(like).__name__}")
    try:
        with np.load(Path(path), allow_pickle=False) as z:
            meta = _meta_of(z, Path(path))
            arrays = _leaves_of(z, Path(path), like)
    except CheckpointError:
        raise
    except (zipfile.BadZipFile, EOFError, KeyError, ValueError) as exc:
        raise _corrupt(Path(path), exc) from exc
    return with_leaves(like, arrays), meta
