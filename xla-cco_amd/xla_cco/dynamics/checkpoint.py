"""Crash-safe checkpoints of (batched) MD states.

File layout (the reference's, src/xla_cco/dynamics/checkpoint.py:55-205, so that
either code base reads the other's files): one ``.npz`` archive without pickles,
entries ``leaf_000``, ``leaf_001``, ... with the state's arrays in field order
(the snapshot and trace records expanded in theirs), ``__n_leaves__`` (int64) and
``__meta__`` (0-d unicode holding a JSON object).

Writing goes to ``<path>.tmp``, which is flushed to disk and then renamed over
``path``; a failure at any point removes the temporary file and leaves ``path`` as
it was.  Reading maps every archive-level failure (not a zip, truncated, an entry
missing, metadata that is not JSON) to :class:`CheckpointError`, a ValueError, and
checks each stored array against the shape and dtype of a template state.
"""

from __future__ import annotations

import contextlib
import json
import os
import zipfile
from pathlib import Path
from typing import Any, Dict, Iterator, List, Tuple

import numpy as np

from xla_cco.dynamics.state import MDState, leaves, with_leaves

META_KEY = "__meta__"
COUNT_KEY = "__n_leaves__"
LEAF_FORMAT = "leaf_{:03d}"
TMP_SUFFIX = ".tmp"

# what a damaged archive raises inside numpy's reader
_ARCHIVE_ERRORS = (zipfile.BadZipFile, EOFError, KeyError, ValueError, json.JSONDecodeError)


class CheckpointError(ValueError):
    """A checkpoint that cannot be read, or that does not fit its template state."""


def tmp_path_of(path: Path) -> Path:
    """The temporary file ``save_state`` writes before renaming it over ``path``."""
    return Path(f"{path}{TMP_SUFFIX}")


def _entries(state: MDState, meta: Dict[str, Any]) -> Dict[str, np.ndarray]:
    arrays = [np.asarray(a) for a in leaves(state)]
    out = {LEAF_FORMAT.format(i): a for i, a in enumerate(arrays)}
    out[COUNT_KEY] = np.int64(len(arrays))
    out[META_KEY] = np.asarray(json.dumps(meta, sort_keys=True))
    return out


def save_state(path: Path, state: MDState, meta: Dict[str, Any]) -> None:
    """Write ``state`` and the JSON-serialisable ``meta`` to ``path`` atomically.

    TypeError for a state that is not an MDState or metadata that is not a
    JSON-serialisable dict (nothing is written then).
    """
    if not isinstance(state, MDState):
        raise TypeError(f"save_state needs an MDState, not {type(state).__name__}")
    if not isinstance(meta, dict):
        raise TypeError(f"checkpoint metadata must be a dict, not {type(meta).__name__}")
    entries = _entries(state, meta)  # json raises TypeError on unserialisable values
    target = Path(path)
    staging = tmp_path_of(target)
    try:
        with open(staging, "wb") as fh:
            np.savez(fh, **entries)
            fh.flush()
            os.fsync(fh.fileno())
        os.replace(staging, target)
    except BaseException:
        with contextlib.suppress(OSError):
            os.unlink(staging)
        raise


@contextlib.contextmanager
def _archive(path: Path) -> Iterator[Any]:
    """The open archive; damage inside it surfaces as CheckpointError."""
    try:
        with np.load(path, allow_pickle=False) as z:
            yield z
    except CheckpointError:
        raise
    except _ARCHIVE_ERRORS as exc:
        raise CheckpointError(
            f"{path}: truncated or corrupt checkpoint ({type(exc).__name__}: {exc})"
        ) from exc


def _read_meta(z: Any, path: Path) -> Dict[str, Any]:
    if META_KEY not in z.files or COUNT_KEY not in z.files:
        raise CheckpointError(f"{path} is not an MD checkpoint (entries {META_KEY}, {COUNT_KEY} absent)")
    meta = json.loads(str(z[META_KEY][()]))
    if not isinstance(meta, dict):
        raise CheckpointError(f"{path}: the metadata is not a JSON object ({type(meta).__name__})")
    return meta


def _read_leaves(z: Any, path: Path, like: MDState) -> List[np.ndarray]:
    template = [np.asarray(a) for a in leaves(like)]
    stored = int(z[COUNT_KEY])
    if stored != len(template):
        raise CheckpointError(f"{path}: {stored} leaf arrays stored, the template state has {len(template)}")
    out = []
    for i, want in enumerate(template):
        got = np.asarray(z[LEAF_FORMAT.format(i)])
        if got.shape != want.shape or got.dtype != want.dtype:
            raise CheckpointError(
                f"{path}: leaf {i} is {got.dtype}{list(got.shape)}, the template wants {want.dtype}{list(want.shape)}"
            )
        out.append(got)
    return out


def read_meta(path: Path) -> Dict[str, Any]:
    """The metadata alone.  CheckpointError for a damaged file, OSError if it cannot be opened."""
    with _archive(Path(path)) as z:
        return _read_meta(z, Path(path))


def load_state(path: Path, like: MDState) -> Tuple[MDState, Dict[str, Any]]:
    """(state, metadata), the state shaped like the template ``like``.

    CheckpointError (a ValueError) for a damaged file or one whose arrays do not
    match the template; TypeError if ``like`` is not an MDState; OSError if the
    file cannot be opened.
    """
    if not isinstance(like, MDState):
        raise TypeError(f"load_state needs an MDState template, not {type(like).__name__}")
    with _archive(Path(path)) as z:
        meta = _read_meta(z, Path(path))
        arrays = _read_leaves(z, Path(path), like)
    return with_leaves(like, arrays), meta
