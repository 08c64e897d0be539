"""ConfigPack reader/writer (the C++ binary format) and fractional stacking.

Format (src/xla_cco/io/configpack.py:74-150, C++ ConfigurationPack):
``PREFIX.ConfigPack`` is a concatenation of records, each
``<u2 dim``, ``<f8 cell``, ``<f8 basis[dim][dim]``, ``<i4 N``, then N times
(``<f8 x[dim]``, ``u1 name[4]``); ``PREFIX.ConfigPackIndex`` is ``<i8 count``
followed by ``count`` ``<i8`` byte offsets.  Positions are fractional with
respect to the stored basis.
"""

from __future__ import annotations

import dataclasses
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

PACK_SUFFIX = ".ConfigPack"
INDEX_SUFFIX = ".ConfigPackIndex"
MAX_DIMENSION = 4
NAME_BYTES = 4
_UNIT_RTOL = 1e-12


@dataclasses.dataclass(frozen=True)
class PackedConfig:
    """One stored configuration: basis (d, d), fractional positions (N, d), cell, names."""

    basis: np.ndarray
    relative: np.ndarray
    cell_size: float
    names: Tuple[bytes, ...]


def _files(prefix: Path) -> Tuple[Path, Path]:
    return Path(str(prefix) + PACK_SUFFIX), Path(str(prefix) + INDEX_SUFFIX)


def _offsets(index_path: Path) -> np.ndarray:
    raw = np.fromfile(index_path, dtype="<i8")
    if raw.size < 1:
        raise ValueError(f"truncated index (no count): {index_path}")
    count = int(raw[0])
    if count < 0 or raw.size != count + 1:
        raise ValueError(f"index {index_path} declares {count} configs but holds {raw.size - 1} offsets")
    return raw[1:]


def _record_types(dim: int) -> Tuple[np.dtype, np.dtype]:
    head = np.dtype([("cell", "<f8"), ("basis", "<f8", (dim, dim)), ("n", "<i4")])
    atom = np.dtype([("x", "<f8", (dim,)), ("name", "u1", (NAME_BYTES,))])
    return head, atom


def _parse(buf: np.ndarray, offset: int) -> PackedConfig:
    size = int(buf.size)
    if not 0 <= offset <= size - 2:
        raise ValueError(f"offset {offset} outside pack ({size} B)")
    dim = int(np.frombuffer(buf, dtype="<u2", count=1, offset=offset)[0])
    if not 1 <= dim <= MAX_DIMENSION:
        raise ValueError(f"bad dimension {dim} at offset {offset}")
    head, atom = _record_types(dim)
    pos = offset + 2
    if pos + head.itemsize > size:
        raise ValueError(f"truncated header at offset {offset}")
    h = np.frombuffer(buf, dtype=head, count=1, offset=pos)[0]
    n = int(h["n"])
    if n < 0:
        raise ValueError(f"negative particle count {n} at offset {offset}")
    pos += head.itemsize
    if pos + n * atom.itemsize > size:
        raise ValueError(f"config at offset {offset} overruns pack ({pos + n * atom.itemsize} > {size} B)")
    recs = np.frombuffer(buf, dtype=atom, count=n, offset=pos)
    names = tuple(bytes(r) for r in np.asarray(recs["name"], dtype=np.uint8))
    return PackedConfig(
        basis=np.array(h["basis"], dtype=np.float64),
        relative=np.array(recs["x"], dtype=np.float64),
        cell_size=float(h["cell"]),
        names=names,
    )


def _serialize(cfg: PackedConfig) -> bytes:
    basis, rel = np.asarray(cfg.basis), np.asarray(cfg.relative)
    for label, arr in (("basis", basis), ("relative", rel)):
        if arr.dtype != np.float64:
            raise TypeError(f"{label} must be float64, got {arr.dtype}")
    if rel.ndim != 2 or not 1 <= rel.shape[1] <= MAX_DIMENSION:
        raise ValueError(f"relative must be (N, d), d in 1..4; got {rel.shape}")
    n, d = rel.shape
    if basis.shape != (d, d):
        raise ValueError(f"basis shape {basis.shape} != ({d}, {d})")
    if len(cfg.names) != n:
        raise ValueError(f"{len(cfg.names)} names for {n} particles")
    if any(len(nm) != NAME_BYTES for nm in cfg.names):
        raise ValueError(f"each name must be exactly {NAME_BYTES} bytes")
    head, atom = _record_types(d)
    h = np.zeros(1, dtype=head)
    h["cell"] = cfg.cell_size
    h["basis"] = basis
    h["n"] = n
    recs = np.zeros(n, dtype=atom)
    recs["x"] = rel
    if n:
        recs["name"] = np.frombuffer(b"".join(cfg.names), dtype=np.uint8).reshape(n, NAME_BYTES)
    return np.array(d, dtype="<u2").tobytes() + h.tobytes() + recs.tobytes()


def read_configpack(prefix: Path, indices: Optional[Sequence[int]] = None) -> List[PackedConfig]:
    """Read configurations (all, or ``indices`` in that order)."""
    pack_path, index_path = _files(Path(prefix))
    if not (pack_path.is_file() and index_path.is_file()):
        raise FileNotFoundError(
            f"need both {pack_path} and {index_path} (pack exists: {pack_path.is_file()}, "
            f"index exists: {index_path.is_file()})"
        )
    offs = _offsets(index_path)
    sel = list(range(offs.size)) if indices is None else list(indices)
    if any(not 0 <= i < offs.size for i in sel):
        raise ValueError(f"config index outside [0, {offs.size})")
    if not sel:
        return []
    if pack_path.stat().st_size == 0:
        raise ValueError(f"empty pack for {offs.size} configs: {pack_path}")
    buf = np.memmap(pack_path, dtype=np.uint8, mode="r")
    try:
        return [_parse(buf, int(offs[i])) for i in sel]
    finally:
        del buf


def write_configpack(prefix: Path, configs: Sequence[PackedConfig]) -> None:
    """Write configurations as a ConfigPack file pair."""
    prefix = Path(prefix)
    if not prefix.parent.is_dir():
        raise FileNotFoundError(f"missing parent directory: {prefix.parent}")
    if len(configs) == 0:
        raise ValueError("refusing to write an empty configuration sequence")
    blobs = [_serialize(c) for c in configs]
    sizes = np.array([len(b) for b in blobs], dtype="<i8")
    offs = np.zeros(len(blobs), dtype="<i8")
    offs[1:] = np.cumsum(sizes)[:-1]
    pack_path, index_path = _files(prefix)
    pack_path.write_bytes(b"".join(blobs))
    np.concatenate(([len(blobs)], offs)).astype("<i8").tofile(index_path)


def stack_fractional(configs: Sequence[PackedConfig]) -> Tuple[np.ndarray, np.ndarray]:
    """((B, N, d) fractional positions, (d, d) unit-volume basis) of equal-shape configs."""
    if len(configs) == 0:
        raise ValueError("no configurations to stack")
    shape = np.asarray(configs[0].relative).shape
    units = []
    for c in configs:
        rel, basis = np.asarray(c.relative), np.asarray(c.basis)
        if rel.dtype != np.float64 or basis.dtype != np.float64:
            raise TypeError("positions and basis must be float64")
        if rel.shape != shape:
            raise ValueError(f"mixed shapes {rel.shape} and {shape}")
        d = shape[1]
        if basis.shape != (d, d):
            raise ValueError(f"basis shape {basis.shape} != ({d}, {d})")
        det = float(np.linalg.det(basis))
        if not det > 0.0:
            raise ValueError(f"non-positive basis determinant {det}")
        units.append(basis / det ** (1.0 / d))
    for ub in units[1:]:
        if not np.allclose(ub, units[0], rtol=_UNIT_RTOL, atol=_UNIT_RTOL):
            raise ValueError("bases differ across configurations")
    return np.stack([np.asarray(c.relative) for c in configs]), units[0]
