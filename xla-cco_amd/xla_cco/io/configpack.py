"""ConfigPack files: the C++ ConfigurationPack binary format.

The byte layout is fixed by the C++ code and the reference's reader
(src/xla_cco/io/configpack.py:74-150): ``PREFIX.ConfigPack`` holds records back
to back, each

    uint16 dim | float64 cell | float64 basis[dim][dim] | int32 N |
    N x (float64 x[dim] | 4 name bytes)

little-endian, positions fractional in the stored basis; ``PREFIX.ConfigPackIndex``
is an int64 record count followed by one int64 byte offset per record.

Here a record is encoded by a small codec object per dimension (field offsets
computed once), the writer streams records to the file while it collects the
offsets, and the reader validates every offset and length against the file size
before it decodes.  Errors follow the reference's contract: a missing file is
FileNotFoundError, a malformed or truncated file and inconsistent shapes are
ValueError, non-float64 arrays are TypeError.
"""

from __future__ import annotations

import dataclasses
import functools
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

PACK_SUFFIX = ".ConfigPack"
INDEX_SUFFIX = ".ConfigPackIndex"
MAX_DIMENSION = 4  # the C++ MaxDimension
NAME_BYTES = 4
_UNIT_RTOL = 1e-12  # agreement of unit-volume bases across a stack


@dataclasses.dataclass(frozen=True)
class PackedConfig:
    """A stored configuration: basis (d, d), fractional positions (N, d), cell size, names."""

    basis: np.ndarray
    relative: np.ndarray
    cell_size: float
    names: Tuple[bytes, ...]


class _Codec:
    """Byte layout of the records of one dimension."""

    def __init__(self, dim: int) -> None:
        self.dim = dim
        # after the uint16 dim: cell, basis, N
        self.fixed = np.dtype({"names": ["cell", "basis", "n"],
                               "formats": ["<f8", ("<f8", (dim, dim)), "<i4"]})
        self.atom = np.dtype({"names": ["x", "tag"], "formats": [("<f8", (dim,)), ("V", NAME_BYTES)]})
        self.head_bytes = 2 + self.fixed.itemsize

    def encode(self, cfg: PackedConfig) -> bytes:
        basis, rel = np.asarray(cfg.basis), np.asarray(cfg.relative)
        n = rel.shape[0]
        fixed = np.zeros((), dtype=self.fixed)
        fixed["cell"], fixed["basis"], fixed["n"] = cfg.cell_size, basis, n
        atoms = np.zeros(n, dtype=self.atom)
        atoms["x"] = rel
        atoms["tag"] = np.array(list(cfg.names), dtype=f"V{NAME_BYTES}") if n else atoms["tag"]
        return np.uint16(self.dim).astype("<u2").tobytes() + fixed.tobytes() + atoms.tobytes()

    def decode(self, data: np.ndarray, at: int) -> PackedConfig:
        end = data.size
        if at + self.head_bytes > end:
            raise ValueError(f"record header at byte {at} runs past the end of the pack ({end} B)")
        fixed = data[at + 2: at + self.head_bytes].view(self.fixed)[0]
        n = int(fixed["n"])
        if n < 0:
            raise ValueError(f"record at byte {at} declares {n} particles")
        body = at + self.head_bytes
        stop = body + n * self.atom.itemsize
        if stop > end:
            raise ValueError(f"record at byte {at} needs {stop} B, the pack has {end} B")
        atoms = data[body:stop].view(self.atom)
        return PackedConfig(
            basis=np.array(fixed["basis"], dtype=np.float64),
            relative=np.array(atoms["x"], dtype=np.float64).reshape(n, self.dim),
            cell_size=float(fixed["cell"]),
            names=tuple(t.tobytes() for t in atoms["tag"]),
        )


@functools.lru_cache(maxsize=None)
def _codec(dim: int) -> _Codec:
    return _Codec(dim)


def _pair(prefix: Path) -> Tuple[Path, Path]:
    text = str(prefix)
    return Path(text + PACK_SUFFIX), Path(text + INDEX_SUFFIX)


def _check_config(cfg: PackedConfig) -> int:
    """Validate one configuration for writing; return its dimension."""
    basis, rel = np.asarray(cfg.basis), np.asarray(cfg.relative)
    if basis.dtype != np.float64 or rel.dtype != np.float64:
        raise TypeError(f"basis and relative must be float64 (got {basis.dtype}, {rel.dtype})")
    if rel.ndim != 2 or not 1 <= rel.shape[1] <= MAX_DIMENSION:
        raise ValueError(f"relative positions must be (N, d) with 1 <= d <= {MAX_DIMENSION}, got {rel.shape}")
    n, dim = rel.shape
    if basis.shape != (dim, dim):
        raise ValueError(f"a {dim}-d configuration needs a ({dim}, {dim}) basis, got {basis.shape}")
    if len(cfg.names) != n:
        raise ValueError(f"{n} particles but {len(cfg.names)} names")
    for tag in cfg.names:
        if len(tag) != NAME_BYTES:
            raise ValueError(f"particle names are exactly {NAME_BYTES} bytes, got {len(tag)}")
    return dim


def _offsets(index_path: Path, pack_bytes: int) -> np.ndarray:
    words = np.fromfile(index_path, dtype="<i8")
    if words.size == 0:
        raise ValueError(f"index {index_path} has no record count")
    count = int(words[0])
    if count < 0 or words.size - 1 != count:
        raise ValueError(f"index {index_path}: count {count} with {words.size - 1} offsets")
    offs = words[1:]
    if offs.size and pack_bytes == 0:
        raise ValueError(f"index {index_path} lists {offs.size} records for an empty pack")
    return offs


def read_configpack(prefix: Path, indices: Optional[Sequence[int]] = None) -> List[PackedConfig]:
    """Read all records of a pack, or those at ``indices`` (in that order, repeats allowed)."""
    pack_path, index_path = _pair(Path(prefix))
    missing = [p for p in (pack_path, index_path) if not p.is_file()]
    if missing:
        raise FileNotFoundError(f"ConfigPack file(s) not found: {', '.join(map(str, missing))}")
    size = pack_path.stat().st_size
    offs = _offsets(index_path, size)
    wanted = range(offs.size) if indices is None else [int(i) for i in indices]
    bad = [i for i in wanted if i < 0 or i >= offs.size]
    if bad:
        raise ValueError(f"record index {bad[0]} outside [0, {offs.size})")
    if len(wanted) == 0:
        return []
    data = np.memmap(pack_path, dtype=np.uint8, mode="r")
    try:
        out = []
        for i in wanted:
            at = int(offs[i])
            if at < 0 or at + 2 > size:
                raise ValueError(f"record {i} starts at byte {at}, outside the {size}-byte pack")
            dim = int(data[at: at + 2].view("<u2")[0])
            if not 1 <= dim <= MAX_DIMENSION:
                raise ValueError(f"record {i} at byte {at}: dimension {dim} not in 1..{MAX_DIMENSION}")
            out.append(_codec(dim).decode(data, at))
        return out
    finally:
        del data


def write_configpack(prefix: Path, configs: Sequence[PackedConfig]) -> None:
    """Write ``configs`` as the file pair ``prefix.ConfigPack`` / ``prefix.ConfigPackIndex``."""
    prefix = Path(prefix)
    if not prefix.parent.is_dir():
        raise FileNotFoundError(f"directory {prefix.parent} does not exist")
    if len(configs) == 0:
        raise ValueError("no configurations to write (an empty pack is refused)")
    dims = [_check_config(c) for c in configs]  # validate everything before touching a file
    pack_path, index_path = _pair(prefix)
    offsets = np.empty(len(configs) + 1, dtype="<i8")
    offsets[0] = len(configs)
    with open(pack_path, "wb") as fh:
        at = 0
        for k, (cfg, dim) in enumerate(zip(configs, dims)):
            blob = _codec(dim).encode(cfg)
            offsets[k + 1] = at
            fh.write(blob)
            at += len(blob)
    offsets.tofile(index_path)


def _unit_basis(basis: np.ndarray) -> np.ndarray:
    dim = basis.shape[0]
    det = float(np.linalg.det(basis))
    if not det > 0.0:
        raise ValueError(f"basis determinant {det} is not positive")
    return basis * det ** (-1.0 / dim)


def stack_fractional(configs: Sequence[PackedConfig]) -> Tuple[np.ndarray, np.ndarray]:
    """Stack equal-shape records: ((B, N, d) fractional positions, (d, d) unit-volume basis).

    Positions are returned as stored; the basis is scaled to unit volume and must
    agree across the records to 1e-12.
    """
    if len(configs) == 0:
        raise ValueError("no configurations to stack")
    rels = [np.asarray(c.relative) for c in configs]
    bases = [np.asarray(c.basis) for c in configs]
    if any(a.dtype != np.float64 for a in rels + bases):
        raise TypeError("positions and basis must be float64")
    shape = rels[0].shape
    odd = next((r.shape for r in rels if r.shape != shape), None)
    if odd is not None:
        raise ValueError(f"configurations of shapes {shape} and {odd} cannot be stacked")
    dim = shape[1]
    odd_basis = next((b.shape for b in bases if b.shape != (dim, dim)), None)
    if odd_basis is not None:
        raise ValueError(f"a {dim}-d configuration needs a ({dim}, {dim}) basis, got {odd_basis}")
    units = [_unit_basis(b) for b in bases]
    ref = units[0]
    if not all(np.allclose(u, ref, rtol=_UNIT_RTOL, atol=_UNIT_RTOL) for u in units[1:]):
        raise ValueError("the configurations' bases differ (after scaling to unit volume)")
    return np.stack(rels), ref

