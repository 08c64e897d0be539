"""Autotuner for the device relax: launch layout and tile knobs chosen by measurement.

No reference counterpart: the reference leaves tiling to XLA and sizes its k-chunks
with a fixed memory heuristic (``auto_chunk_size``, src/xla_cco/energy.py:86-101).
Here every knob the tuner touches leaves every lane's bits unchanged, so a tuned
relax returns exactly what an untuned one returns; only the time differs:

* ``rho_chunk`` -- particles whose phase tables one rho workgroup stages in LDS at a
  time (16, 32 or 64); each thread sums its particles in the same order for any
  chunk, while shorter chunks let more workgroups share a CU;
* ``shards`` -- lanes split into 1 to 4 shards on separate streams, so one shard's
  HBM-bound L-BFGS update overlaps another's evaluation (kernels index lanes, so a
  lane's arithmetic does not depend on its shard);
* ``graph`` -- evaluation steps replayed from a HIP graph per launch (0 or 8).

``tune_relax`` times the first ``probe_steps`` evaluation steps of the actual batch
for each candidate (best of two), picks the fastest, and records it under a key of
the workload's shape class: method, d, N, the k-count M, the lane count rounded up
to a power of two, the precision and the L-BFGS memory.  Choices live in memory and,
when ``XLA_CCO_TUNING_CACHE`` names a file, in that JSON file across processes.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import itertools
import json
import os
import threading
import time
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Tuple, Union

import numpy as np

from xla_cco import _native
from xla_cco.energy import StealthyEnergy
from xla_cco.minimize.descent import DescentSettings
from xla_cco.minimize.lbfgs import LBFGSSettings

Settings = Union[DescentSettings, LBFGSSettings]

OPT_SHARDS = 2  # CCO_OPT_SHARDS
OPT_GRAPH = 3  # CCO_OPT_GRAPH
EOPT_RHO_CHUNK = 1  # CCO_EOPT_RHO_CHUNK
CACHE_ENV = "XLA_CCO_TUNING_CACHE"


@dataclasses.dataclass(frozen=True)
class TuneChoice:
    """Knob values of one workload class (0 shards: the library's own default)."""

    rho_chunk: int
    shards: int
    graph: int
    probe_ms: float = 0.0


_lock = threading.Lock()
_choices: Dict[str, TuneChoice] = {}
_loaded_from: Optional[str] = None


def _pow2(x: int) -> int:
    return 1 << max(0, int(x) - 1).bit_length()


def workload_key(energy: StealthyEnergy, b: int, n: int, d: int, settings: Settings) -> str:
    """The shape class a choice is kept under."""
    method = "descent" if isinstance(settings, DescentSettings) else "lbfgs"
    precision = getattr(settings, "precision", "fp64")
    memory = getattr(settings, "memory", 0)
    return f"{method}/d{d}/n{n}/m{energy.m}/b{_pow2(b)}/{precision}/mem{memory}"


def _cache_path() -> Optional[Path]:
    name = os.environ.get(CACHE_ENV)
    return Path(name) if name else None


def _load() -> None:
    global _loaded_from
    path = _cache_path()
    if path is None or _loaded_from == str(path):
        return
    _loaded_from = str(path)
    if not path.is_file():
        return
    try:
        data = json.loads(path.read_text())
    except (OSError, ValueError):
        return  # an unreadable cache is ignored and rewritten
    for key, val in data.items():
        try:
            _choices.setdefault(key, TuneChoice(**val))
        except TypeError:
            continue


def _store() -> None:
    path = _cache_path()
    if path is None:
        return
    tmp = Path(f"{path}.tmp")
    tmp.write_text(json.dumps({k: dataclasses.asdict(v) for k, v in sorted(_choices.items())}, indent=1))
    os.replace(tmp, path)


def cached_choice(key: str) -> Optional[TuneChoice]:
    with _lock:
        _load()
        return _choices.get(key)


def forget() -> None:
    """Drop the in-memory choices (the cache file, if any, is left alone)."""
    global _loaded_from
    with _lock:
        _choices.clear()
        _loaded_from = None


def _rho_chunks(energy: StealthyEnergy) -> List[int]:
    """Chunk lengths the energy's LDS budget allows, longest first."""
    L = _native.lib()
    cur = C.c_int64(0)
    _native.check(L.cco_energy_get_option(energy.handle(), EOPT_RHO_CHUNK, C.byref(cur)))
    ok = []
    for c in (64, 32, 16):
        if L.cco_energy_set_option(energy.handle(), EOPT_RHO_CHUNK, c) == 0:
            ok.append(c)
    _native.check(L.cco_energy_set_option(energy.handle(), EOPT_RHO_CHUNK, int(cur.value)))
    return ok


def candidates(energy: StealthyEnergy, b: int) -> Iterable[Tuple[int, int, int]]:
    """(rho_chunk, shards, graph) triples worth timing for ``b`` lanes."""
    shards = [s for s in (1, 2, 3, 4) if s == 1 or b >= 16 * s]
    return itertools.product(_rho_chunks(energy), shards, (0, 8))


def apply(session, choice: TuneChoice) -> None:
    """Set a session's (and its energy's) knobs; call before ``init``."""
    L = _native.lib()
    _native.check(L.cco_energy_set_option(session.energy.handle(), EOPT_RHO_CHUNK, choice.rho_chunk))
    _native.check(L.cco_relax_set_option(session.handle, OPT_SHARDS, choice.shards))
    _native.check(L.cco_relax_set_option(session.handle, OPT_GRAPH, choice.graph))


def _probe(session, s0: np.ndarray, steps: int) -> float:
    """Seconds for ``steps`` evaluation steps of a fresh init (device synchronised)."""
    L = _native.lib()
    ctx = _native.context(session.energy.device)
    session.init(s0)
    ctx.synchronize()
    t0 = time.perf_counter()
    na = C.c_int(0)
    _native.check(L.cco_relax_step(session.handle, int(steps), C.byref(na)))
    ctx.synchronize()
    return time.perf_counter() - t0


def tune_relax(energy: StealthyEnergy, s0: np.ndarray, settings: Settings, probe_steps: int = 96,
               reps: int = 2, store: bool = True) -> TuneChoice:
    """Time every candidate on the first ``probe_steps`` steps of relaxing ``s0``."""
    from xla_cco.minimize import RelaxSession  # the session class lives in the package root

    if not isinstance(energy, StealthyEnergy):
        raise TypeError("the autotuner times the built-in StealthyEnergy kernels")
    s0 = np.ascontiguousarray(s0)
    b, n, d = s0.shape
    best: Optional[TuneChoice] = None
    sess = RelaxSession(energy, b, n, d, settings)
    try:
        for chunk, shards, graph in candidates(energy, b):
            trial = TuneChoice(chunk, shards, graph)
            apply(sess, trial)
            t = min(_probe(sess, s0, probe_steps) for _ in range(reps))
            if best is None or t < best.probe_ms * 1e-3:
                best = dataclasses.replace(trial, probe_ms=t * 1e3)
    finally:
        sess.close()
    assert best is not None
    if store:
        key = workload_key(energy, b, n, d, settings)
        with _lock:
            _load()
            _choices[key] = best
            _store()
    return best


def choice_for(energy: StealthyEnergy, s0: np.ndarray, settings: Settings) -> TuneChoice:
    """The recorded choice of this workload class, tuning on ``s0`` the first time."""
    b, n, d = np.shape(s0)
    hit = cached_choice(workload_key(energy, b, n, d, settings))
    return hit if hit is not None else tune_relax(energy, s0, settings)
