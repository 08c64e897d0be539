"""Autotuner for the device relax: launch layout and tile knobs chosen by measurement.

No reference counterpart: the reference leaves tiling to XLA and sizes its k-chunks
with a fixed memory heuristic (``auto_chunk_size``, src/xla_cco/energy.py:86-101).

By default (``exact=True``) every knob the tuner touches leaves every lane's bits
unchanged, so a tuned relax returns exactly what an untuned one returns; only the
time differs:

* ``rho_chunk`` -- particles whose phase tables one rho workgroup stages in LDS at a
  time (16, 32 or 64); each thread sums its particles in the same order for any
  chunk, while shorter chunks let more workgroups share a CU;
* ``shards`` -- lanes split into 1 to 4 shards on separate streams, so one shard's
  HBM-bound L-BFGS update overlaps another's evaluation (kernels index lanes, so a
  lane's arithmetic does not depend on its shard);
* ``graph`` -- evaluation steps replayed from a HIP graph per launch (0 or 8).

``exact=False`` also tunes two knobs that regroup sums, so a lane's bits change at
rounding level while every lane still stops on the same tests:

* ``rho_split`` -- particles per density split (the default, or 2x, 4x, 8x it):
  the split partial sums are added in another grouping (density and energy
  differences of a few ulp);
* ``advance`` -- the L-BFGS direction kernel (global-memory or register two-loop,
  Gram form): the same direction from different summation orders.  Along a
  trajectory the energies agree to about 1e-9 relative; a lane can end on a
  different evaluation count (a line search accepting at another trial step).

``tune_relax`` times the first ``probe_steps`` evaluation steps of the actual batch
for each candidate (best of two; the advance variants after ``memory`` steps, when
the history is full), picks the fastest, and records it under a key of the
workload's shape class: method, d, N, the k-count M, the lane count rounded up to a
power of two, the precision (with the split-precision pair storage), the L-BFGS memory
and the exactness.  Choices live in
memory and, when ``XLA_CCO_TUNING_CACHE`` names a file, in that JSON file across
processes.
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

OPT_ADVANCE_VARIANT = 1  # CCO_OPT_ADVANCE_VARIANT
OPT_SHARDS = 2  # CCO_OPT_SHARDS
OPT_GRAPH = 3  # CCO_OPT_GRAPH
EOPT_RHO_CHUNK = 1  # CCO_EOPT_RHO_CHUNK
EOPT_RHO_SPLIT = 4  # CCO_EOPT_RHO_SPLIT
SPLIT_UNIT = 128  # rho split lengths are multiples of this
CACHE_ENV = "XLA_CCO_TUNING_CACHE"


@dataclasses.dataclass(frozen=True)
class TuneChoice:
    """Knob values of one workload class (0 shards, rho_split or advance: the
    library's own default)."""

    rho_chunk: int
    shards: int
    graph: int
    probe_ms: float = 0.0
    rho_split: int = 0
    advance: int = 0


_lock = threading.Lock()
_choices: Dict[str, TuneChoice] = {}
_loaded_from: Optional[str] = None


def _pow2(x: int) -> int:
    return 1 << max(0, int(x) - 1).bit_length()


def workload_key(energy: StealthyEnergy, b: int, n: int, d: int, settings: Settings,
                 exact: bool = True) -> str:
    """The shape class a choice is kept under."""
    method = "descent" if isinstance(settings, DescentSettings) else "lbfgs"
    precision = getattr(settings, "precision", "fp64")
    if precision == "mixed" and getattr(settings, "split_history", "fp64") == "fp32":
        precision = "mixed-fp32pairs"  # a different advance stream, timed apart
    memory = getattr(settings, "memory", 0)
    tail = "" if exact else "/rounding"
    return f"{method}/d{d}/n{n}/m{energy.m}/b{_pow2(b)}/{precision}/mem{memory}{tail}"


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


def default_split(n: int) -> int:
    """The library's default particles per rho split for ``n`` particles (128 up to
    N = 2048, else the multiple of 128 that gives at most 16 splits)."""
    if n <= 16 * SPLIT_UNIT:
        return SPLIT_UNIT
    return -(-n // (16 * SPLIT_UNIT)) * SPLIT_UNIT


def split_candidates(n: int) -> List[int]:
    """Split lengths worth timing: 0 (the default), then 2x, 4x and 8x the default
    while each still makes fewer splits than the one before."""
    base = default_split(n)
    out = [0]
    last = -(-n // base)
    for f in (2, 4, 8):
        ns = -(-n // (f * base))
        if ns >= last:
            break
        out.append(f * base)
        last = ns
    return out


def advance_candidates(session) -> List[int]:
    """L-BFGS advance variants the session accepts (none for descent)."""
    if not isinstance(session.settings, LBFGSSettings):
        return []
    L = _native.lib()
    ok = [v for v in (1, 2, 3) if L.cco_relax_set_option(session.handle, OPT_ADVANCE_VARIANT, v) == 0]
    _native.check(L.cco_relax_set_option(session.handle, OPT_ADVANCE_VARIANT, 0))
    return ok


def energy_options(energy: StealthyEnergy) -> Dict[int, int]:
    """The energy's tunable option values, to hand back to ``restore_energy``."""
    L = _native.lib()
    out = {}
    for key in (EOPT_RHO_CHUNK, EOPT_RHO_SPLIT):
        v = C.c_int64(0)
        _native.check(L.cco_energy_get_option(energy.handle(), key, C.byref(v)))
        out[key] = int(v.value)
    return out


def restore_energy(energy: StealthyEnergy, saved: Dict[int, int]) -> None:
    L = _native.lib()
    for key, v in saved.items():
        _native.check(L.cco_energy_set_option(energy.handle(), key, v))


def apply(session, choice: TuneChoice) -> None:
    """Set a session's (and its energy's) knobs; call before ``init``.  An
    ``advance`` of 0 leaves the session's variant as it is."""
    L = _native.lib()
    _native.check(L.cco_energy_set_option(session.energy.handle(), EOPT_RHO_CHUNK, choice.rho_chunk))
    _native.check(L.cco_energy_set_option(session.energy.handle(), EOPT_RHO_SPLIT, choice.rho_split))
    _native.check(L.cco_relax_set_option(session.handle, OPT_SHARDS, choice.shards))
    _native.check(L.cco_relax_set_option(session.handle, OPT_GRAPH, choice.graph))
    if choice.advance:
        _native.check(L.cco_relax_set_option(session.handle, OPT_ADVANCE_VARIANT, choice.advance))


def _probe(session, s0: np.ndarray, steps: int, skip: int = 0) -> float:
    """Seconds for ``steps`` evaluation steps of a fresh init, after ``skip``
    untimed ones (device synchronised)."""
    L = _native.lib()
    ctx = _native.context(session.energy.device)
    session.init(s0)
    na = C.c_int(0)
    if skip:
        _native.check(L.cco_relax_step(session.handle, int(skip), C.byref(na)))
    ctx.synchronize()
    t0 = time.perf_counter()
    _native.check(L.cco_relax_step(session.handle, int(steps), C.byref(na)))
    ctx.synchronize()
    return time.perf_counter() - t0


def tune_relax(energy: StealthyEnergy, s0: np.ndarray, settings: Settings, probe_steps: int = 96,
               reps: int = 2, store: bool = True, exact: bool = True) -> TuneChoice:
    """Time every candidate on the first ``probe_steps`` steps of relaxing ``s0``;
    ``exact=False`` then also times the rho split lengths and the L-BFGS advance
    variants with the best layout.  The energy's options are left as they were."""
    from xla_cco.minimize import RelaxSession  # the session class lives in the package root

    if not isinstance(energy, StealthyEnergy):
        raise TypeError("the autotuner times the built-in StealthyEnergy kernels")
    s0 = np.ascontiguousarray(s0)
    b, n, d = s0.shape
    best: Optional[TuneChoice] = None
    saved = energy_options(energy)
    sess = RelaxSession(energy, b, n, d, settings)

    def trial_time(trial: TuneChoice, skip: int = 0) -> float:
        apply(sess, trial)
        return min(_probe(sess, s0, probe_steps, skip) for _ in range(reps))

    try:
        for chunk, shards, graph in candidates(energy, b):
            trial = TuneChoice(chunk, shards, graph, rho_split=saved[EOPT_RHO_SPLIT])
            t = trial_time(trial)
            if best is None or t < best.probe_ms * 1e-3:
                best = dataclasses.replace(trial, probe_ms=t * 1e3)
        assert best is not None
        if not exact:
            timed = {sp: trial_time(dataclasses.replace(best, rho_split=sp)) for sp in split_candidates(n)}
            sp = min(timed, key=timed.get)
            best = dataclasses.replace(best, rho_split=sp, probe_ms=timed[sp] * 1e3)
            variants = advance_candidates(sess)
            if len(variants) > 1:
                skip = int(getattr(settings, "memory", 0))  # the history full
                timed = {v: trial_time(dataclasses.replace(best, advance=v), skip) for v in variants}
                v = min(timed, key=timed.get)
                best = dataclasses.replace(best, advance=v, probe_ms=timed[v] * 1e3)
    finally:
        sess.close()
        restore_energy(energy, saved)
    if store:
        key = workload_key(energy, b, n, d, settings, exact)
        with _lock:
            _load()
            _choices[key] = best
            _store()
    return best


def choice_for(energy: StealthyEnergy, s0: np.ndarray, settings: Settings,
               exact: bool = True) -> TuneChoice:
    """The recorded choice of this workload class, tuning on ``s0`` the first time."""
    b, n, d = np.shape(s0)
    hit = cached_choice(workload_key(energy, b, n, d, settings, exact))
    return hit if hit is not None else tune_relax(energy, s0, settings, exact=exact)
