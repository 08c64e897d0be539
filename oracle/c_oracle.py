"""ctypes wrapper of the C oracle (oracle/_build/libcco_oracle.so) -- TEST INFRASTRUCTURE.

Used by tests/ and by bench.py's cpu_baseline leg only.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libcco_oracle.so"
_lib: Optional[C.CDLL] = None

_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)


def load(path: Path) -> C.CDLL:
    """Open a build of cco_oracle.c and declare its entry points."""
    h = C.CDLL(str(path))
    h.oracle_energy_grad_batch.argtypes = [
        _PD, C.c_int, C.c_int, _PD, _PD, C.c_int, C.c_int, C.c_double, _PD, _PD, C.c_int, C.c_int
    ]
    h.oracle_relax.argtypes = [
        _PD, C.c_int, C.c_int, _PD, _PD, C.c_int, C.c_int, C.c_double, C.c_int,
        _PD, _PI64, _PD, _PI64, C.c_int64, _PD, _PD, _PI64, _PI64, _PI64, C.c_int, C.c_int,
    ]
    return h


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
        _lib = load(LIB)
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_PD)


def energy_grad_batch(
    s: np.ndarray, n: np.ndarray, w: np.ndarray, volume: float, threads: int = 1, phases: int = 0
) -> Tuple[np.ndarray, np.ndarray]:
    """Phi and dPhi/ds of (B, N, d) configurations.

    ``phases`` 0 evaluates cos/sin of 2 pi n.s per (k, j) as the reference does;
    1 multiplies per-axis power tables (the accuracy class of the HIP kernels).
    """
    s = np.ascontiguousarray(s, dtype=np.float64)
    nf = np.ascontiguousarray(n, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    b, nn, d = s.shape
    phi = np.empty(b)
    g = np.empty_like(s)
    lib().oracle_energy_grad_batch(_d(s), b, nn, _d(nf), _d(w), nf.shape[0], d, volume, _d(phi), _d(g),
                                   threads, int(phases))
    return phi, g


def relax(
    s0: np.ndarray,
    n: np.ndarray,
    w: np.ndarray,
    volume: float,
    settings: object,
    eval_cap: int = 2**62,
    threads: int = 1,
    handle: Optional[C.CDLL] = None,
    phases: int = 0,
):
    """Batched relax with an oracle.DescentSettings / oracle.LBFGSSettings.

    ``phases`` selects the energy's phase evaluation as in energy_grad_batch.

    ``handle`` selects another build of cco_oracle.c (default: the Makefile's).
    """
    from oracle import oracle as O

    s0 = np.ascontiguousarray(s0, dtype=np.float64)
    nf = np.ascontiguousarray(n, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    b, nn, d = s0.shape
    dcfg = np.zeros(8)
    dint = np.zeros(4, dtype=np.int64)
    lcfg = np.zeros(12)
    lint = np.zeros(8, dtype=np.int64)
    if isinstance(settings, O.DescentSettings):
        method = 0
        dcfg[:6] = [settings.step_init, settings.step_min, settings.shrink, settings.grow,
                    settings.rel_tol, settings.energy_tol]
        dint[0] = settings.max_evals
    elif isinstance(settings, O.LBFGSSettings):
        method = 1
        lcfg[: len(settings.rescales)] = settings.rescales
        lcfg[8], lcfg[9] = settings.energy_tol, settings.grad_tol
        lint[:5] = [len(settings.rescales), settings.memory, settings.max_evals_per_pass,
                    settings.stagnation_patience, settings.max_linesearch_steps]
    else:
        raise TypeError("unsupported settings")
    s = np.empty_like(s0)
    e = np.empty(b)
    ne = np.empty(b, dtype=np.int64)
    ni = np.empty(b, dtype=np.int64)
    st = np.empty(b, dtype=np.int64)
    (handle or lib()).oracle_relax(
        _d(s0), b, nn, _d(nf), _d(w), nf.shape[0], d, volume, method,
        _d(dcfg), dint.ctypes.data_as(_PI64), _d(lcfg), lint.ctypes.data_as(_PI64), int(eval_cap),
        _d(s), _d(e), ne.ctypes.data_as(_PI64), ni.ctypes.data_as(_PI64), st.ctypes.data_as(_PI64),
        threads, int(phases),
    )
    return s, e, ne, ni, st
