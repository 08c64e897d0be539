"""ctypes binding of the gfx950 HIP library (``_native/libcco_hip.so``).

The library is the only compute path: there is no CPU fallback.  Loading
fails with ``NativeLibraryError`` when the library has not been built, and
creating a context fails with ``RuntimeError`` when no GPU is visible
(the reference's fail-fast device guard, src/xla_cco/_device.py:137).
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path
from typing import Dict, Optional

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_native" / "libcco_hip.so"

CCO_OK = 0
CCO_ERR_ARG = 1
CCO_ERR_DTYPE = 2
CCO_ERR_DEVICE = 3
CCO_ERR_BOUND = 4
CCO_ERR_STATE = 5
CCO_PROFILE_STAGES = 10  # include/cco.h: entries of cco_relax_profile's arrays

METHOD_DESCENT = 0
METHOD_LBFGS = 1


class NativeLibraryError(ImportError):
    """The HIP library is missing or does not export the expected ABI."""


class FlowSettingsC(C.Structure):
    _fields_ = [
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("dt_init", C.c_double),
        ("dt_min", C.c_double),
        ("dt_max", C.c_double),
        ("t_max", C.c_double),
        ("energy_tol", C.c_double),
        ("grad_tol", C.c_double),
        ("max_steps", C.c_int64),
        ("arclength", C.c_int32),
        ("record_every", C.c_int32),
        ("max_records", C.c_int32),
    ]


class DescentSettingsC(C.Structure):
    _fields_ = [
        ("step_init", C.c_double),
        ("step_min", C.c_double),
        ("shrink", C.c_double),
        ("grow", C.c_double),
        ("rel_tol", C.c_double),
        ("energy_tol", C.c_double),
        ("max_evals", C.c_int64),
        ("record_every", C.c_int32),
        ("max_records", C.c_int32),
    ]


class LBFGSSettingsC(C.Structure):
    _fields_ = [
        ("memory", C.c_int32),
        ("n_rescales", C.c_int32),
        ("rescales", C.c_double * 8),
        ("max_evals_per_pass", C.c_int64),
        ("energy_tol", C.c_double),
        ("grad_tol", C.c_double),
        ("stagnation_patience", C.c_int32),
        ("max_linesearch_steps", C.c_int32),
        ("record_every", C.c_int32),
        ("max_records", C.c_int32),
    ]


class MDSettingsC(C.Structure):
    _fields_ = [
        ("andersen_p", C.c_double),
        ("gamma", C.c_double),
        ("phase_floor", C.c_double),
        ("scheme", C.c_int32),
        ("andersen_mode", C.c_int32),
        ("wind_every", C.c_int32),
        ("record_rho_every", C.c_int32),
        ("max_rho_records", C.c_int32),
        ("n_record_steps", C.c_int32),
    ]


_PDn = C.POINTER(C.c_double)
_PI64n = C.POINTER(C.c_int64)


class MDStateIOC(C.Structure):
    _fields_ = [
        ("r", _PDn),
        ("v", _PDn),
        ("f", _PDn),
        ("lane_f64", _PDn),
        ("lane_i64", _PI64n),
        ("key", C.POINTER(C.c_uint32)),
        ("rho_prev", _PDn),
        ("wind_abs", _PDn),
        ("wind_net", _PDn),
        ("n_valid", _PI64n),
        ("sum_s", _PDn),
        ("snap_step", _PI64n),
        ("snap_time", _PDn),
        ("snap_phi", _PDn),
        ("snap_kinetic", _PDn),
        ("snap_wind_abs", _PDn),
        ("snap_wind_net", _PDn),
        ("snap_rho", _PDn),
        ("snap_r", _PDn),
        ("trace_rho", _PDn),
        ("trace_time", _PDn),
    ]


_P = C.c_void_p
_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)
_PI32 = C.POINTER(C.c_int32)
_PI = C.POINTER(C.c_int)

# name -> (restype, argtypes); every symbol declared in include/cco.h
SIGNATURES = {
    "cco_version": (C.c_int, []),
    "cco_last_error": (C.c_char_p, []),
    "cco_device_count": (C.c_int, [_PI]),
    "cco_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "cco_ctx_destroy": (C.c_int, [_P]),
    "cco_ctx_synchronize": (C.c_int, [_P]),
    "cco_dev_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "cco_dev_free": (C.c_int, [_P, _P]),
    "cco_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "cco_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "cco_energy_create": (
        C.c_int,
        [_P, _PI64, _PD, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(_P)],
    ),
    "cco_energy_destroy": (C.c_int, [_P]),
    "cco_energy_info": (C.c_int, [_P, _PI, _PI, _PI, _PD, _PI]),
    "cco_value_grad_density": (C.c_int, [_P, _PD, C.c_int, C.c_int, _PD, _PD, _PD, _PD]),
    "cco_value_grad_density_dev": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
    "cco_value_grad_density_dev_noalloc": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
    "cco_noalloc_scratch_bytes": (C.c_size_t, []),
    "cco_energy_set_option": (C.c_int, [_P, C.c_int, C.c_int64]),
    "cco_energy_get_option": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int64)]),
    "cco_relax_create": (
        C.c_int,
        [
            _P,
            _P,
            C.c_int,
            C.c_int,
            C.c_int,
            C.c_int,
            C.POINTER(DescentSettingsC),
            C.POINTER(LBFGSSettingsC),
            C.POINTER(_P),
        ],
    ),
    "cco_relax_destroy": (C.c_int, [_P]),
    "cco_relax_init": (C.c_int, [_P, _PD]),
    "cco_relax_init_dev": (C.c_int, [_P, _P]),
    "cco_relax_run": (C.c_int, [_P, C.c_int64, C.c_int]),
    "cco_relax_step": (C.c_int, [_P, C.c_int64, _PI]),
    "cco_relax_pending": (C.c_int, [_P, _PI, _PI32, _PD]),
    "cco_relax_submit": (C.c_int, [_P, _PD, _PD]),
    "cco_relax_result": (C.c_int, [_P, _PD, _PD, _PD, _PI64, _PI64, _PI64]),
    "cco_relax_result_dev": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "cco_relax_trajectory": (C.c_int, [_P, _PI64, _PI64, _PD, _PD, _PD]),
    "cco_relax_stats": (C.c_int, [_P, _PI64, _PI64]),
    "cco_relax_counters": (C.c_int, [_P, _PI64]),
    "cco_relax_set_option": (C.c_int, [_P, C.c_int, C.c_int64]),
    "cco_relax_get_option": (C.c_int, [_P, C.c_int, _PI64]),
    "cco_relax_uneven_create": (
        C.c_int,
        [_P, C.c_int, C.POINTER(_P), C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int,
         C.POINTER(DescentSettingsC), C.POINTER(LBFGSSettingsC), C.POINTER(_P)],
    ),
    "cco_relax_set_precision": (C.c_int, [_P, C.c_int, C.c_double]),
    "cco_relax_precision_state": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cco_energy_set_precision": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "cco_hessian": (C.c_int, [_P, _PD, C.c_int, C.c_int, _PD]),
    "cco_hessian_dev": (C.c_int, [_P, _P, C.c_int, C.c_int, _P]),
    "cco_flow_create": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P, _PD, C.POINTER(_P)]),
    "cco_flow_set_method": (C.c_int, [_P, C.c_int]),
    "cco_flow_init": (C.c_int, [_P, _PD, _PD]),
    "cco_flow_result": (C.c_int, [_P, _PD, _PD, _PD, _PD, _PD, _PI64, _PI64, _PI64]),
    "cco_flow_trajectory": (C.c_int, [_P, _PI64, _PD, _PD, _PD, _PD]),
    "cco_md_create": (
        C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _PD, C.c_double, C.c_double, C.POINTER(_P)]
    ),
    "cco_md_run": (C.c_int, [_P, C.c_int64, C.c_int, _PI]),
    "cco_md_configure": (
        C.c_int, [_P, C.POINTER(MDSettingsC), C.c_int, C.c_int, _PD, _PI64]
    ),
    "cco_md_put": (C.c_int, [_P, C.POINTER(MDStateIOC)]),
    "cco_md_get": (C.c_int, [_P, C.POINTER(MDStateIOC)]),
    "cco_sk_moments": (
        C.c_int, [_P, _PD, C.c_int, C.c_int, C.c_int, _PI64, _PI32, C.c_int, _PD, _PD]
    ),
    "cco_pair_stats": (
        C.c_int, [_P, _PD, C.c_int, C.c_int, C.c_int, _PD, _PD, C.c_int, _PI64, _PD]
    ),
    "cco_relax_set_profiling": (C.c_int, [_P, C.c_int]),
    "cco_relax_profile": (C.c_int, [_P, _PD, _PI64]),
    "cco_energy_profile": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _PD]),
    "cco_copy_dev": (C.c_int, [_P, _P, _P, C.c_size_t, _PD]),
}

_lib: Optional[C.CDLL] = None
_lock = threading.Lock()


def lib() -> C.CDLL:
    """Load the HIP library once (raises NativeLibraryError when missing)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise NativeLibraryError(
                f"HIP library not built: {LIB_PATH} is missing "
                "(run `python xla-cco_amd/build_native.py`)"
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name, None)
            if fn is None:
                raise NativeLibraryError(f"{LIB_PATH} does not export {name}")
            fn.restype = res
            fn.argtypes = args
        _lib = handle
        return handle


def check(rc: int) -> None:
    """Map a C-ABI return code to the reference's exception types."""
    if rc == CCO_OK:
        return
    msg = lib().cco_last_error().decode(errors="replace")
    if rc == CCO_ERR_ARG:
        raise ValueError(msg)
    if rc == CCO_ERR_DTYPE:
        raise TypeError(msg)
    raise RuntimeError(msg or f"cco error {rc}")


def dptr(a: np.ndarray) -> "C._Pointer[C.c_double]":
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PD)


def i64ptr(a: np.ndarray) -> "C._Pointer[C.c_int64]":
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PI64)


def u32ptr(a: np.ndarray) -> "C._Pointer[C.c_uint32]":
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Context:
    """One device and its stream (cco_ctx)."""

    def __init__(self, device: int = 0) -> None:
        L = lib()
        count = C.c_int(0)
        check(L.cco_device_count(C.byref(count)))
        if count.value == 0:
            raise RuntimeError("no GPU visible to the HIP runtime; the HIP path has no CPU fallback")
        h = _P()
        check(L.cco_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self) -> None:
        check(lib().cco_ctx_synchronize(self.handle))

    def __del__(self) -> None:  # pragma: no cover - interpreter teardown order
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.cco_ctx_destroy(self.handle)
        except Exception:
            pass


_contexts: Dict[int, Context] = {}


def context(device: int = 0) -> Context:
    """The process-wide context of ``device`` (created on first use)."""
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _contexts.setdefault(device, ctx)
            ctx = _contexts[device]
    return ctx


class DeviceArray:
    """A float64/int64 device buffer on a context (cco_dev_alloc), for the *_dev entry points."""

    def __init__(self, ctx: Context, shape, dtype=np.float64) -> None:
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _P()
        check(lib().cco_dev_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        check(lib().cco_memcpy_h2d(ctx.handle, d.ptr, a.ctypes.data_as(_P), d.nbytes))
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib().cco_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(_P), self.ptr, self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr is not None and _lib is not None:
            _lib.cco_dev_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.free()
        except Exception:
            pass
