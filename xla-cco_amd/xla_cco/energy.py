"""Stealthy energy on fractional positions, evaluated by the gfx950 HIP library.

Phi = (1/V) sum_k w_k |rho_k|^2, rho_k = sum_j exp(2 pi i n_k . s_j) over a
padded half-space k-set (src/xla_cco/energy.py:3-11).  ``value_and_grad``
and ``value_grad_density`` return the analytic gradient
dPhi/ds_j = (2/V) sum_k w_k (Im rho_k cos th_kj - Re rho_k sin th_kj) 2 pi n_k
(energy.py:195-225), which equals the C++ gradient -B F_j.

All arithmetic runs on the GPU (``cco_value_grad_density`` of include/cco.h);
this module only validates inputs, owns the k-set upload and converts
arrays.  Inputs may be NumPy arrays (host) and results come back as NumPy.
"""

from __future__ import annotations

import abc
import ctypes as C
from typing import Optional, Tuple

import numpy as np

from xla_cco import _native
from xla_cco._types import require_float64
from xla_cco.box import Box
from xla_cco.kvectors import KSet


class AbstractEnergy(abc.ABC):
    """Extension point for energies on fractional positions (energy.py:115-137).

    Subclasses implement ``__call__`` and ``value_and_grad``; ``relax`` calls
    only ``value_and_grad`` (on the host, once per pending lane) when the
    energy is not a ``StealthyEnergy``; the minimizer itself still runs on the
    GPU.
    """

    @abc.abstractmethod
    def __call__(self, s: np.ndarray) -> float:
        raise NotImplementedError

    def value_and_grad(self, s: np.ndarray) -> Tuple[float, np.ndarray]:
        raise NotImplementedError(
            f"{type(self).__name__} must implement value_and_grad (no autodiff backend)"
        )

    def hessian(self, s: np.ndarray) -> np.ndarray:
        """d^2 Phi / ds ds flattened row-major to (n d, n d) (energy.py:134-137)."""
        raise NotImplementedError(
            f"{type(self).__name__} must implement hessian (no autodiff backend)"
        )


class StealthyEnergy(AbstractEnergy):
    """Stealthy energy over a padded k-set (energy.py:140-184).

    Attributes mirror the reference: ``n_vectors`` (M, d) integer-valued
    float64 (zero rows are padding), ``weights`` (M,) (0 on padding),
    ``volume``; ``chunk_size`` and ``n_max`` are accepted for interface
    parity.  The device tiles over (k, j) itself, so neither changes the
    arithmetic (every evaluation uses per-axis phase recurrences, the
    reference's table path of energy.py:227-270, seeded by exact sincospi).
    """

    def __init__(
        self,
        n_vectors: np.ndarray,
        weights: np.ndarray,
        volume: float,
        chunk_size: Optional[int] = None,
        n_max: Optional[int] = None,
        device: int = 0,
    ) -> None:
        n_vectors = np.asarray(n_vectors)
        weights = np.asarray(weights)
        require_float64(n_vectors, "n_vectors")
        require_float64(weights, "weights")
        if n_vectors.ndim != 2 or n_vectors.shape[1] not in (1, 2, 3):
            raise ValueError(f"n_vectors must have shape (m, d), d in 1..3, got {n_vectors.shape}")
        if weights.shape != (n_vectors.shape[0],):
            raise ValueError("weights must have shape (m,)")
        if n_max is not None and n_max < 0:
            raise ValueError(f"n_max must be non-negative, got {n_max}")
        if n_max is not None and n_vectors.size and np.abs(n_vectors).max() > n_max:
            raise ValueError(
                f"n_max={n_max} is smaller than max|n|={np.abs(n_vectors).max():g}"
            )
        if np.any(n_vectors != np.round(n_vectors)):
            raise ValueError("n_vectors must be integer-valued")
        if chunk_size is not None and chunk_size < 1:
            raise ValueError("chunk_size must be >= 1")
        self.n_vectors = np.ascontiguousarray(n_vectors)
        self.weights = np.ascontiguousarray(weights)
        self.volume = float(volume)
        self.chunk_size = chunk_size
        self.n_max = n_max
        self.device = device
        self._handle: Optional[C.c_void_p] = None

    # ----- device handle -----
    @property
    def dim(self) -> int:
        return int(self.n_vectors.shape[1])

    @property
    def m(self) -> int:
        return int(self.n_vectors.shape[0])

    def handle(self) -> C.c_void_p:
        """Upload the k-set on first use (cco_energy_create)."""
        if self._handle is None:
            ctx = _native.context(self.device)
            n = np.ascontiguousarray(np.rint(self.n_vectors).astype(np.int64))
            h = C.c_void_p()
            _native.check(
                _native.lib().cco_energy_create(
                    ctx.handle,
                    _native.i64ptr(n),
                    _native.dptr(self.weights),
                    self.m,
                    self.dim,
                    self.volume,
                    1,
                    C.byref(h),
                )
            )
            self._handle = h
        return self._handle

    def __del__(self) -> None:  # pragma: no cover - interpreter teardown order
        try:
            if self._handle is not None and _native._lib is not None:
                _native._lib.cco_energy_destroy(self._handle)
        except Exception:
            pass

    def set_precision(self, precision: str) -> bool:
        """Select FP64 ("fp64", default) or split-precision ("mixed") evaluations.

        "mixed" runs rho and the forces as complex GEMMs on the FP16 matrix
        cores with hi/lo operand splits and FP32 accumulation (Phi, the density
        sums and the outputs stay FP64).  Returns whether the k-set allows it;
        when it does not, evaluations stay FP64.
        """
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        applied = C.c_int(0)
        _native.check(
            _native.lib().cco_energy_set_precision(
                self.handle(), 1 if precision == "mixed" else 0, C.byref(applied)
            )
        )
        return bool(applied.value)

    # ----- evaluation -----
    def _check(self, s: np.ndarray, batched: bool) -> np.ndarray:
        require_float64(s, "s")
        want = 3 if batched else 2
        if s.ndim != want or s.shape[-1] != self.dim:
            raise ValueError(f"s must have shape {'(b, n, d)' if batched else '(n, d)'} with d={self.dim}")
        return np.ascontiguousarray(s)

    def _eval(
        self, s: np.ndarray, grad: bool, density: bool
    ) -> Tuple[np.ndarray, Optional[np.ndarray], Optional[np.ndarray], Optional[np.ndarray]]:
        b, n = s.shape[0], s.shape[1]
        phi = np.empty(b)
        g = np.empty_like(s) if grad else None
        re = np.empty((b, self.m)) if density else None
        im = np.empty((b, self.m)) if density else None
        _native.check(
            _native.lib().cco_value_grad_density(
                self.handle(),
                _native.dptr(s),
                b,
                n,
                _native.dptr(phi),
                _native.dptr(g) if g is not None else None,
                _native.dptr(re) if re is not None else None,
                _native.dptr(im) if im is not None else None,
            )
        )
        return phi, g, re, im

    def __call__(self, s: np.ndarray) -> float:
        """Phi(s) (energy.py:186-193)."""
        s = self._check(np.asarray(s) if not hasattr(s, "dtype") else s, batched=False)
        phi, _, _, _ = self._eval(s[None], grad=False, density=False)
        return float(phi[0])

    def value_and_grad(self, s: np.ndarray) -> Tuple[float, np.ndarray]:
        """(Phi, dPhi/ds) (energy.py:195-204)."""
        s = self._check(s, batched=False)
        phi, g, _, _ = self._eval(s[None], grad=True, density=False)
        assert g is not None
        return float(phi[0]), g[0]

    def value_grad_density(
        self, s: np.ndarray
    ) -> Tuple[float, np.ndarray, np.ndarray, np.ndarray]:
        """(Phi, dPhi/ds, Re rho, Im rho) from one pass (energy.py:206-225).

        Padding rows (n = 0) return re = N, im = 0, as the reference does.
        """
        s = self._check(s, batched=False)
        phi, g, re, im = self._eval(s[None], grad=True, density=True)
        assert g is not None and re is not None and im is not None
        return float(phi[0]), g[0], re[0], im[0]

    def hessian(self, s: np.ndarray) -> np.ndarray:
        """Closed-form Hessian flattened row-major to (n d, n d) (energy.py:272-299).

        H[(j,a),(l,b)] = (2 (2 pi)^2 / V) sum_k w_k n_ka n_kb [cos th_kj cos th_kl
        + sin th_kj sin th_kl - delta_jl (Re rho_k cos th_kj + Im rho_k sin th_kj)],
        computed on the GPU (one FP64 GEMM per block pair).
        """
        s = self._check(s, batched=False)
        return self.hessian_batch(s[None])[0]

    def hessian_batch(self, s: np.ndarray) -> np.ndarray:
        """Hessians of a batch (b, n, d) -> (b, n d, n d)."""
        s = self._check(s, batched=True)
        b, n, d = s.shape
        out = np.empty((b, n * d, n * d))
        _native.check(
            _native.lib().cco_hessian(self.handle(), _native.dptr(s), b, n, _native.dptr(out))
        )
        return out

    # ----- batched forms (one launch for all configurations) -----
    def energy_batch(self, s: np.ndarray) -> np.ndarray:
        s = self._check(s, batched=True)
        return self._eval(s, grad=False, density=False)[0]

    def value_and_grad_batch(self, s: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        s = self._check(s, batched=True)
        phi, g, _, _ = self._eval(s, grad=True, density=False)
        assert g is not None
        return phi, g

    def value_grad_density_batch(
        self, s: np.ndarray
    ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        s = self._check(s, batched=True)
        phi, g, re, im = self._eval(s, grad=True, density=True)
        assert g is not None and re is not None and im is not None
        return phi, g, re, im


def collective_density(
    s: np.ndarray, n_vectors: np.ndarray, chunk_size: Optional[int] = None
) -> Tuple[np.ndarray, np.ndarray]:
    """(Re rho, Im rho) with rho_k = sum_j exp(2 pi i n_k . s_j) (energy.py:327-358).

    Evaluated by the HIP density pass over a unit-weight copy of the k rows.
    """
    require_float64(s, "s")
    require_float64(n_vectors, "n_vectors")
    s = np.asarray(s)
    n_vectors = np.asarray(n_vectors)
    if s.ndim != 2 or n_vectors.ndim != 2 or s.shape[1] != n_vectors.shape[1]:
        raise ValueError("dimension mismatch")
    e = StealthyEnergy(n_vectors, np.ones(n_vectors.shape[0]), 1.0, chunk_size=chunk_size)
    _, _, re, im = e._eval(np.ascontiguousarray(s)[None], grad=False, density=True)
    assert re is not None and im is not None
    return re[0], im[0]


def auto_chunk_size(
    n_particles: int, n_vectors: int, budget_bytes: int = 1 << 30, bytes_per_pair: int = 96
) -> Optional[int]:
    """A ``chunk_size`` whose (c, N) work arrays fit ``budget_bytes`` (energy.py:86-101).

    None when the whole (M, N) matrix fits.  Kept for interface parity: the
    GPU kernels tile (k, j) themselves, so ``chunk_size`` never changes the
    arithmetic here.
    """
    assert n_particles >= 1 and n_vectors >= 1, "sizes must be positive"
    assert budget_bytes >= 1 and bytes_per_pair >= 1, "budget must be positive"
    size = max(1, budget_bytes // (n_particles * bytes_per_pair))
    return None if size >= n_vectors else size


def make_stealthy_energy(
    box: Box, kset: KSet, chunk_size: Optional[int] = None, pad_to: int = 64
) -> StealthyEnergy:
    """StealthyEnergy from a box and a k-set, rows padded to ``pad_to`` (energy.py:361-414).

    Raises ValueError on an empty k-set.
    """
    require_float64(np.asarray(box.basis), "box.basis")
    assert pad_to >= 1, "pad_to must be >= 1"
    assert chunk_size is None or chunk_size >= 1, "chunk_size must be >= 1"
    m = int(kset.n.shape[0])
    if m == 0:
        raise ValueError("empty k-set: no k-vectors in the shell")
    total = -(-m // pad_to) * pad_to
    n_vec = np.zeros((total, kset.n.shape[1]), dtype=np.float64)
    n_vec[:m] = kset.n
    wts = np.zeros(total, dtype=np.float64)
    wts[:m] = kset.weights
    return StealthyEnergy(n_vec, wts, box.volume, chunk_size=chunk_size, n_max=None)
