"""The allocation-free evaluation against the allocating one, through the C ABI.

cco_value_grad_density_dev_noalloc runs the kernels of
cco_value_grad_density_dev with its partial sums in the library's fixed
per-GPU scratch array.  Two calls in a row, the second smaller (fewer lanes,
fewer particles), must give bit-identical Phi, gradients and densities to the
allocating entry point on the same inputs; a batch beyond the scratch size and
a split-precision energy are refused before any work.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from xla_cco import _native
from xla_cco.box import cubic_box
from xla_cco.energy import make_stealthy_energy
from xla_cco.kvectors import enumerate_kvectors, k_outer_from_ka, ka_from_chi

pytestmark = pytest.mark.gpu


def _energy(n: int, d: int, chi: float):
    box = cubic_box(n, d)
    ks = enumerate_kvectors(box.basis, 0.0, k_outer_from_ka(ka_from_chi(chi, 0.05, d), 0.05, d))
    return make_stealthy_energy(box, ks)


def _eval(fn, e, s: np.ndarray, with_grad: bool, with_rho: bool):
    ctx = _native.context(0)
    b, n, d = s.shape
    mi = C.c_int()
    mp = C.c_int()
    di = C.c_int()
    vol = C.c_double()
    nm = (C.c_int * 3)()
    _native.check(_native.lib().cco_energy_info(e.handle(), C.byref(mi), C.byref(mp), C.byref(di), C.byref(vol), nm))
    m_pad = mp.value
    ds = _native.DeviceArray.from_host(ctx, s)
    phi = _native.DeviceArray(ctx, (b,))
    grad = _native.DeviceArray(ctx, (b, n, d)) if with_grad else None
    re = _native.DeviceArray(ctx, (b, m_pad)) if with_rho else None
    im = _native.DeviceArray(ctx, (b, m_pad)) if with_rho else None
    rc = fn(e.handle(), ds.ptr, b, n, phi.ptr, grad.ptr if grad else None,
            re.ptr if re else None, im.ptr if im else None)
    out = {"rc": rc}
    if rc == 0:
        out["phi"] = phi.to_host()
        if grad:
            out["grad"] = grad.to_host()
        if re:
            out["re"], out["im"] = re.to_host(), im.to_host()
    for a in (ds, phi, grad, re, im):
        if a is not None:
            a.free()
    return out


@pytest.mark.parametrize("d,chi", [(3, 0.3), (2, 0.45)])
def test_noalloc_equals_allocating_entry_point(gpu_ctx, d: int, chi: float) -> None:
    L = _native.lib()
    e = _energy(400, d, chi)
    rng = np.random.default_rng(7 + d)
    cases = [(24, 400, True, True), (5, 150, False, False), (3, 150, True, False)]
    for b, n, with_grad, with_rho in cases:
        s = rng.random((b, n, d))
        print(f"noalloc d={d} B={b} N={n} grad={with_grad} rho={with_rho}", flush=True)
        got = _eval(L.cco_value_grad_density_dev_noalloc, e, s, with_grad, with_rho)
        want = _eval(L.cco_value_grad_density_dev, e, s, with_grad, with_rho)
        assert got["rc"] == 0 and want["rc"] == 0, _native.lib().cco_last_error()
        for key in want:
            assert np.array_equal(got[key], want[key]), key


def test_noalloc_refuses_oversized_batch_and_split_precision(gpu_ctx) -> None:
    L = _native.lib()
    cap = int(L.cco_noalloc_scratch_bytes())
    assert cap == 64 << 20
    e = _energy(2000, 3, 0.45)
    # 16 B x 16 splits x ~5000 box entries per lane: 1.3 MB a lane, so 256 lanes exceed 64 MiB
    s = np.random.default_rng(1).random((256, 2000, 3))
    out = _eval(L.cco_value_grad_density_dev_noalloc, e, s, True, False)
    assert out["rc"] == _native.CCO_ERR_ARG
    assert b"scratch" in L.cco_last_error()
    small = _energy(200, 2, 0.3)
    if small.set_precision("mixed"):
        out = _eval(L.cco_value_grad_density_dev_noalloc, small, np.random.default_rng(2).random((2, 200, 2)), True, False)
        assert out["rc"] == _native.CCO_ERR_STATE
        small.set_precision("fp64")
        out = _eval(L.cco_value_grad_density_dev_noalloc, small, np.random.default_rng(2).random((2, 200, 2)), True, False)
        assert out["rc"] == 0
