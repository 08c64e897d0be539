"""The C-ABI library loads and exports every symbol include/cco.h declares.

CPU only: no compute call is made (there may be no GPU here).
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "cco.h"


def _declared() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cco_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary() -> None:
    names = _declared()
    for must in (
        "cco_energy_create",
        "cco_value_grad_density",
        "cco_value_grad_density_dev",
        "cco_relax_create",
        "cco_relax_run",
        "cco_relax_result",
    ):
        assert must in names


def test_library_exports_every_declared_symbol() -> None:
    from xla_cco import _native

    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # the ctypes table covers the header exactly
    assert sorted(_native.SIGNATURES) == _declared()


def test_library_is_gfx950_only() -> None:
    """The shared object carries a gfx950 code object and no other GPU target."""
    from xla_cco import _native

    blob = _native.LIB_PATH.read_bytes()
    assert b"gfx950" in blob
    for other in (b"gfx942", b"gfx90a", b"gfx1100", b"gfx908"):
        assert other not in blob


def test_loader_reports_missing_library(monkeypatch: pytest.MonkeyPatch, tmp_path: Path) -> None:
    from xla_cco import _native

    monkeypatch.setattr(_native, "LIB_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(_native.NativeLibraryError):
        _native.lib()


def test_version_and_device_count_callable() -> None:
    from xla_cco import _native

    L = _native.lib()
    assert L.cco_version() >= 100
    c = ctypes.c_int(-1)
    assert L.cco_device_count(ctypes.byref(c)) == 0
    assert c.value >= 0


def test_profile_stage_count_agrees() -> None:
    """cco_relax_profile fills CCO_PROFILE_STAGES entries: the ctypes constant and
    bench.py's stage names must size their arrays alike."""
    import sys

    from xla_cco import _native

    m = re.search(r"#define CCO_PROFILE_STAGES (\d+)", HEADER.read_text())
    assert m is not None
    sys.path.insert(0, str(ROOT))
    import bench

    assert int(m.group(1)) == _native.CCO_PROFILE_STAGES == len(bench.STAGES)
