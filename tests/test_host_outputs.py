"""Host outputs: ConfigPack format, tables, success mask, isotropic k-set, binning (CPU).

The pack reader/writer must reproduce the reference's bundled files byte for
byte (tests/golden/inputs are copies of tests/data/inputs of the reference).
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from xla_cco.analysis.structure import _bin_moments, isotropic_directions
from xla_cco.cli import relax as cli
from xla_cco.io.configpack import PackedConfig, read_configpack, stack_fractional, write_configpack
from xla_cco.io.tables import write_function_table

INPUTS = Path(__file__).resolve().parent / "golden" / "inputs"


@pytest.mark.parametrize("name", ["poisson_d1_n50", "poisson_d2_n64_pert", "poisson_d3_n100"])
def test_configpack_round_trip_is_byte_identical(tmp_path: Path, name: str) -> None:
    cfgs = read_configpack(INPUTS / name)
    assert len(cfgs) == 5
    write_configpack(tmp_path / "x", cfgs)
    for suf in (".ConfigPack", ".ConfigPackIndex"):
        assert (tmp_path / ("x" + suf)).read_bytes() == (INPUTS / (name + suf)).read_bytes()
    s, b = stack_fractional(cfgs)
    assert s.shape[0] == 5 and np.allclose(abs(np.linalg.det(b)), 1.0)


def test_configpack_rejects_bad_input(tmp_path: Path) -> None:
    with pytest.raises(FileNotFoundError):
        read_configpack(tmp_path / "missing")
    with pytest.raises(ValueError):
        write_configpack(tmp_path / "e", [])
    bad = PackedConfig(np.eye(2), np.zeros((3, 2)), 1.0, (b"ab",) * 3)
    with pytest.raises(ValueError):
        write_configpack(tmp_path / "b", [bad])
    (tmp_path / "t.ConfigPack").write_bytes(b"\x02\x00")
    np.array([1, 0], dtype="<i8").tofile(tmp_path / "t.ConfigPackIndex")
    with pytest.raises(ValueError):
        read_configpack(tmp_path / "t")


def test_function_table_format(tmp_path: Path) -> None:
    p = tmp_path / "t.txt"
    write_function_table(p, [np.array([1.0, 2.5]), np.array([float("nan"), -0.0])])
    assert p.read_text() == (
        "1.0000000000e+00\tnan\t0.0000000000e+00\t0.0000000000e+00\t\n"
        "2.5000000000e+00\t-0.0000000000e+00\t0.0000000000e+00\t0.0000000000e+00\t\n"
    )
    write_function_table(p, [np.array([-float("nan")])])
    assert p.read_text().startswith("-nan\t")
    with pytest.raises(ValueError):
        write_function_table(p, [np.array([np.inf])])


def test_success_mask_is_strict() -> None:
    e = np.array([9.99999968e-21, np.nextafter(9.99999968e-21, 0.0), 1e-30, 1.0])
    assert list(cli.success_mask(e, 9.99999968e-21)) == [False, True, True, False]


@pytest.mark.parametrize("d,n", [(1, 20), (2, 30), (3, 27)])
def test_isotropic_kset_matches_bruteforce(d: int, n: int) -> None:
    basis = (n ** (1.0 / d)) * np.eye(d)
    n_dir, n_mult, ksq = isotropic_directions(basis, 5.0 if d > 1 else 5.0, 40.0)
    rows = np.concatenate([np.arange(1, L + 1)[:, None] * nd[None, :] for nd, L in zip(n_dir, n_mult)])
    ref_rows, ref_ksq = O.isotropic_kset(basis, 5.0, 40.0)
    assert np.array_equal(rows, ref_rows)
    assert np.array_equal(ksq, ref_ksq)


def test_bin_moments_direct() -> None:
    rng = np.random.default_rng(0)
    k = rng.uniform(0.1, 1.0, size=200)
    s1 = rng.uniform(size=200)
    s2 = s1 * s1 + rng.uniform(0, 0.1, size=200)
    mk, ms, ek, es = _bin_moments(k, s1, s2, 3, 0.1)
    idx = np.floor(k / 0.1).astype(int)
    for j, b in enumerate(sorted(set(idx))):
        sel = idx == b
        c = sel.sum()
        assert mk[j] == pytest.approx(k[sel].mean(), rel=1e-12)
        assert ms[j] == pytest.approx(s1[sel].mean(), rel=1e-12)
        var_s = (s2[sel].mean() - s1[sel].mean() ** 2) / (3 * c)
        assert es[j] == pytest.approx(math.sqrt(var_s), rel=1e-9)


def test_hp_and_g2_tables() -> None:
    edges = np.array([0.0, 1.0, 3.0])
    hp = cli.hp_table(np.array([2, 0]), edges)
    assert np.allclose(hp[:, 0], [0.5, 2.0]) and np.allclose(hp[:, 1], [1.0, 0.0])
    assert np.allclose(hp[:, 2], [0.5, 1.0]) and np.allclose(hp[:, 3], [1.0 / math.sqrt(2), 0.0])
    g2 = cli.g2_table(np.array([1.5, 0.7]), np.array([4, 0]), edges)
    assert np.allclose(g2[:, 3], [0.75, 0.0])


def test_cli_validation() -> None:
    p = cli.build_parser()
    base = ["--dim", "2", "--ka", "1.0", "--phi", "0.05", "--load", "x", "--save", "y"]
    with pytest.raises(ValueError):
        cli.validate_args(p.parse_args(base + ["--grow", "2.0"]))
    with pytest.raises(ValueError):
        cli.validate_args(p.parse_args(base + ["--device", "cpu"]))
    with pytest.raises(ValueError):
        cli.validate_args(p.parse_args(base + ["--record-every", "2"]))
    cli.validate_args(p.parse_args(base))
