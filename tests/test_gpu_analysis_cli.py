"""GPU analysis (S(k) moments, pair statistics) and the relax CLI end to end.

Pair statistics and S(k) are checked against brute-force oracles
(tests/oracle.py:158-209 of the reference); the CLI mirrors
tests/integration/test_cli.py (log keys, strict success set, outputs).
"""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from xla_cco.analysis.pairs import (
    g2_from_histogram,
    nearest_neighbor_distances,
    nearest_neighbor_distances_batch,
    pair_distance_histogram,
)
from xla_cco.analysis.structure import isotropic_directions, isotropic_structure_factor, sk_moments
from xla_cco.box import Box, cubic_box
from xla_cco.cli import relax as cli
from xla_cco.io.configpack import read_configpack

pytestmark = pytest.mark.gpu
INPUTS = Path(__file__).resolve().parent / "golden" / "inputs"

LOG_KEYS = {"args", "N", "d", "M", "chi", "a", "K1", "K2", "tol", "method", "settings",
            "status_counts", "n_success", "n_evals", "wall_seconds", "jax_version", "device"}


@pytest.mark.parametrize("d,n", [(1, 24), (2, 30), (3, 27)])
def test_sk_moments_match_oracle(gpu_ctx, d: int, n: int) -> None:
    basis = (n ** (1.0 / d)) * np.eye(d)
    n_dir, n_mult, ksq = isotropic_directions(basis, 5.0, 60.0)
    rows = np.concatenate([np.arange(1, L + 1)[:, None] * nd[None, :] for nd, L in zip(n_dir, n_mult)])
    s = np.random.default_rng(d).uniform(size=(3, n, d))
    m1, m2 = sk_moments(s, n_dir, n_mult)
    sk = np.stack([O.structure_factor(s[i], rows) for i in range(3)])
    assert np.allclose(m1, sk.mean(axis=0), rtol=1e-9, atol=1e-12)
    assert np.allclose(m2, (sk * sk).mean(axis=0), rtol=1e-9, atol=1e-12)


def test_isotropic_structure_factor_runs_and_bins(gpu_ctx) -> None:
    box = cubic_box(64, 2)
    s = np.random.default_rng(0).uniform(size=(4, 64, 2))
    k, sk, ek, es = isotropic_structure_factor(s, box)
    assert k.shape == sk.shape == ek.shape == es.shape and k.size > 100
    assert np.all(np.diff(k) > 0) and np.all(sk >= 0)
    # Poisson points: S(k) ~ 1 on average at large k
    assert 0.8 < float(np.mean(sk[k > 50])) < 1.2


@pytest.mark.parametrize("d,n", [(1, 17), (2, 40), (3, 30)])
def test_pair_stats_match_oracle(gpu_ctx, d: int, n: int) -> None:
    box = cubic_box(n, d)
    s = np.random.default_rng(5 + d).uniform(size=(3, n, d))
    side = float(box.basis[0, 0])
    edges = np.linspace(0.0, 0.5 * side, 37)
    counts = pair_distance_histogram(s, box, edges)
    ref = np.zeros(edges.size - 1, dtype=np.int64)
    for i in range(3):
        pd, nn_ref = O.min_image_distances(s[i], box.basis)
        ref += np.histogram(pd, bins=edges)[0]
        nn = nearest_neighbor_distances(s[i], box)
        assert np.allclose(nn, nn_ref, rtol=1e-14, atol=1e-14)
    assert np.array_equal(counts, ref)
    nnb = nearest_neighbor_distances_batch(s, box)
    assert nnb.shape == (3, n)


def test_g2_of_poisson_is_flat(gpu_ctx) -> None:
    box = cubic_box(400, 2)
    s = np.random.default_rng(1).uniform(size=(8, 400, 2))
    edges = np.linspace(0.5, 4.0, 15)
    counts = pair_distance_histogram(s, box, edges)
    g2 = g2_from_histogram(counts, edges, 400, 8, 400.0, 2)
    assert np.all(np.abs(g2 - 1.0) < 0.08)


def test_single_particle_nearest_image(gpu_ctx) -> None:
    box = cubic_box(1, 2)
    nn = nearest_neighbor_distances(np.array([[0.3, 0.4]]), box)
    assert nn[0] == pytest.approx(1.0, rel=1e-14)


@pytest.mark.parametrize("method", ["lbfgs", "descent"])
def test_cli_end_to_end_d3_n100(gpu_ctx, tmp_path: Path, method: str) -> None:
    """The CPU-runnable reference case: d=3 N=100, Ka=0.5963271031078479, phi=0.05 (M=13)."""
    save = tmp_path / "out" / "S"
    args = ["--dim", "3", "--ka", "0.5963271031078479", "--phi", "0.05",
            "--load", str(INPUTS / "poisson_d3_n100"), "--save", str(save), "--method", method,
            "--tol", "9.99999968e-21", "--stop-tol", "9.99999968e-21"]
    if method == "descent":
        args += ["--grow", "2.0", "--max-steps", "200000"]
    assert cli.main(args) == 0
    log = json.loads(Path(str(save) + "_log.json").read_text())
    assert LOG_KEYS <= set(log)
    assert log["M"] == 13 and log["N"] == 100 and log["n_success"] == 5
    e = np.load(str(save) + "_energies.npy")
    assert e.shape == (5,) and np.all(e < 9.99999968e-21)
    out = read_configpack(save)
    succ = read_configpack(Path(str(save) + "_Success"))
    init = read_configpack(Path(str(save) + "_InitConfig"))
    assert len(out) == len(succ) == len(init) == 5
    side = 100 ** (1 / 3)
    assert out[0].cell_size == pytest.approx(side) and np.allclose(out[0].basis, side * np.eye(3))
    assert np.all(out[0].relative >= 0.0) and np.all(out[0].relative < 1.0)
    src = read_configpack(INPUTS / "poisson_d3_n100")
    assert np.array_equal(init[2].relative, src[2].relative)
    for name in ("_Sk.txt", "_g2.txt", "_Hp.txt"):
        rows = Path(str(save) + name).read_text().splitlines()
        assert len(rows) > 10 and all(len(r.split("\t")) == 5 for r in rows)
    # the relaxed positions really are ground states
    from xla_cco import cubic_box, enumerate_kvectors, k_outer_from_ka, make_stealthy_energy
    box = cubic_box(100, 3)
    energy = make_stealthy_energy(box, enumerate_kvectors(box.basis, 0.0, k_outer_from_ka(0.5963271031078479, 0.05, 3)))
    phi = energy.energy_batch(np.stack([c.relative for c in out]))
    assert np.all(phi < 1e-19)


def test_cli_scaled_input_and_errors(gpu_ctx, tmp_path: Path) -> None:
    first = tmp_path / "a" / "S"
    base = ["--dim", "2", "--ka", "0.3", "--phi", "0.05", "--max-steps", "50"]
    assert cli.main(base + ["--load", str(INPUTS / "poisson_d2_n64"), "--save", str(first)]) == 0
    # this CLI's outputs (basis scaled by N^(1/d)) are valid inputs with --input-basis auto
    assert cli.main(base + ["--load", str(first), "--save", str(tmp_path / "b" / "S")]) == 0
    assert cli.main(base + ["--load", str(first), "--save", str(tmp_path / "c" / "S"),
                            "--input-basis", "unit"]) == 1
    assert cli.main(base + ["--dim", "3", "--load", str(INPUTS / "poisson_d2_n64"),
                            "--save", str(tmp_path / "d" / "S")]) == 1
