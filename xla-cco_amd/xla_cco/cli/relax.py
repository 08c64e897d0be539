"""``python -m xla_cco.cli.relax``: relax the configurations of a ConfigPack on the GPU.

A drop-in for the reference's driver (src/xla_cco/cli/relax.py:101-546): the same
command-line arguments, the same output files and the same log keys, so scripts
written for it run unchanged.  For the output prefix SAVE:

* ``SAVE.ConfigPack`` (relaxed), ``SAVE_InitConfig.ConfigPack`` (inputs) and
  ``SAVE_Success.ConfigPack`` (energy < ``--tol``), with the C++ header convention
  (cell N^(1/d), basis times N^(1/d));
* ``SAVE_Sk.txt``, ``SAVE_g2.txt`` and ``SAVE_Hp.txt`` over the Success set, in the
  C++ four-column layout (S(k) with the C++ binning and cutoffs 5 / 400; g2 on
  [0, 0.2 L] and Hp on [0, max nearest-neighbour distance], fixed uniform bins);
* ``SAVE_energies.npy``, ``SAVE_log.json`` and, when recording, ``SAVE_traj.npz``.

The job is a small pipeline: the arguments are checked against a table of rules,
the pack is read and its cell convention resolved, the k-set and the device energy
are built, the configurations are relaxed in batches on the GPU, and the outputs
are written.  There is no CPU path: ``--device cpu`` is refused.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time
from pathlib import Path
from typing import Any, Callable, Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from xla_cco import _native
from xla_cco.analysis.pairs import g2_from_histogram, nearest_neighbor_distances_batch, pair_distance_histogram
from xla_cco.analysis.structure import isotropic_structure_factor
from xla_cco.box import Box, cubic_box, wrap_fractional
from xla_cco.energy import StealthyEnergy, make_stealthy_energy
from xla_cco.io.configpack import PackedConfig, read_configpack, stack_fractional, write_configpack
from xla_cco.io.tables import write_function_table
from xla_cco.kvectors import chi_from_count, enumerate_kvectors, k_outer_from_ka
from xla_cco.minimize import relax
from xla_cco.minimize import result as _result
from xla_cco.minimize.descent import DescentSettings
from xla_cco.minimize.lbfgs import LBFGSSettings

Settings = Union[DescentSettings, LBFGSSettings]

CPP_PARITY_TOL = 9.99999968e-21  # float32(1e-20): the C++ driver's success threshold
_VOLUME_ATOL = 1e-12  # cell-volume match of the two basis conventions
_G2_RANGE = 0.2  # g2 up to this fraction of the box side
_SK_CUTOFFS = (5.0, 400.0)  # direction and modulus cutoffs of the isotropic k-set
_STATUS_NAME = {v: k for k, v in vars(_result).items() if k.startswith("STATUS_") and isinstance(v, int)}


# ---------------- arguments ----------------

_OPTIONS: Sequence[Tuple[Tuple[str, ...], Dict[str, Any]]] = (
    (("--dim",), dict(type=int, required=True, help="dimension d (1-3)")),
    (("--ka",), dict(type=float, required=True, help="K2 * a (outer radius)")),
    (("--phi",), dict(type=float, required=True, help="packing fraction")),
    (("--k1a",), dict(type=float, default=0.0, help="K1 * a (default 0)")),
    (("--load",), dict(type=Path, required=True, help="input ConfigPack prefix")),
    (("--input-basis",), dict(choices=("auto", "unit", "scaled"), default="auto",
                              help="cell convention of the input: volume 1, N, or either")),
    (("--save",), dict(type=Path, required=True, help="output prefix")),
    (("--method",), dict(choices=("lbfgs", "descent"), default="lbfgs")),
    (("--tol",), dict(type=float, default=1e-20, help="success threshold on the energy")),
    (("--max-steps",), dict(type=int, default=None,
                            help="evaluation budget: L-BFGS per pass, descent in total")),
    (("--grow",), dict(type=float, default=1.0, help="descent step growth factor")),
    (("--stop-tol",), dict(type=float, default=0.0, help="stop a configuration below this energy (0: never)")),
    (("--batch-size",), dict(type=int, default=1024, help="configurations per relax call")),
    (("--chunk-size",), dict(type=int, default=None, help="accepted for compatibility")),
    (("--device",), dict(choices=("gpu", "cpu"), default="gpu")),
    (("--record-every",), dict(type=int, default=0, help="trajectory stride (iterations)")),
    (("--max-records",), dict(type=int, default=0, help="trajectory rows per configuration")),
    (("--g2-bins",), dict(type=int, default=200, help="bins of the g2 grid")),
    (("--hp-bins",), dict(type=int, default=200, help="bins of the Hp grid")),
)


def build_parser() -> argparse.ArgumentParser:
    """The argument parser (the reference's option names and defaults)."""
    parser = argparse.ArgumentParser(
        prog="python -m xla_cco.cli.relax",
        description="Relax ConfigPack configurations to stealthy ground states on an MI355X. "
        "Success is energy < --tol; an L-BFGS run that ends in a line-search failure far "
        "below --tol has converged to round-off.",
    )
    for flags, kw in _OPTIONS:
        parser.add_argument(*flags, **kw)
    return parser


# (test, message) pairs; a rule fails when its test is False
_RULES: Sequence[Tuple[Callable[[argparse.Namespace], bool], Callable[[argparse.Namespace], str]]] = (
    (lambda a: a.dim in (1, 2, 3), lambda a: f"--dim is 1, 2 or 3 (got {a.dim})"),
    (lambda a: a.ka > 0.0, lambda a: f"--ka must be > 0 (got {a.ka})"),
    (lambda a: a.phi > 0.0, lambda a: f"--phi must be > 0 (got {a.phi})"),
    (lambda a: 0.0 <= a.k1a < a.ka, lambda a: f"--k1a must lie in [0, --ka) (got {a.k1a}, --ka {a.ka})"),
    (lambda a: a.tol > 0.0, lambda a: f"--tol must be > 0 (got {a.tol})"),
    (lambda a: a.stop_tol >= 0.0, lambda a: f"--stop-tol must be >= 0 (got {a.stop_tol})"),
    (lambda a: a.max_steps is None or a.max_steps >= 1, lambda a: f"--max-steps must be >= 1 (got {a.max_steps})"),
    (lambda a: a.method == "descent" or a.grow == 1.0, lambda a: "--grow only applies with --method descent"),
    (lambda a: a.grow >= 1.0, lambda a: f"--grow must be >= 1 (got {a.grow})"),
    (lambda a: a.batch_size >= 1, lambda a: f"--batch-size must be >= 1 (got {a.batch_size})"),
    (lambda a: a.g2_bins >= 1 and a.hp_bins >= 1, lambda a: "--g2-bins and --hp-bins must be >= 1"),
    (lambda a: a.chunk_size is None or a.chunk_size >= 1, lambda a: f"--chunk-size must be >= 1 (got {a.chunk_size})"),
    (lambda a: a.record_every >= 0 and a.max_records >= 0, lambda a: "--record-every and --max-records must be >= 0"),
    (lambda a: (a.record_every > 0) == (a.max_records > 0),
     lambda a: "--record-every and --max-records are given together (both > 0) or not at all"),
    (lambda a: a.device == "gpu", lambda a: "--device cpu is not available: the relax runs on the GPU"),
)


def validate_args(args: argparse.Namespace) -> None:
    """ValueError naming the first argument that breaks a rule."""
    for ok, message in _RULES:
        if not ok(args):
            raise ValueError(message(args))


def select_device(device: str) -> str:
    """Open the GPU for ``--device gpu`` (RuntimeError naming the cause if none is
    visible); ``cpu`` is refused with ValueError, never served by a CPU path."""
    if device == "cpu":
        raise ValueError("--device cpu is not available: the relax runs on the GPU")
    if device != "gpu":
        raise ValueError(f"--device must be gpu or cpu, got {device!r}")
    try:
        _native.context(0)
    except RuntimeError as exc:
        raise RuntimeError(f"--device gpu: {exc}") from exc
    return "gpu"


def wrap_fractional_batch(s: np.ndarray) -> np.ndarray:
    """Fractional positions of a (b, n, d) batch wrapped into [0, 1)."""
    if np.ndim(s) != 3:
        raise ValueError(f"expected a (b, n, d) batch, got shape {np.shape(s)}")
    return wrap_fractional(np.asarray(s))


# ---------------- tables and masks ----------------


def success_mask(energy: np.ndarray, tol: float) -> np.ndarray:
    """Energy strictly below ``tol`` (the C++ ``E < E_max``)."""
    if not tol > 0.0:
        raise ValueError(f"tol must be > 0, got {tol}")
    return np.asarray(energy) < tol


def _centres_and_halves(edges: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    half = 0.5 * (edges[1:] - edges[:-1])
    return edges[:-1] + half, half


def hp_table(counts: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """Nearest-neighbour density table: r_mid, p = count / (total width), half width, p / sqrt(count)."""
    c = np.asarray(counts, dtype=np.float64)
    mid, half = _centres_and_halves(np.asarray(edges, dtype=np.float64))
    density = c / (c.sum() * 2.0 * half)
    err = np.zeros_like(density)
    hit = c > 0
    err[hit] = density[hit] / np.sqrt(c[hit])
    return np.column_stack([mid, density, half, err])


def g2_table(g2: np.ndarray, counts: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """Pair-correlation table: r_mid, g2, half width, g2 / sqrt(count)."""
    g = np.asarray(g2, dtype=np.float64)
    c = np.asarray(counts, dtype=np.float64)
    mid, half = _centres_and_halves(np.asarray(edges, dtype=np.float64))
    err = np.zeros_like(g)
    hit = c > 0
    err[hit] = g[hit] / np.sqrt(c[hit])
    return np.column_stack([mid, g, half, err])


# ---------------- the job ----------------


@dataclasses.dataclass
class _Inputs:
    configs: List[PackedConfig]  # with unit-volume bases
    s0: np.ndarray  # (B, N, d) wrapped fractional positions
    basis: np.ndarray  # unit-volume basis shared by all configurations

    @property
    def n(self) -> int:
        return int(self.s0.shape[1])


def _unit_volume_config(cfg: PackedConfig, i: int, n: int, dim: int, convention: str) -> PackedConfig:
    """The record with its basis scaled to unit volume, per the allowed convention."""
    volume = abs(float(np.linalg.det(np.asarray(cfg.basis))))
    unit_ok = convention in ("auto", "unit") and abs(volume - 1.0) <= _VOLUME_ATOL
    scaled_ok = convention in ("auto", "scaled") and abs(volume / n - 1.0) <= _VOLUME_ATOL * max(1.0, float(n))
    if unit_ok:
        return cfg
    if scaled_ok:
        return dataclasses.replace(cfg, basis=np.asarray(cfg.basis, dtype=np.float64) / float(n) ** (1.0 / dim))
    raise ValueError(
        f"configuration {i} has cell volume {volume!r}, which --input-basis {convention} does not accept: "
        f"volume 1 is a unit basis (the C++ input mode scales it by N^(1/d)), volume N = {n} a basis "
        f"already scaled by N^(1/d) (the convention of this driver's outputs)"
    )


def _read_inputs(prefix: Path, dim: int, convention: str) -> _Inputs:
    records = read_configpack(prefix)
    if not records:
        raise ValueError(f"{prefix} holds no configurations")
    pack_dim = int(np.asarray(records[0].relative).shape[1])
    if pack_dim != dim:
        raise ValueError(f"--dim {dim} but the pack holds {pack_dim}-d configurations")
    s0, _ = stack_fractional(records)
    n = int(s0.shape[1])
    units = [_unit_volume_config(c, i, n, dim, convention) for i, c in enumerate(records)]
    return _Inputs(units, wrap_fractional(s0), np.asarray(units[0].basis, dtype=np.float64))


@dataclasses.dataclass
class _Physics:
    box: Box
    energy: StealthyEnergy
    info: Dict[str, Any]  # the log's N, d, M, chi, a, K1, K2


def _build_physics(inputs: _Inputs, args: argparse.Namespace) -> _Physics:
    n, dim = inputs.n, args.dim
    if n < 2:
        raise ValueError(f"a stealthy relax needs N >= 2 particles, got {n}")
    if np.array_equal(inputs.basis, np.eye(dim)):
        box = cubic_box(n, dim)
    else:
        box = Box(basis=inputs.basis * float(n) ** (1.0 / dim))
    k_out = k_outer_from_ka(args.ka, args.phi, dim)
    k_in = k_outer_from_ka(args.k1a, args.phi, dim) if args.k1a > 0.0 else 0.0
    kset = enumerate_kvectors(np.asarray(box.basis), k_in, k_out)
    m = int(kset.n.shape[0])
    if m == 0:
        raise ValueError(f"no k-vector with {k_in} < |k| < {k_out}: the k-set is empty")
    info = {"N": n, "d": dim, "M": m, "chi": chi_from_count(m, n, dim), "a": args.ka / k_out,
            "K1": k_in, "K2": k_out}
    return _Physics(box, make_stealthy_energy(box, kset, chunk_size=args.chunk_size), info)


def _settings(args: argparse.Namespace) -> Settings:
    shared: Dict[str, Any] = dict(energy_tol=args.stop_tol, record_every=args.record_every,
                                  max_records=args.max_records)
    if args.method == "descent":
        budget = {} if args.max_steps is None else {"max_evals": args.max_steps}
        return DescentSettings(grow=args.grow, **shared, **budget)
    budget = {} if args.max_steps is None else {"max_evals_per_pass": args.max_steps}
    return LBFGSSettings(**shared, **budget)


_RESULT_FIELDS = ("energy", "n_evals", "n_iters", "status")
_TRAJ_FIELDS = ("s", "energy", "step", "iteration", "count")


def _relax_all(energy: StealthyEnergy, s0: np.ndarray, settings: Settings, batch: int) -> Dict[str, np.ndarray]:
    """Relax in batches of ``batch`` configurations; per-field arrays in input order."""
    pieces: Dict[str, List[np.ndarray]] = {}
    for lo in range(0, s0.shape[0], batch):
        res = relax(energy, np.ascontiguousarray(s0[lo: lo + batch]), settings)
        fields = {"s": wrap_fractional(res.s)}
        fields.update({f: getattr(res, f) for f in _RESULT_FIELDS})
        if res.trajectory is not None:
            fields.update({f"traj_{f}": np.asarray(getattr(res.trajectory, f)) for f in _TRAJ_FIELDS})
        for key, value in fields.items():
            pieces.setdefault(key, []).append(value)
    return {key: np.concatenate(parts) for key, parts in pieces.items()}


def _as_packed(templates: Sequence[PackedConfig], s: np.ndarray, side: float) -> List[PackedConfig]:
    """Records with the C++ header convention: cell = side, basis times side."""
    return [PackedConfig(basis=np.asarray(t.basis, dtype=np.float64) * side,
                         relative=np.ascontiguousarray(x, dtype=np.float64), cell_size=side, names=t.names)
            for t, x in zip(templates, s)]


def _write_structure(prefix: str, s: np.ndarray, box: Box, args: argparse.Namespace) -> Dict[str, Any]:
    """S(k), g2 and Hp tables of the Success configurations."""
    n, dim = s.shape[1], s.shape[2]
    side = float(n) ** (1.0 / dim)
    k_mod, s_k, k_err, s_err = isotropic_structure_factor(s, box, *_SK_CUTOFFS)
    write_function_table(Path(f"{prefix}_Sk.txt"), [k_mod, s_k, k_err, s_err])
    edges = np.linspace(0.0, _G2_RANGE * side, args.g2_bins + 1)
    counts = pair_distance_histogram(s, box, edges)
    g2 = g2_from_histogram(counts, edges, n, s.shape[0], float(n), dim)
    write_function_table(Path(f"{prefix}_g2.txt"), list(g2_table(g2, counts, edges).T))
    nn = nearest_neighbor_distances_batch(s, box).ravel()
    hp_edges = np.linspace(0.0, float(max(nn.max(), 1e-12)), args.hp_bins + 1)
    hp_counts = np.histogram(nn, bins=hp_edges)[0]
    write_function_table(Path(f"{prefix}_Hp.txt"), list(hp_table(hp_counts, hp_edges).T))
    return {"analysis": "written", "sk_nan_error_bins": int(np.isnan(s_err).sum())}


def _tallies(out: Dict[str, np.ndarray], tol: float) -> Dict[str, Any]:
    codes, counts = np.unique(out["status"], return_counts=True)
    evals = out["n_evals"]
    return {
        "status_counts": {_STATUS_NAME.get(int(c), f"STATUS_{int(c)}"): int(k) for c, k in zip(codes, counts)},
        "n_success": int(success_mask(out["energy"], tol).sum()),
        "n_evals": {"min": int(evals.min()), "mean": float(evals.mean()), "median": float(np.median(evals)),
                    "max": int(evals.max())},
    }


def run(args: argparse.Namespace) -> Dict[str, Any]:
    """Run the whole job for parsed ``args``; returns the log written to SAVE_log.json."""
    validate_args(args)
    clock = [time.perf_counter()]
    device = select_device(args.device)
    inputs = _read_inputs(args.load, args.dim, args.input_basis)
    phys = _build_physics(inputs, args)
    settings = _settings(args)
    clock.append(time.perf_counter())
    print(f"relaxing {inputs.s0.shape[0]} configurations, N={inputs.n} M={phys.info['M']}, on the GPU", flush=True)
    out = _relax_all(phys.energy, inputs.s0, settings, args.batch_size)
    clock.append(time.perf_counter())

    prefix = str(args.save)
    Path(prefix).parent.mkdir(parents=True, exist_ok=True)
    side = float(inputs.n) ** (1.0 / args.dim)
    write_configpack(Path(prefix), _as_packed(inputs.configs, out["s"], side))
    write_configpack(Path(f"{prefix}_InitConfig"), _as_packed(inputs.configs, inputs.s0, side))
    np.save(f"{prefix}_energies.npy", out["energy"])
    log = _tallies(out, args.tol)
    ok = success_mask(out["energy"], args.tol)
    if ok.any():
        chosen = [c for c, keep in zip(inputs.configs, ok) if keep]
        write_configpack(Path(f"{prefix}_Success"), _as_packed(chosen, out["s"][ok], side))
        log.update(_write_structure(prefix, out["s"][ok], phys.box, args))
    else:
        log["analysis"] = "skipped: no Success configurations"
    traj = {k[len("traj_"):]: v for k, v in out.items() if k.startswith("traj_")}
    if traj:
        np.savez(f"{prefix}_traj.npz", **traj)
    clock.append(time.perf_counter())
    log.update(phys.info)
    log.update({
        "args": {k: str(v) if isinstance(v, Path) else v for k, v in vars(args).items()},
        "tol": args.tol,
        "method": args.method,
        "settings": dataclasses.asdict(settings),
        "n_configs": int(inputs.s0.shape[0]),
        "wall_seconds": {"setup": clock[1] - clock[0], "relax": clock[2] - clock[1], "output": clock[3] - clock[2]},
        "jax_version": "not used (HIP gfx950 library)",
        "device": f"{device}:0 (HIP gfx950)",
    })
    Path(f"{prefix}_log.json").write_text(json.dumps(log, indent=2) + "\n")
    return log


def main(argv: Optional[Sequence[str]] = None) -> int:
    """Entry point: 0 on success, 1 (message on stderr) on a usage or runtime error."""
    args = build_parser().parse_args(argv)
    try:
        log = run(args)
    except (ValueError, TypeError, RuntimeError, OSError) as exc:
        print(f"ERROR: {exc}", file=sys.stderr, flush=True)
        return 1
    print(json.dumps(log, indent=2), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
