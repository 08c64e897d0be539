"""Benchmark: converged configurations per second of the batched FP64 relax on one MI355X.

Workload (one "step"): relax 256 configurations of N=1000 particles in d=2
(chi=0.45, M=894 half-space k-vectors, unit density) from uniform random
positions with three-pass L-BFGS until E < 1e-20 (or another stop), all in
FP64 on the GPU.  value = configurations with E < 1e-20 per second of wall
time, inputs already resident in device memory when the timed region starts.

    python bench.py [--gpus 1] [--steps 2] [--warmup 1]

Prints one JSON line.  Per-kernel device times come from HIP events recorded
on the library's streams around the launches of every 4th evaluation step
(--profile-every) inside the timed region; the roofline object describes the
kernel with the largest total time.  The CPU
baseline relaxes one configuration of the same workload per host thread (at
most 16) to completion with the oracle's C restatement and reports converged
configurations per wall second.
"""

from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "xla-cco_amd"))
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0
METRIC = "converged configs/s (E<1e-20), mixed vs FP64 on 1 GPU; energy+force evals/s"
STAGES = ("rho_kernel", "reduce_kernel", "force_kernel", "lbfgs_advance_kernel", "compact_kernel",
          "md_prepare_kernel", "lbfgs_gram_dot_kernel", "lbfgs_gram_solve_kernel", "lbfgs_gram_comb_kernel",
          "md_tracked_rho_reduce")


def build_workload(d: int, n: int, chi: float):
    from xla_cco import cubic_box, enumerate_kvectors, k_outer_from_ka, ka_from_chi, make_stealthy_energy

    box = cubic_box(n, d)
    ks = enumerate_kvectors(box.basis, 0.0, k_outer_from_ka(ka_from_chi(chi, 0.05, d), 0.05, d))
    return box, ks, make_stealthy_energy(box, ks)


def cpu_baseline(ks, volume: float, d: int, n: int, seconds: float):
    """Relax configurations of the same workload to completion with the C oracle on the host cores.

    One configuration per OpenMP thread (at most 16), the oracle's L-BFGS with per-axis
    phase tables (the reference's table path restated in C; 4x faster on the CPU than
    its cos/sin form), run until every configuration stops; converged configs/s =
    configurations with E < 1e-20 / wall seconds.  ``seconds`` bounds the run: each
    configuration may take at most the evaluations that fit in it (never reached on the
    default workload, about 10 s for about 760 evaluations per configuration).
    """
    from oracle import c_oracle, oracle as O

    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        threads = os.cpu_count() or 1
    threads = max(1, min(threads, 16))
    s0 = np.random.default_rng(99).uniform(size=(threads, n, d))
    settings = O.LBFGSSettings(energy_tol=1e-20)
    t0 = time.perf_counter()
    c_oracle.relax(s0[:1], ks.n, ks.weights, volume, settings, eval_cap=3, phases=1)
    per_eval = (time.perf_counter() - t0) / 3.0
    cap = max(4, int(seconds / max(per_eval, 1e-6)))
    t0 = time.perf_counter()
    _, e, ne, _, st = c_oracle.relax(
        s0, ks.n, ks.weights, volume, settings, eval_cap=cap, threads=threads, phases=1
    )
    wall = time.perf_counter() - t0
    converged = int(np.sum(e < 1e-20))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), model)
    except OSError:  # pragma: no cover
        pass
    return {
        "value": converged / wall,
        "unit": "configs/s",
        "cores": threads,
        "kind": "port",
        "converged": f"{converged}/{threads}",
        "evals_per_s": float(ne.sum()) / wall,
        "evals_per_converged": float(ne.sum()) / max(converged, 1),
        "host": {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))},
        "sample": (
            f"C oracle (oracle/cco_oracle.c, per-axis phase tables, same L-BFGS) relaxing {threads} "
            f"configurations of the workload to completion, one per OpenMP thread "
            f"({converged} converged, {int(ne.sum())} evaluations, {wall:.1f} s wall, cap {cap} "
            f"evaluations per configuration not reached: {bool(np.all(ne < cap))}); "
            f"value = converged configurations / wall seconds"
        ),
    }


# Converged configs/s the reference published for its CLI L-BFGS path on an A100 80GB
# PCIe: d=3, N=100, 1200 configurations, batches of 1024, tol 9.99999968e-21, no early
# stop (benchmarks/RESULTS.md:116-120).
PUBLISHED_SHAPE = {0.35: {"value": 25.4, "wall_s": 47.3, "total_evals": 5087505},
                   0.45: {"value": 20.7, "wall_s": 57.9, "total_evals": 5637167}}


def published_shape(chi: float, reps: int = 2, autotune: bool = False) -> dict:
    """The reference's published converged-configs/s workload, run the way its CLI runs it.

    1200 uniform-random configurations (seed 7; the reference's cco_runs inputs are not
    public), d=3, N=100, phi=0.05, chi -> Ka; LBFGSSettings() with no early stop
    (--stop-tol 0), relaxed in batches of 1024 as relax_batches does; success is
    energy < 9.99999968e-21.  Host-to-device input upload is inside the timed region,
    as in the reference's measurement.  Best of ``reps`` warm runs.
    """
    from xla_cco.minimize import LBFGSSettings, relax

    tol = 9.99999968e-21
    _, ks, energy = build_workload(3, 100, chi)
    s0 = np.random.default_rng(7).uniform(size=(1200, 100, 3))
    relax(energy, s0[:64], LBFGSSettings())  # warm the session paths
    if autotune:  # tune both batch shapes outside the timed region
        for lo in range(0, 1200, 1024):
            relax(energy, np.ascontiguousarray(s0[lo: lo + 1024]), LBFGSSettings(max_evals_per_pass=20),
                  autotune=True)
    best, out = float("inf"), None
    for _ in range(reps):
        t0 = time.perf_counter()
        parts = [relax(energy, np.ascontiguousarray(s0[lo: lo + 1024]), LBFGSSettings(), autotune=autotune)
                 for lo in range(0, 1200, 1024)]
        wall = time.perf_counter() - t0
        if wall < best:
            best, out = wall, parts
    energies = np.concatenate([p.energy for p in out])
    evals = int(sum(int(p.n_evals.sum()) for p in out))
    ok = int(np.sum(energies < tol))
    # the same 1200 configurations as one batch (its L-BFGS history, 0.6 GB, is
    # far inside one GPU's 288 GB): no second, mostly idle batch of 176 lanes
    one, one_out = float("inf"), None
    for _ in range(reps):
        t0 = time.perf_counter()
        res = relax(energy, s0, LBFGSSettings(), autotune=autotune)
        wall = time.perf_counter() - t0
        if wall < one:
            one, one_out = wall, res
    ok_one = int(np.sum(one_out.energy < tol))
    ref = PUBLISHED_SHAPE[chi]
    return {"value": ok / best, "unit": "configs/s", "converged": f"{ok}/1200", "wall_s": best,
            "total_evals": evals, "M": int(ks.n.shape[0]),
            "reference": {"value": ref["value"], "hardware": "A100 80GB PCIe", "total_evals": ref["total_evals"],
                          "source": "benchmarks/RESULTS.md:116-120"},
            "vs_reference": (ok / best) / ref["value"],
            "one_batch": {"value": ok_one / one, "converged": f"{ok_one}/1200", "wall_s": one,
                          "total_evals": int(one_out.n_evals.sum()),
                          "note": "all 1200 configurations in one relax call instead of batches of 1024"}}


# evals/s the reference published for relax on an A100-SXM4-80GB (benchmarks/RESULTS.md:55-66):
# d=3, N=100, 1000 configurations of default_rng(0), per-(k,j) trig path, warm wall time
PUBLISHED_RELAX_RATES = {
    ("lbfgs", 0.05): {"evals": 455419, "wall_s": 1.58, "evals_per_s": 2.88e5},
    ("lbfgs", 0.45): {"evals": 557049, "wall_s": 15.50, "evals_per_s": 3.594e4},
    ("descent_400", 0.05): {"evals": 400000, "wall_s": 0.10, "evals_per_s": 3.998e6},
    ("descent_400", 0.45): {"evals": 400000, "wall_s": 0.16, "evals_per_s": 2.477e6},
}


def published_relax_rates(reps: int = 2) -> dict:
    """The reference's published relax benchmark rows, on the same inputs.

    relax of default_rng(0).random((1000, 100, 3)) at chi 0.05 and 0.45 (phi=0.05) with
    LBFGSSettings(max_evals_per_pass=200) and DescentSettings(max_evals=400), as
    benchmarks/bench_energy.py:147-170 times them: host arrays in, results out, best of
    ``reps`` warm runs.  evaluations/s = total n_evals / wall seconds.
    """
    from xla_cco.minimize import DescentSettings, LBFGSSettings, relax

    out = {}
    s0 = np.random.default_rng(0).random((1000, 100, 3))
    for chi in (0.05, 0.45):
        _, _, energy = build_workload(3, 100, chi)
        for name, st in (("lbfgs", LBFGSSettings(max_evals_per_pass=200)),
                         ("descent_400", DescentSettings(max_evals=400))):
            relax(energy, s0[:64], st)  # warm the session paths
            best, evals = float("inf"), 0
            for _ in range(reps):
                t0 = time.perf_counter()
                res = relax(energy, s0, st)
                wall = time.perf_counter() - t0
                if wall < best:
                    best, evals = wall, int(res.n_evals.sum())
            ref = PUBLISHED_RELAX_RATES[(name, chi)]
            out[f"{name}_chi{chi}"] = {
                "evals": evals, "wall_s": best, "evals_per_s": evals / best,
                "reference": dict(ref, hardware="A100-SXM4-80GB", source="benchmarks/RESULTS.md:55-66"),
                "wall_vs_reference": ref["wall_s"] / best,
            }
    return out


def energy_times(d: int, n: int, chi: float, B: int, steps: int, warmup: int):
    """Seconds per batched value_and_grad of B device-resident configurations, FP64 and
    split precision; (times, M, split precision applied)."""
    from xla_cco import _native

    ctx = _native.context(0)
    _, ks, e64 = build_workload(d, n, chi)
    _, _, emp = build_workload(d, n, chi)
    mixed_ok = emp.set_precision("mixed")
    s = _native.DeviceArray.from_host(ctx, np.random.default_rng(17).uniform(size=(B, n, d)))
    phi = _native.DeviceArray(ctx, (B,))
    grad = _native.DeviceArray(ctx, (B, n, d))
    L = _native.lib()
    out = {}
    for name, e in (("fp64", e64), ("mixed", emp)):
        for _ in range(warmup):
            _native.check(L.cco_value_grad_density_dev(e.handle(), s.ptr, B, n, phi.ptr, grad.ptr, None, None))
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            _native.check(L.cco_value_grad_density_dev(e.handle(), s.ptr, B, n, phi.ptr, grad.ptr, None, None))
        ctx.synchronize()
        out[name] = (time.perf_counter() - t0) / steps
    for a in (s, phi, grad):
        a.free()
    return out, int(ks.n.shape[0]), mixed_ok


def bench_energy(args) -> int:
    """Batched energy+force evaluations at one shape, FP64 and split precision: lane-evals/s."""
    d, n, chi, B = args.d, args.n, args.chi, args.lanes
    out, M, mixed_ok = energy_times(d, n, chi, B, args.steps, args.warmup)
    flop = 8.0 * M * n * (1 + d) * B
    line = {
        "metric": "energy+force lane-evaluations/s (batched value_and_grad on device-resident positions)",
        "value": B / out["fp64"],
        "unit": "lane-evaluations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["fp64"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"value_and_grad of {B} uniform-random configurations, d={d} N={n} chi={chi} (M={M})",
                   "lanes": B, "M": M},
        "algorithmic_tflops": flop / out["fp64"] / 1e12,
        "mixed": {"applied": mixed_ok, "value": B / out["mixed"], "ms_per_step": out["mixed"] * 1e3,
                  "vs_fp64": out["fp64"] / out["mixed"]},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_uneven(args) -> int:
    """Uneven batch: 512 configurations of different N, d and chi in one session."""
    from xla_cco import _native
    from xla_cco.minimize import LBFGSSettings, UnevenSession

    L = _native.lib()
    ctx = _native.context(0)
    B = 512
    tol = 1e-20
    rng = np.random.default_rng(2024)
    specs = [(int(rng.integers(100, 2001)), int(rng.integers(2, 4)), float(rng.choice([0.1, 0.2, 0.3, 0.45])))
             for _ in range(B)]
    cache = {}
    energies, m_real = [], []
    for n, d, chi in specs:
        if (n, d, chi) not in cache:
            _, ks, e = build_workload(d, n, chi)
            cache[(n, d, chi)] = (e, int(ks.n.shape[0]))
        energies.append(cache[(n, d, chi)][0])
        m_real.append(cache[(n, d, chi)][1])
    settings = LBFGSSettings(energy_tol=tol)
    sess = UnevenSession(energies, [(n, d) for n, d, _ in specs], settings)
    _native.check(L.cco_relax_set_option(sess.handle, 2, args.shards))
    bound = 1 + 3 * (settings.max_evals_per_pass + settings.max_linesearch_steps + 1)
    inputs = [
        [np.random.default_rng(5000 + 97 * k + i).uniform(size=(n, d)) for i, (n, d, _) in enumerate(specs)]
        for k in range(args.warmup + args.steps)
    ]
    for w in range(args.warmup):
        sess.init(inputs[w])
        sess.run(bound, 256)
    _native.check(L.cco_relax_set_profiling(sess.handle, 0 if args.no_profile else args.profile_every))
    elapsed, converged, lane_evals = 0.0, 0, 0
    for k in range(args.steps):
        sess.init(inputs[args.warmup + k])  # host upload outside the timed region
        ctx.synchronize()
        t0 = time.perf_counter()
        sess.run(bound, 256)
        ctx.synchronize()
        elapsed += time.perf_counter() - t0
        res = sess.result()
        converged += sum(1 for r in res if float(r.energy) < tol)
        cnt = (C.c_int64 * 4)()
        _native.check(L.cco_relax_counters(sess.handle, cnt))
        lane_evals += cnt[0]
    ms = (C.c_double * len(STAGES))()
    launches = (C.c_int64 * len(STAGES))()
    _native.check(L.cco_relax_profile(sess.handle, ms, launches))
    stage_avg = {STAGES[i]: ms[i] / launches[i] for i in range(len(STAGES)) if launches[i] > 0}
    flops = sum(8.0 * m_real[i] * n * (1 + d) for i, (n, d, _) in enumerate(specs))
    sess.close()
    # the same batch with split precision (lanes that fit the one-workgroup-per-
    # lane kernel start split, all turn FP64 at E <= 1e-8): one warm, one timed relax
    mixed = None
    if not args.no_mixed:
        ms_ = UnevenSession(energies, [(n, d) for n, d, _ in specs], LBFGSSettings(energy_tol=tol, precision="mixed"))
        _native.check(L.cco_relax_set_option(ms_.handle, 2, args.shards))
        _native.check(L.cco_relax_set_option(ms_.handle, 5, args.split_history))
        for _ in range(2):
            ms_.init(inputs[args.warmup])
            split0 = ms_.precision_state()[1]
            ctx.synchronize()
            t0 = time.perf_counter()
            ms_.run(bound, 256)
            ctx.synchronize()
            t_m = time.perf_counter() - t0
        conv_m = sum(1 for r in ms_.result() if float(r.energy) < tol)
        ms_.close()
        mixed = {"value": conv_m / t_m, "unit": "configs/s", "converged": f"{conv_m}/{B}",
                 "lanes_split_at_start": split0, "vs_fp64": (conv_m / t_m) / (converged / elapsed),
                 "split_history": "fp32" if args.split_history else "fp64"}
    line = {
        "metric": METRIC,
        "value": converged / elapsed,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"uneven batch: {B} configurations, N ~ U{{100..2000}}, d ~ U{{2,3}}, "
            "chi ~ U{0.1,0.2,0.3,0.45}, one session, FP64 L-BFGS to E<1e-20",
            "lanes": B,
            "distinct_energies": len(cache),
            "energy_tol": tol,
        },
        "converged": f"{converged}/{B * args.steps}",
        "energy_force_evals_per_s": lane_evals / elapsed,
        "mean_flop_per_lane_eval": flops / B,
        "kernel_avg_ms": stage_avg,
        "mixed_vs_fp64": mixed,
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_flow(args) -> int:
    """DOPRI5 gradient flow of 32 lanes at d=3, N=10000, chi=0.45: lane-evaluations/s."""
    from xla_cco import _native, kvectors as K
    from xla_cco.box import cubic_box
    from xla_cco.energy import make_stealthy_energy
    from xla_cco.minimize.flow import FlowSettings

    ctx = _native.context(0)
    L = _native.lib()
    d, n, chi, B = 3, 10000, 0.45, 32
    box = cubic_box(n, d)
    k_c = K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, k_c)
    M = int(ks.n.shape[0])
    energy = make_stealthy_energy(box, ks)
    s0 = np.ascontiguousarray(np.random.default_rng(41).uniform(size=(B, n, d)))
    fl = FlowSettings()
    fs = _native.FlowSettingsC(fl.rtol, fl.atol, fl.dt_init, fl.dt_min, fl.dt_max, fl.t_max, fl.energy_tol,
                               fl.grad_tol, int(fl.max_steps), 0, 0, 0)
    bas = np.ascontiguousarray(box.basis)

    def run(mixed: bool):
        """Timed evaluation steps of a flow session (split precision: lanes turn FP64 at Phi <= 1e-8)."""
        h = C.c_void_p()
        _native.check(L.cco_flow_create(ctx.handle, energy.handle(), B, n, d, C.byref(fs), _native.dptr(bas),
                                        C.byref(h)))
        try:
            if mixed:
                _native.check(L.cco_relax_set_precision(h, 1, 1e-8))
            _native.check(L.cco_flow_init(h, _native.dptr(s0), None))
            na = C.c_int(0)
            _native.check(L.cco_relax_step(h, args.warmup, C.byref(na)))  # evaluations, lanes resident
            ctx.synchronize()
            _native.check(L.cco_relax_set_profiling(h, 4))
            t0 = time.perf_counter()
            _native.check(L.cco_relax_step(h, args.steps, C.byref(na)))
            ctx.synchronize()
            wall = time.perf_counter() - t0
            ms = (C.c_double * len(STAGES))()
            launches = (C.c_int64 * len(STAGES))()
            _native.check(L.cco_relax_profile(h, ms, launches))
            return wall, ms, launches, int(na.value)
        finally:
            L.cco_relax_destroy(h)

    wall, ms, launches, running = run(False)
    mixed = None
    if not args.no_mixed:
        wall_m, _, _, _ = run(True)
        mixed = {"value": B * args.steps / wall_m, "vs_fp64": wall / wall_m}
    stage_avg = {STAGES[i]: ms[i] / launches[i] for i in range(len(STAGES)) if launches[i] > 0}
    if "lbfgs_advance_kernel" in stage_avg:  # the advance stage of a flow session
        stage_avg["flow_advance_kernel"] = stage_avg.pop("lbfgs_advance_kernel")
    lane_evals = B * args.steps
    flop = 8.0 * M * n * (1 + d) * lane_evals
    line = {
        "metric": "gradient-flow lane-evaluations/s (DOPRI5 stages, one energy+force evaluation each)",
        "value": lane_evals / wall,
        "unit": "lane-evaluations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"DOPRI5 gradient flow (rtol 1e-6, atol 1e-14, time parametrisation), {B} lanes, "
                               f"d={d} N={n} chi={chi} (M={M}), uniform random positions", "lanes": B, "M": M,
                   "running_after": running},
        "algorithmic_tflops": flop / wall / 1e12,
        # the 8 M N (1 + d) convention counts about twice the flops the factorised force
        # scan executes, so this is a rate of algorithmic work, not FP64 utilisation
        # (rocprofv3 instruction counts give the executed FP64 work)
        "algorithmic_frac_of_fp64_peak": flop / wall / 1e12 / PEAK_FP64_TFLOPS,
        "kernel_avg_ms": stage_avg,
        "mixed": mixed,
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_md_drift(args) -> int:
    """The long NVE run of the large MD shape: split-precision flow, ground states, then
    args.md_steps velocity-Verlet steps in FP64 and in split precision from the same
    states; energy drift of each and their difference."""
    from xla_cco import _native, kvectors as K
    from xla_cco.box import cubic_box
    from xla_cco.dynamics import MDSettings, Waves, init_state, prng_key
    from xla_cco.dynamics.run import MDSession
    from xla_cco.energy import make_stealthy_energy
    from xla_cco.minimize import LBFGSSettings, relax
    from xla_cco.minimize.flow import FlowSettings, gradient_flow_batch

    ctx = _native.context(0)
    d, n, chi, B = 3, 10000, 0.45, 32
    box = cubic_box(n, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d))
    M = int(ks.n.shape[0])
    energy = make_stealthy_energy(box, ks)
    s0 = np.random.default_rng(41).uniform(size=(B, n, d))

    def note(msg: str) -> None:
        print(msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    fl = gradient_flow_batch(energy, s0, box.basis, FlowSettings(max_steps=args.flow_steps, precision="mixed"))
    t_flow = time.perf_counter() - t0
    note(f"split flow: {args.flow_steps} steps in {t_flow:.1f} s, mean Phi {fl.energy.mean():.4e}")
    t0 = time.perf_counter()
    gs = relax(energy, fl.s, LBFGSSettings(energy_tol=1e-20, precision="mixed"))
    t_relax = time.perf_counter() - t0
    ok = int(np.sum(gs.energy < 1e-20))
    note(f"split relax to E < 1e-20: {ok}/{B} in {t_relax:.1f} s")
    r0 = gs.s @ box.basis
    v0 = np.random.default_rng(42).normal(scale=np.sqrt(1e-4), size=r0.shape)
    kin0 = 0.5 * np.sum(v0 * v0, axis=(1, 2))
    nv = np.asarray(ks.n, dtype=np.float64)
    waves = Waves(n_vectors=nv, kvec=nv @ (2.0 * np.pi * np.linalg.inv(box.basis).T),
                  constrained=np.ones(M, bool), n_constrained=M)
    keys = np.stack([prng_key(i) for i in range(B)])
    chunk = 10_000
    out = {}
    for prec in ("fp64", "mixed"):
        st = MDSettings(dt=0.01, remove_com=False, drift_tol=float("inf"), precision=prec, chunk_steps=chunk)
        lanes = init_state(energy, waves, box.basis, r0, v0, 1e-4, keys, st)
        dev = MDSession(energy, waves, box.basis, st, B, n, d)
        dev.put(lanes)
        ctx.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < args.md_steps:
            k = min(chunk, args.md_steps - done)
            dev.run(k)
            done += k
            ctx.synchronize()
            note(f"{prec}: {done} steps, {time.perf_counter() - t0:.1f} s")
        wall = time.perf_counter() - t0
        out[prec] = (wall, dev.get(lanes))
        dev.close()
    (w64, a), (wmx, m) = out["fp64"], out["mixed"]
    e_a = a.phi + 0.5 * np.sum(a.v * a.v, axis=(1, 2))
    e_m = m.phi + 0.5 * np.sum(m.v * m.v, axis=(1, 2))
    line = {
        "metric": "NVE MD energy drift over a long run, FP64 vs split-precision forces",
        "value": B * args.md_steps / wmx,
        "unit": "lane-steps/s (split precision)",
        "n_gpus": 1,
        "steps": args.md_steps,
        "warmup": 0,
        "ms_per_step": wmx / args.md_steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 positions/velocities, split-precision (fp16 hi+lo MFMA) forces",
        "data": "synthetic",
        "config": {"workload": f"split-precision DOPRI5 flow ({args.flow_steps} steps) and L-BFGS relax of {B} "
                               f"uniform-random configurations, d={d} N={n} chi={chi} (M={M}), then NVE velocity "
                               f"Verlet from the relaxed states, v ~ N(0, 1e-4), dt=0.01, {args.md_steps} steps",
                   "lanes": B, "M": M},
        "flow": {"seconds": t_flow, "mean_phi": float(fl.energy.mean())},
        "relax": {"seconds": t_relax, "converged": f"{ok}/{B}"},
        "fp64": {"lane_steps_per_s": B * args.md_steps / w64, "ms_per_step": w64 / args.md_steps * 1e3,
                 "max_drift_over_ke0": float(np.max(a.max_drift / kin0)),
                 "final_energy_change_over_ke0": float(np.max(np.abs(e_a - a.e0) / kin0)),
                 "status_ok": int(np.sum(a.status == 0))},
        "mixed": {"lane_steps_per_s": B * args.md_steps / wmx, "ms_per_step": wmx / args.md_steps * 1e3,
                  "vs_fp64": w64 / wmx,
                  "max_drift_over_ke0": float(np.max(m.max_drift / kin0)),
                  "final_energy_change_over_ke0": float(np.max(np.abs(e_m - m.e0) / kin0)),
                  "status_ok": int(np.sum(m.status == 0))},
        "mixed_vs_fp64_final": {"energy_rel": float(np.max(np.abs(e_m - e_a) / np.abs(e_a))),
                                "max_abs_dr": float(np.max(np.abs(m.r - a.r)))},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def bench_md(args) -> int:
    """NVE MD steps of 32 lanes at d=3, N=10000, chi=0.45 (the large MD shape): lane-steps/s."""
    from xla_cco import _native, kvectors as K
    from xla_cco.box import cubic_box
    from xla_cco.dynamics import MDSettings, init_state, make_waves, prng_key
    from xla_cco.dynamics.run import MDSession
    from xla_cco.energy import make_stealthy_energy

    ctx = _native.context(0)
    d, n, chi, B = 3, 10000, 0.45, 32
    box = cubic_box(n, d)
    k_c = K.k_outer_from_ka(K.ka_from_chi(chi, 0.05, d), 0.05, d)
    ks = K.enumerate_kvectors(box.basis, 0.0, k_c)
    M = int(ks.n.shape[0])
    energy = make_stealthy_energy(box, ks)
    # the constrained waves take their rho from the force pass; the others (|k| up to
    # md_track K, every step) from the tracked-wave rho pass
    waves = make_waves(box.basis, k_c, args.md_track * k_c)
    rng = np.random.default_rng(31)
    r0 = rng.uniform(size=(B, n, d)) @ box.basis
    v0 = rng.normal(scale=1e-2, size=(B, n, d))  # T = 1e-4 per component
    keys = np.stack([prng_key(i) for i in range(B)])
    out = {}
    for prec in ("fp64", "mixed"):
        st = MDSettings(dt=0.01, remove_com=False, chunk_steps=10_000, drift_tol=float("inf"), precision=prec,
                        wind_every=1)
        lanes = init_state(energy, waves, box.basis, r0, v0, 1e-4, keys, st)
        # one device session holds the lanes; the timed region is the device steps only
        dev = MDSession(energy, waves, box.basis, st, B, n, d)
        dev.put(lanes)
        dev.run(args.warmup)
        ctx.synchronize()
        t0 = time.perf_counter()
        dev.run(args.steps)
        ctx.synchronize()
        out[prec] = (time.perf_counter() - t0, dev.get(lanes))
        dev.close()
    wall, res = out["fp64"]
    wall_m, res_m = out["mixed"]
    lane_steps = B * args.steps
    m_trk = int(waves.kvec.shape[0]) - int(waves.n_constrained)  # waves summed by the tracked pass
    flop = 8.0 * (M * n * (1 + d) + m_trk * n) * lane_steps
    e0 = res.phi + 0.5 * np.sum(res.v * res.v, axis=(1, 2))
    e0m = res_m.phi + 0.5 * np.sum(res_m.v * res_m.v, axis=(1, 2))
    line = {
        "metric": "MD lane-steps/s (one energy+force evaluation per lane-step)",
        "value": lane_steps / wall,
        "unit": "lane-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"NVE velocity Verlet, {B} lanes, d={d} N={n} chi={chi} (M={M}), dt=0.01, "
                               "uniform random positions, v ~ N(0, 1e-4)", "lanes": B, "M": M,
                   "tracked_waves": int(waves.kvec.shape[0]), "k_track_over_K": args.md_track},
        "algorithmic_tflops": flop / wall / 1e12,
        # the 8 M N (1 + d) convention counts about twice the flops the factorised force
        # scan executes, so this is a rate of algorithmic work, not FP64 utilisation
        # (rocprofv3 instruction counts give the executed FP64 work)
        "algorithmic_frac_of_fp64_peak": flop / wall / 1e12 / PEAK_FP64_TFLOPS,
        "mixed": {"value": lane_steps / wall_m, "vs_fp64": wall / wall_m,
                  "energy_diff_rel": float(np.max(np.abs(e0m - e0) / np.abs(e0)))},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lanes", type=int, default=256)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--d", type=int, default=2)
    ap.add_argument("--chi", type=float, default=0.45)
    ap.add_argument("--cpu-seconds", type=float, default=60.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel event timing")
    ap.add_argument(
        "--profile-every",
        type=int,
        default=4,
        help="time the kernels of every k-th evaluation step with HIP events (events between "
        "launches cost device time; 1 times every step)",
    )
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="mixed: split-precision MFMA evaluations while E > --switch, FP64 below")
    ap.add_argument("--switch", type=float, default=1e-8, help="energy below which mixed lanes turn FP64")
    ap.add_argument("--split-history", type=int, choices=(0, 1), default=0,
                    help="split-precision L-BFGS lanes keep FP64 (s, y) pairs (0) or FP32 pairs (1)")
    ap.add_argument(
        "--advance-variant",
        type=int,
        default=0,
        help="L-BFGS advance: 0 auto, 1 global-memory two-loop, 2 register two-loop, 3 Gram form",
    )
    ap.add_argument(
        "--shards",
        type=int,
        default=0,
        help="lane shards, one stream each (0 auto: 3 at this batch size; 1 = a single stream)",
    )
    ap.add_argument("--graph", type=int, default=0,
                    help="evaluation steps per HIP-graph replay (0: plain launches)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick rho chunk, shards and graph steps with xla_cco.minimize.tuning "
                         "(timed on the first warmup input, outside the timed region)")
    ap.add_argument("--tune-rounding", action="store_true",
                    help="with --autotune, also pick the rho split length and the L-BFGS "
                         "advance variant (results change at rounding level)")
    ap.add_argument(
        "--uneven",
        action="store_true",
        help="time the uneven batch instead: 512 configurations, N ~ U{100..2000}, "
        "d ~ U{2,3}, chi ~ U{0.1,0.2,0.3,0.45}, one session",
    )
    ap.add_argument("--energy", action="store_true",
                    help="time batched energy+force evaluations (FP64 and split precision) at --d/--n/--chi/--lanes")
    ap.add_argument("--md", action="store_true",
                    help="time NVE MD at d=3 N=10000 chi=0.45 with 32 lanes instead (steps = MD steps)")
    ap.add_argument("--flow", action="store_true",
                    help="time the DOPRI5 gradient flow at d=3 N=10000 chi=0.45 with 32 lanes (steps = evaluations)")
    ap.add_argument("--md-drift", action="store_true",
                    help="the long NVE run of the large MD shape: split flow, relax, then --md-steps "
                         "steps in FP64 and split precision; energy drift of each")
    ap.add_argument("--md-steps", type=int, default=100_000)
    ap.add_argument("--flow-steps", type=int, default=60)
    ap.add_argument("--md-track", type=float, default=1.0,
                    help="with --md: track every wave with |k| < md_track K, wound every step "
                         "(1: the constrained waves only; the reference MD command line tracks 2K)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the split-precision relax reported beside the FP64 value")
    ap.add_argument("--no-energy-compare", action="store_true",
                    help="skip the FP64 vs split-precision energy+force evaluation timings")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the reference's published d=3 N=100 x 1200 workload reported beside the value")
    ap.add_argument(
        "--calibrate-copy",
        type=int,
        default=0,
        metavar="BYTES",
        help="first run the library's plain copy kernel over BYTES (HBM counter calibration)",
    )
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU only (asked for {args.gpus}): lanes are independent, "
            "so a multi-GPU run is G copies of this one with the lanes split across ranks.",
            file=sys.stderr,
        )
        return 2

    from xla_cco import _native
    from xla_cco.minimize import LBFGSSettings, RelaxSession

    if args.uneven:
        return bench_uneven(args)
    if args.energy:
        return bench_energy(args)
    if args.flow:
        return bench_flow(args)
    if args.md_drift:
        return bench_md_drift(args)
    if args.md:
        return bench_md(args)

    L = _native.lib()
    ctx = _native.context(0)
    if args.calibrate_copy > 0:
        nbytes = (args.calibrate_copy // 32) * 32
        src = _native.DeviceArray(ctx, (nbytes // 8,))
        dst = _native.DeviceArray(ctx, (nbytes // 8,))
        t = C.c_double(0.0)
        for _ in range(2):
            _native.check(L.cco_copy_dev(ctx.handle, dst.ptr, src.ptr, nbytes, C.byref(t)))
        print(f"copy_kernel {nbytes} bytes: {t.value:.4f} ms, {2 * nbytes / t.value / 1e6:.1f} GB/s",
              file=sys.stderr)
        src.free()
        dst.free()
    d, n, chi, B = args.d, args.n, args.chi, args.lanes
    box, ks, energy = build_workload(d, n, chi)
    M = int(ks.n.shape[0])
    tol = 1e-20
    settings = LBFGSSettings(energy_tol=tol, precision=args.precision, precision_switch=args.switch)
    choice = None
    if args.autotune:
        from xla_cco.minimize import tuning

        choice = tuning.tune_relax(energy, np.random.default_rng(1000).uniform(size=(B, n, d)), settings,
                                   store=False, exact=not args.tune_rounding)
    sess = RelaxSession(energy, B, n, d, settings)
    _native.check(L.cco_relax_set_option(sess.handle, 1, args.advance_variant))
    _native.check(L.cco_relax_set_option(sess.handle, 2, args.shards))
    _native.check(L.cco_relax_set_option(sess.handle, 3, args.graph))
    _native.check(L.cco_relax_set_option(sess.handle, 5, args.split_history))
    if choice is not None:
        tuning.apply(sess, choice)
        args.shards, args.graph = choice.shards, choice.graph
    bound = 1 + 3 * (settings.max_evals_per_pass + settings.max_linesearch_steps + 1)

    inputs = [
        _native.DeviceArray.from_host(ctx, np.random.default_rng(1000 + i).uniform(size=(B, n, d)))
        for i in range(args.warmup + args.steps)
    ]
    out_s = _native.DeviceArray(ctx, (B, n, d))
    out_e = _native.DeviceArray(ctx, (B,))
    out_g = _native.DeviceArray(ctx, (B,))
    out_i = [_native.DeviceArray(ctx, (B,), np.int64) for _ in range(3)]

    def one_step(dev_s0) -> None:
        _native.check(L.cco_relax_init_dev(sess.handle, dev_s0.ptr))
        _native.check(L.cco_relax_run(sess.handle, bound, 256))
        _native.check(
            L.cco_relax_result_dev(
                sess.handle, out_s.ptr, out_e.ptr, out_g.ptr, out_i[0].ptr, out_i[1].ptr, out_i[2].ptr
            )
        )

    for w in range(args.warmup):
        one_step(inputs[w])
    ctx.synchronize()

    def timed(shards: int, first_input: int, nsteps: int) -> dict:
        """nsteps timed relaxes with `shards` lane shards; totals, counters and stage times."""
        _native.check(L.cco_relax_set_option(sess.handle, 2, shards))
        _native.check(L.cco_relax_set_profiling(sess.handle, 0 if args.no_profile else args.profile_every))
        r = {"converged": 0, "lane_evals": 0, "iters": 0, "pairs": 0, "launches": 0, "elapsed": 0.0}
        for k in range(nsteps):
            ctx.synchronize()
            t0 = time.perf_counter()
            one_step(inputs[first_input + k])
            ctx.synchronize()
            r["elapsed"] += time.perf_counter() - t0
            e = out_e.to_host()
            r["converged"] += int(np.sum(e < tol))
            cnt = (C.c_int64 * 4)()
            _native.check(L.cco_relax_counters(sess.handle, cnt))
            r["lane_evals"] += cnt[0]
            r["iters"] += cnt[1]
            r["pairs"] += cnt[2]
            r["launches"] += cnt[3]
        opt = C.c_int64(0)
        _native.check(L.cco_relax_get_option(sess.handle, 1, C.byref(opt)))
        r["variant"] = int(opt.value)
        _native.check(L.cco_relax_get_option(sess.handle, 2, C.byref(opt)))
        r["shards"] = int(opt.value)
        ms = (C.c_double * len(STAGES))()
        launches = (C.c_int64 * len(STAGES))()
        _native.check(L.cco_relax_profile(sess.handle, ms, launches))
        r["ms"], r["timed_launches"] = list(ms), list(launches)
        return r

    nd = n * d

    def roofline_of(r: dict):
        """Per-kernel averages and the roofline of the kernel with the largest total time."""
        if args.no_profile or r["timed_launches"][0] == 0:
            return {}, None
        ms, launches = r["ms"], r["timed_launches"]
        # algorithmic work per launch of each kernel
        per_step = 1.0 / max(r["launches"], 1)
        lane_evals_per_launch = r["lane_evals"] * per_step
        gram = r["variant"] == 3
        # the FP64 energy kernels run on the vector ALUs (FP64 VALU peak = FP64 matrix peak);
        # their work is the algorithmic 8 M N (1 + d) convention, which counts more flops
        # than the factorised kernels execute (see the instruction-count profiles)
        work = {
            "rho_kernel": ("fp64", 8.0 * M * n * lane_evals_per_launch),
            "force_kernel": ("fp64", 8.0 * M * n * d * lane_evals_per_launch),
            # fused two-loop: s_i, y_i read twice per visited pair; Gram form: line search only
            "lbfgs_advance_kernel": (
                "hbm",
                ((0.0 if gram else 32.0 * nd * r["pairs"]) + 8.0 * nd * (3 * r["lane_evals"] + 12 * r["iters"]))
                * per_step,
            ),
            # Gram form: one read of s_i, y_i per pair in each of the dot and direction passes
            "lbfgs_gram_dot_kernel": ("hbm", (16.0 * nd * r["pairs"] + 24.0 * nd * r["iters"]) * per_step),
            "lbfgs_gram_comb_kernel": ("hbm", (16.0 * nd * r["pairs"] + 48.0 * nd * r["iters"]) * per_step),
        }
        stage_avg = {STAGES[i]: ms[i] / launches[i] for i in range(len(STAGES)) if launches[i] > 0}
        dominant = max((i for i in range(len(STAGES)) if launches[i] > 0), key=lambda i: ms[i])
        name = STAGES[dominant]
        avg_ms = stage_avg[name]
        bound_kind, per_launch = work.get(name, ("hbm", 0.0))
        if bound_kind == "fp64":
            achieved = per_launch / (avg_ms * 1e-3) / 1e12
            peak, unit = PEAK_FP64_TFLOPS, "TFLOP/s"
        else:
            achieved = per_launch / (avg_ms * 1e-3) / 1e9
            peak, unit = PEAK_HBM_GBS, "GB/s"
        traffic = None
        tfile = ROOT / "profiles" / "traffic.json"
        if tfile.exists():
            try:
                tdat = json.loads(tfile.read_text()).get(f"d{d}_n{n}_chi{chi}_b{B}_s{r['shards']}", {})
                traffic = tdat.get(name, {}).get("bytes_per_launch")
            except (ValueError, AttributeError):
                traffic = None
        return stage_avg, {
            "kernel": name,
            "bound": bound_kind,
            "achieved": achieved,
            "peak": peak,
            "unit": unit,
            "frac": achieved / peak,
            "traffic": traffic,
            # HBM bytes per launch are not counted in this run: they come from rocprofv3
            # --pmc FETCH_SIZE / WRITE_SIZE runs of this workload, scaled by the same
            # counters on a plain copy kernel, committed as profiles/traffic.json
            "traffic_source": "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; not this run)"
            if traffic is not None else None,
            "avg_launch_ms": avg_ms,
            "work_per_launch": per_launch,
            # lane shards run on separate streams: a launch covers 1/streams of the lanes
            # and its event-timed duration includes waiting for and sharing the CUs with
            # the other shards' kernels
            "streams": r["shards"],
        }

    main_run = timed(args.shards, args.warmup, args.steps)
    converged, lane_evals, elapsed = main_run["converged"], main_run["lane_evals"], main_run["elapsed"]
    variant_run, shards_run = main_run["variant"], main_run["shards"]
    value = converged / elapsed
    evals_per_conv = lane_evals / max(converged, 1)
    stage_avg, roofline = roofline_of(main_run)
    # the same kernels alone on one stream (one more timed relax of the first input)
    roofline_one = None
    if roofline is not None and shards_run > 1:
        _, roofline_one = roofline_of(timed(1, args.warmup, 1))

    # the metric's comparison: the same relax with split-precision evaluations
    # (lanes turn FP64 at E <= switch), one warm and one timed relax of the first
    # timed input, in a session of its own
    mixed = None
    if args.precision == "fp64" and not args.no_mixed:
        ms_ = RelaxSession(energy, B, n, d, LBFGSSettings(energy_tol=tol, precision="mixed",
                                                           precision_switch=args.switch))
        _native.check(L.cco_relax_set_option(ms_.handle, 2, args.shards))
        _native.check(L.cco_relax_set_option(ms_.handle, 5, args.split_history))
        for _ in range(2):
            ctx.synchronize()
            t0 = time.perf_counter()
            _native.check(L.cco_relax_init_dev(ms_.handle, inputs[args.warmup].ptr))
            _native.check(L.cco_relax_run(ms_.handle, bound, 256))
            _native.check(L.cco_relax_result_dev(ms_.handle, out_s.ptr, out_e.ptr, out_g.ptr, out_i[0].ptr,
                                                 out_i[1].ptr, out_i[2].ptr))
            ctx.synchronize()
            t_mixed = time.perf_counter() - t0
        conv_m = int(np.sum(out_e.to_host() < tol))
        mixed = {"value": conv_m / t_mixed, "unit": "configs/s", "converged": f"{conv_m}/{B}",
                 "vs_fp64": (conv_m / t_mixed) / value, "switch": args.switch,
                 "split_history": "fp32" if args.split_history else "fp64",
                 "note": "split-precision rho and forces (fp16 hi+lo MFMA) and FP32 L-BFGS pairs until "
                         "E <= switch, then FP64" if args.split_history else
                         "split-precision rho and forces (fp16 hi+lo MFMA) until E <= switch, then FP64"}
        ms_.close()

    mixed_bound = None
    if not args.no_profile and sum(main_run["ms"]) > 0:
        energy_ms = sum(main_run["ms"][STAGES.index(k)] for k in ("rho_kernel", "reduce_kernel", "force_kernel"))
        mixed_bound = sum(main_run["ms"]) / max(sum(main_run["ms"]) - energy_ms, 1e-12)

    # energy+force evaluations alone, FP64 against split precision, at the relax
    # workload's shape and at the MD / flow lanes' shape (d=3 N=10^4 chi=0.45 x32)
    evals = None
    if not args.no_energy_compare:
        evals = {}
        for key, (ed, en, echi, eB) in (("relax_shape", (d, n, chi, B)), ("d3_n10000_b32", (3, 10000, 0.45, 32))):
            tt, eM, ok = energy_times(ed, en, echi, eB, 10, 2)
            evals[key] = {"workload": f"value_and_grad of {eB} configurations, d={ed} N={en} chi={echi} (M={eM})",
                          "fp64_ms": tt["fp64"] * 1e3, "mixed_ms": tt["mixed"] * 1e3,
                          "mixed_vs_fp64": tt["fp64"] / tt["mixed"] if ok else None}

    published = None
    if not args.no_published:
        published = {f"chi{c}": published_shape(c, autotune=args.autotune) for c in (0.35, 0.45)}
        published["relax_rates"] = published_relax_rates()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ks, box.volume, d, n, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched FP64 L-BFGS relax to E<1e-20, d={d} N={n} chi={chi} (M={M}), {B} uniform-random configurations",
            "d": d,
            "N": n,
            "chi": chi,
            "M": M,
            "lanes": B,
            "method": "lbfgs (memory 100, rescales 1e4/1/1e-2, zoom line search)",
            "advance": {1: "two-loop, global memory", 2: "two-loop, register-resident",
                        3: "Gram form (dot / recursion / direction kernels)"}.get(variant_run),
            "energy_tol": tol,
            "precision": args.precision if args.precision == "fp64" else f"mixed (FP64 below E={args.switch:g})",
            "streams": shards_run,
            "graph_steps": args.graph,
            "autotune": None if choice is None else dataclasses.asdict(choice),
        },
        "converged": f"{converged}/{B * args.steps}",
        "energy_force_evals_per_s": lane_evals / elapsed,
        "evals_per_converged": evals_per_conv,
        "kernel_avg_ms": stage_avg,
        "roofline": roofline,
        "roofline_one_stream": roofline_one,
        "mixed_vs_fp64": mixed,
        # the split-precision relax can gain at most 1 / (1 - energy share of a step) here:
        # the FP64 L-BFGS advance (the dominant kernel) is the same on both paths
        "mixed_bound_note": "the L-BFGS advance and compaction are FP64 on both paths, so free energy "
                            "evaluations would speed the relax up by at most mixed_speedup_bound "
                            "(event-timed kernel time of a step / its non-energy part)",
        "mixed_speedup_bound": mixed_bound,
        "energy_evals_fp64_vs_mixed": evals,
        "published_shape": published,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    sess.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
