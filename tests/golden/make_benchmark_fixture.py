"""Extract the reference's published benchmark outcomes into a small JSON fixture.

Reads the data files benchmarks/results_cpu.json, results_gpu.json and
results_threshold_gpu.json of the sdawley1/xla-cco tree (no reference code is
imported or run) and writes ``benchmark_outcomes.json`` beside this script.

Every row was produced by the reference on inputs that NumPy reproduces exactly,
``np.random.default_rng(0).random((B, N, d))`` (benchmarks/bench_energy.py:74-92,
bench_threshold.py:46-63), on a cubic unit-density box with the half-space k-set
at packing fraction 0.05:

* ``value_and_grad``: mean energy and max |dPhi/ds| over B lanes;
* ``descent_400``: mean final energy of ``relax(.., DescentSettings(max_evals=400))``;
* ``relax_lbfgs``: total n_evals and mean final energy of
  ``relax(.., LBFGSSettings(max_evals_per_pass=200))``, for the reference's trig
  path and its phase-table path (their gap is the reference's own round-off spread).

Usage: python tests/golden/make_benchmark_fixture.py [/path/to/reference]
"""

import json
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "benchmarks"
rows = []
for name in ("results_cpu.json", "results_gpu.json"):
    data = json.loads((ref / name).read_text())
    for r in data["results"]:
        if r["variant"] not in ("baseline", "tables"):
            continue
        row = {
            "source": f"benchmarks/{name}",
            "path": "trig" if r["variant"] == "baseline" else "tables",
            "kind": r["kind"],
            "d": r["d"],
            "n": r["n"],
            "chi": r["chi"],
            "m": r["m"],
            "batch": r["batch"] if r["kind"] == "value_and_grad" else r["relax_batch"],
            "energy_mean": r["energy_mean"],
        }
        if r["kind"] == "value_and_grad":
            row["grad_absmax"] = r["grad_absmax"]
        else:
            row["n_evals"] = r["n_evals"]
        rows.append(row)
thr = json.loads((ref / "results_threshold_gpu.json").read_text())
for r in thr["rows"]:
    for path in ("trig", "tables"):
        rows.append({
            "source": "benchmarks/results_threshold_gpu.json",
            "path": path,
            "kind": "relax_lbfgs",
            "d": 3,
            "n": 100,
            "chi": r["chi"],
            "m": r["m"],
            "batch": thr["batch"],
            "n_evals": r[path]["n_evals"],
        })
out = Path(__file__).with_name("benchmark_outcomes.json")
out.write_text(json.dumps({"phi": 0.05, "seed": 0, "max_evals_per_pass": 200,
                           "descent_max_evals": 400, "rows": rows}, indent=1) + "\n")
print(f"wrote {out} ({len(rows)} rows)")
