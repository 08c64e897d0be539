"""Extract the C++ energy fixture from the reference's own test data.

Reads tests/data/md_cxx_reference.json and tests/data/md_cxx_samples.npz of
the sdawley1/xla-cco tree (data files only; no reference code is imported or
run) and writes ``md_cxx_phi.npz`` beside this script: the 15 stored
configurations (N=256, d=2), their basis, and for each the C++ energy
``phi_cpp`` (printed at %.17g by the C++ harness), plus the k-set metadata
(K, its hex bits, Ka, phi, the 50 n-vectors).

Usage: python tests/golden/make_md_cxx_fixture.py [/path/to/reference]
"""

import json
import sys
from pathlib import Path

import numpy as np

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
data = json.loads((ref / "tests/data/md_cxx_reference.json").read_text())
with np.load(ref / "tests/data/md_cxx_samples.npz") as z:
    s, basis, seed, sample = z["s"], z["basis"], z["seed"], z["sample"]
runs = {r["seed"]: r for r in data["runs"]}
phi = np.array([runs[int(a)]["phi_cpp"][int(b)] for a, b in zip(seed, sample)])
out = Path(__file__).with_name("md_cxx_phi.npz")
np.savez(
    out,
    s=s,
    basis=basis,
    phi_cpp=phi,
    K=np.float64(data["K"]),
    K_hex=np.array(data["K_hex"]),
    Ka=np.float64(data["Ka"]),
    phi_fake=np.float64(data["phi_fake"]),
    n_vectors=np.array(data["n_vectors"], dtype=np.int64),
    N=np.int64(data["N"]),
    d=np.int64(data["d"]),
    M=np.int64(data["M"]),
)
print(f"wrote {out} ({s.shape[0]} configs)")
