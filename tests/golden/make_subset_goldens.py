"""Write golden vectors for sort_edges over a sampled edge set E_s
(graphlet_core.cpp:836-851) by running oracle/_ref/ref_subset -- a harness
that calls the reference library, compiled from the reference sources by
oracle/Makefile (`make -C oracle ref`) in this container.

Each case records E_s (the sampled edge ids), the 17 totals and the .micro
rows (one per E_ordered position) in tests/golden/subset/<case>.npz, indexed
by tests/golden/goldens_subset.json.

Usage: python tests/golden/make_subset_goldens.py
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_subset")

import pgd_amd  # noqa: E402
from tests.test_oracle import GOLDENS, golden_graph_path  # noqa: E402

GRAPHS = ["DD21.mtx", "sample_graph.csv", "gnp_2000_0.01", "ba_3000_4", "4-clique-and-edge.txt"]
# (variant, flags, how E_s is drawn)
VARIANTS = [
    ("es10", [], "frac10"),
    ("es10_o_kcore_s2l", ["-o", "kcore", "--s2l"], "frac10"),
    ("es30_o_natural", ["-o", "natural"], "frac30"),
    ("es_dups_o_deg_vol", ["-o", "deg_vol"], "dups"),
    ("es_all_shuffled_n_kcore", ["-n", "kcore"], "all"),
]


def draw(m: int, how: str, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    if how == "frac10":
        return rng.choice(m, size=max(1, m // 10), replace=False).astype(np.int32)
    if how == "frac30":
        return rng.choice(m, size=max(1, (3 * m) // 10), replace=False).astype(np.int32)
    if how == "dups":  # ids may repeat in the reference's E_s
        return rng.integers(0, m, size=max(1, m // 5)).astype(np.int32)
    return rng.permutation(m).astype(np.int32)


def main() -> None:
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out_dir = os.path.join(HERE, "subset")
    os.makedirs(out_dir, exist_ok=True)
    index = {}
    for gi, name in enumerate(GRAPHS):
        m = GOLDENS[name]["m"]
        for vi, (vname, flags, how) in enumerate(VARIANTS):
            tmp = tempfile.mkdtemp(prefix="pgd_sub_")
            try:
                path = golden_graph_path(name, GOLDENS[name], tmp)
                es = draw(m, how, 1000 * gi + vi)
                with open(os.path.join(tmp, "es.txt"), "w") as f:
                    f.write("\n".join(str(int(x)) for x in es) + "\n")
                r = subprocess.run([HARNESS, "es.txt", "-f", os.path.basename(path), "-c", "x.macro", "-m",
                                    "x.micro", "-w", "4", *flags], capture_output=True, text=True, cwd=tmp,
                                   timeout=600)
                if r.returncode != 0:
                    raise RuntimeError(f"{name} {vname}: {r.stderr[-500:]}")
                totals = {}
                with open(os.path.join(tmp, "x.macro")) as f:
                    for line in f:
                        k, v = line.split(" = ")
                        totals[k.strip()] = int(v)
                with open(os.path.join(tmp, "x.micro"), "rb") as f:
                    rows = [ln.split(b",") for ln in f.read().splitlines()[1:] if ln]
                micro = np.array([[int(x) for x in row] for row in rows], dtype=np.uint64).reshape(-1, 10)
                key = f"{name}@{vname}"
                fn = f"{key.replace('@', '__')}.npz"
                np.savez_compressed(os.path.join(out_dir, fn), E_s=es, micro=micro)
                index[key] = {"graph": name, "flags": flags, "draw": how, "count": int(len(es)),
                              "totals": [totals[k] for k in pgd_amd.TOTAL_NAMES], "file": "subset/" + fn}
                print(key, len(es), totals["total_3_tris"], flush=True)
            finally:
                shutil.rmtree(tmp)
    with open(os.path.join(HERE, "goldens_subset.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
