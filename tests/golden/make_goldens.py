"""Write golden vectors by running the reference `pgd`, compiled from the
reference sources by oracle/Makefile (`make -C oracle ref`), in this container.

For every case it records the 17 totals of the .macro file, the printed K,
and the .micro rows (src,dst and the 8 per-edge columns, in E_ordered order)
-- as a compressed array for small graphs, as a SHA-256 of the file text for
large ones.  Inputs are either data files (copied here compressed: the
reference's own fixtures and small hand-made reader cases) or seeded graphs
from pgd_amd's generators (re-generated by the tests, checked by edge-list
hash).

Usage: python tests/golden/make_goldens.py [--only NAME ...]
"""
from __future__ import annotations

import argparse
import gzip
import hashlib
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
REF_SRC = "/root/reference"
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "pgd")

import pgd_amd  # noqa: E402  (generators only: host code)
from tests.golden.graphs import edge_hash, gen_cases, make_graph  # noqa: E402

TOTAL_NAMES = pgd_amd.TOTAL_NAMES

# hand-made reader cases: (file name, text)
READER_CASES = {
    "rd_csv_header.csv": "src,dst\n1,2\n2,3\n3,1\n3,4\n4,5\n5,3\n",
    "rd_one_based.txt": "1 2\n2 3\n3 4\n4 1\n1 3\n2 4\n4 5\n",
    "rd_zero_based.txt": "0 1\n1 2\n2 0\n2 3\n3 4\n4 2\n",
    "rd_dups_loops.txt": "1 2\n2 1\n2 3\n3 3\n1 2\n3 4\n4 1\n1 3\n4 4\n2 4\n",
    "rd_tab.edges": "# comment\n%another\n5\t7\n7\t9\n9\t5\n5\t11\n11\t7\n9\t11\n11\t13\n",
    "rd_weighted.txt": "1 2 0.5\n2 3 1.0\n3 1 2.0\n1 4 3.0\n4 2 1.0\n",
    "rd_triple_space.txt": "1   2\n2   3\n3   1\n3   4\n",
    "rd_comment_mid.txt": "1 2\n% mid comment\n2 3\n\n3 1\n1 4\n",
    "rd_gap_ids.mtx": "%%MatrixMarket matrix coordinate pattern symmetric\n20 20 7\n2 5\n5 9\n9 2\n2 12\n12 5\n9 12\n12 14\n",
    "rd_mtx_dups.mtx": "%%MatrixMarket matrix coordinate pattern symmetric\n6 6 8\n1 2\n2 3\n3 1\n2 1\n3 4\n4 4\n4 5\n5 3\n",
    "rd_mtx_comments.mtx": "%%MatrixMarket matrix coordinate pattern symmetric\n% c\n% d\n5 5 6\n1 2\n2 3\n% inner\n3 1\n3 4\n4 5\n5 3\n",
    # fields a line leaves unset keep the previous line's values
    "rd_carry_tokens.txt": "1 2\n3 \n4 5\n 6\n7 8 9\n3\n5 6\n2 3\n6 1\n 8\n9 \n",
    "rd_carry_csv.csv": "a,b\n1,2\n,3\n4,\n5,6\n2,,3\n6,1\n,7\n8,\n",
    "rd_carry_mtx.mtx": "%%MatrixMarket matrix coordinate pattern symmetric\n6 6 5\n1 2\n3\n4 5\n6\n2 6\n",
    "rd_blank_first.txt": "\n# not a comment block any more\n1 2\n2 3\n3 1\n3 4\n",
}

# reference fixtures copied (compressed) as inputs
FIXTURES = ["data/4-clique.txt", "data/4-clique.edges", "data/4-chordal-cycle.txt", "data/4-clique-and-edge.txt",
            "data/DD21.mtx", "sample_graph.csv", "data/C250-9.mtx", "data/C500-9.mtx"]


def run_ref(path: str, workdir: str, extra: list[str], threads: int) -> dict:
    macro = os.path.join(workdir, "out.macro")
    micro = os.path.join(workdir, "out.micro")
    for f in (macro, micro):
        if os.path.exists(f):
            os.remove(f)
    cmd = [REF_BIN, "-f", path, "-c", macro, "-m", micro, "-w", str(threads), *extra]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=workdir)
    if r.returncode != 0:
        raise RuntimeError(f"reference failed on {path}: {r.stderr[-500:]}")
    totals = {}
    with open(macro) as f:
        for line in f:
            k, v = line.split(" = ")
            totals[k.strip()] = int(v)
    K = None
    for line in r.stdout.splitlines():
        if line.startswith("K = "):
            K = int(line[4:])
    with open(micro, "rb") as f:
        text = f.read()
    rows = [ln.split(b",") for ln in text.splitlines()[1:] if ln]
    arr = np.array([[int(x) for x in row] for row in rows], dtype=np.uint64).reshape(-1, 10)
    return dict(totals=[totals[k] for k in TOTAL_NAMES], K=K, micro=arr,
                micro_sha256=hashlib.sha256(text).hexdigest())


VARIANTS = {"": [], "o_kcore": ["-o", "kcore"], "s2l": ["--s2l"], "n_s2l": ["--s2l_neigh"],
            "o_natural": ["-o", "natural"], "n_kcore": ["-n", "kcore"], "o_deg_vol": ["-o", "deg_vol"],
            "r_csc": ["-r", "csc"], "r_adj": ["-r", "adj"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    if not os.path.exists(REF_BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    index_path = os.path.join(HERE, "goldens.json")
    index = json.load(open(index_path)) if os.path.exists(index_path) else {}
    work = tempfile.mkdtemp(prefix="pgd_golden_")

    def want(name):
        return not args.only or name in args.only

    def record(name, path, variants, meta, keep_micro_rows=25000):
        for vname, flags in variants.items():
            key = name + ("@" + vname if vname else "")
            if not want(key) and not want(name):
                continue
            res = run_ref(path, work, flags, args.threads)
            ent = dict(meta, flags=flags, totals=res["totals"], K=res["K"], m=int(res["micro"].shape[0]),
                       micro_sha256=res["micro_sha256"])
            if res["micro"].shape[0] <= keep_micro_rows:
                fn = f"micro_{key.replace('@', '__')}.npz"
                np.savez_compressed(os.path.join(HERE, fn), micro=res["micro"])
                ent["micro_file"] = fn
            index[key] = ent
            print(key, res["totals"][:4], "K", res["K"], flush=True)

    # reference fixtures and hand-made reader files
    for rel in FIXTURES:
        name = os.path.basename(rel)
        src = os.path.join(REF_SRC, rel)
        gz = os.path.join(HERE, "inputs", name + ".gz")
        with open(src, "rb") as f, open(gz, "wb") as raw, \
                gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as g:  # byte-stable output
            g.write(f.read())
        local = os.path.join(work, name)
        shutil.copy(src, local)
        variants = {"": []} if name not in ("DD21.mtx", "sample_graph.csv") else VARIANTS
        if name.startswith("C"):
            variants = {"": [], "r_csc": ["-r", "csc"]}
        record(name, local, variants, dict(input="inputs/" + name + ".gz"))
    for name, text in READER_CASES.items():
        path = os.path.join(HERE, "inputs", name)
        with open(path, "w") as f:
            f.write(text)
        local = os.path.join(work, name)
        shutil.copy(path, local)
        record(name, local, {"": []}, dict(input="inputs/" + name))
    # seeded synthetic graphs (written as .mtx for the reference)
    for name, kind, prm in gen_cases() + [("bigid_2p24", "bigid", dict(n=1 << 24, leaves=30000, pairs=120000))]:
        if not want(name):
            continue
        n, s, d = make_graph(kind, prm)
        path = os.path.join(work, name + ".mtx")
        pgd_amd.write_mtx(path, n, s, d)
        variants = {"": []}
        if name in ("gnp_2000_0.01", "rmat_12", "ba_3000_4"):
            variants = {"": [], "o_kcore": ["-o", "kcore"], "s2l": ["--s2l"], "n_s2l": ["--s2l_neigh"]}
        record(name, path, variants, dict(gen=kind, params=prm, edges_sha256=edge_hash(s, d)))
    with open(index_path, "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)
    shutil.rmtree(work)


if __name__ == "__main__":
    main()
