"""Pin the CPU oracle before trusting it (CPU only).

Each check is independent of the oracle's own formulas:
  * golden vectors written by the compiled reference (tests/golden);
  * the per-edge columns stored in the reference's sample_graph.csv;
  * known answers for the reference's fixtures (exact-integer values);
  * brute force over every vertex subset of small random graphs;
  * closed forms (cliques, stars) and the reference's own 10 identities
    (graphlet_core.cpp:1276-1285).
"""
from __future__ import annotations

import gzip
import hashlib
import json
import os
import random
import shutil
import tempfile

import numpy as np
import pytest

from tests import brute
from tests.oracle_lib import TOTAL_NAMES, OracleGraph, derive, run_pipeline

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
GOLDENS = json.load(open(os.path.join(GOLD, "goldens.json"))) if os.path.exists(os.path.join(GOLD, "goldens.json")) else {}

# Exact totals for the reference's fixtures (printed order), computed by an
# independent exact-integer evaluation of the paper's Lemmas and by 4-subset
# brute force for the tiny graphs; n as the reference reader counts it.
KNOWN = {
    "data/4-clique.txt": [6, 0, 4, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
    "data/4-chordal-cycle.txt": [5, 1, 2, 2, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0],
    "data/4-clique-and-edge.txt": [7, 3, 4, 3, 3, 0, 1, 0, 3, 0, 0, 0, 0, 0, 0, 1, 0],
    "data/DD21.mtx": [14267, 16502611, 9286, 35397, 81879530, 31553402783, 2116, 9925, 41154, 1267, 13838, 88568,
                      234712803384, 101544802, 203029889, 53278602, 45201167584460],
    "sample_graph.csv": [14267, 16502611, 9286, 35397, 81879530, 31553402783, 2116, 9925, 41154, 1267, 13838, 88568,
                         234712803384, 101544802, 203029889, 53278602, 45201167584460],
}


def input_path(rel: str, tmp: str) -> str:
    """Materialise a golden input (gzip-compressed data files are expanded)."""
    src = os.path.join(GOLD, rel)
    if rel.endswith(".gz"):
        out = os.path.join(tmp, os.path.basename(rel)[:-3])
        with gzip.open(src, "rb") as f, open(out, "wb") as g:
            g.write(f.read())
        return out
    out = os.path.join(tmp, os.path.basename(rel))
    shutil.copy(src, out)
    return out


def golden_graph_path(name: str, ent: dict, tmp: str) -> str:
    if "input" in ent:
        return input_path(ent["input"], tmp)
    import pgd_amd
    from tests.golden.graphs import edge_hash, make_graph
    n, s, d = make_graph(ent["gen"], ent["params"])
    assert edge_hash(s, d) == ent["edges_sha256"], "generator output changed"
    path = os.path.join(tmp, name.split("@")[0] + ".mtx")
    pgd_amd.write_mtx(path, n, s, d)
    return path


def flags_to_args(flags: list[str]) -> dict:
    kw = dict(neigh="deg", s2l_neigh=False, ordering="deg", s2l=False)
    i = 0
    while i < len(flags):
        f = flags[i]
        if f == "-o":
            kw["ordering"] = flags[i + 1]; i += 2
        elif f == "-n":
            kw["neigh"] = flags[i + 1]; i += 2
        elif f == "--s2l":
            kw["s2l"] = True; i += 1
        elif f == "--s2l_neigh":
            kw["s2l_neigh"] = True; i += 1
        elif f == "-r":
            i += 2  # representation does not change any output
        else:
            raise ValueError(f)
    return kw


def micro_matrix(res: dict, deg: np.ndarray) -> np.ndarray:
    ev, eu, order = res["ev"], res["eu"], res["order"]
    v, u = ev[order].astype(np.int64), eu[order].astype(np.int64)
    t = res["tri"].astype(np.uint64)
    s2 = (deg[v].astype(np.uint64) - t) + (deg[u].astype(np.uint64) - t) - np.uint64(2)
    cols = [(v + 1).astype(np.uint64), (u + 1).astype(np.uint64), t, s2, res["k4"], res["chordal"],
            np.zeros_like(t), res["c4"], res["star3"], res["path4"]]
    return np.stack([c.astype(np.uint64) for c in cols], axis=1) if len(t) else np.zeros((0, 10), np.uint64)


def micro_text(mat: np.ndarray) -> bytes:
    head = "% src,dst,triangle,2-star,4-clique,4-chordal-cycle,4-tailed-triangle,4-cycle,3-star,4-path\n"
    body = "".join(",".join(str(int(x)) for x in row) + "\n" for row in mat)
    return (head + body).encode()


@pytest.fixture(scope="module")
def tmpdir_mod():
    d = tempfile.mkdtemp(prefix="pgd_oracle_")
    yield d
    shutil.rmtree(d)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("rel", sorted(KNOWN))
def test_known_answers(rel, tmpdir_mod):
    gz = "inputs/" + os.path.basename(rel) + ".gz"
    path = input_path(gz, tmpdir_mod)
    g = OracleGraph.from_file(path)
    res = run_pipeline(g, micro=False)
    assert res["totals"] == KNOWN[rel]


@pytest.mark.parametrize("name", sorted(GOLDENS))
def test_oracle_matches_reference_goldens(name, tmpdir_mod):
    ent = GOLDENS[name]
    if ent.get("m", 0) > 300000:
        pytest.skip("large golden: covered by the GPU suite")
    path = golden_graph_path(name, ent, tmpdir_mod)
    g = OracleGraph.from_file(path)
    kw = flags_to_args(ent["flags"])
    res = run_pipeline(g, micro=True, **kw)
    assert res["totals"] == ent["totals"], [(TOTAL_NAMES[i], a, b) for i, (a, b) in
                                            enumerate(zip(res["totals"], ent["totals"])) if a != b]
    assert res["K"] == ent["K"]
    deg, _, _ = g.degrees()
    mat = micro_matrix(res, deg)
    if "micro_file" in ent:
        want = np.load(os.path.join(GOLD, ent["micro_file"]))["micro"]
        assert mat.shape == want.shape
        assert np.array_equal(mat, want)
    assert hashlib.sha256(micro_text(mat)).hexdigest() == ent["micro_sha256"]


def test_sample_graph_columns(tmpdir_mod):
    """The 8 per-edge columns stored in sample_graph.csv match the oracle."""
    path = input_path("inputs/sample_graph.csv.gz", tmpdir_mod)
    rows = []
    with open(path) as f:
        f.readline()
        for line in f:
            rows.append([int(x) for x in line.split(",")])
    # the reader renumbers ids in order of first appearance
    ids = {}
    for r in rows:
        if r[0] == r[1]:
            continue
        for x in r[:2]:
            if x not in ids:
                ids[x] = len(ids)
    g = OracleGraph.from_file(path)
    res = run_pipeline(g, micro=True)
    deg, _, _ = g.degrees()
    mat = micro_matrix(res, deg)
    got = {}
    for row in mat:
        a, b = int(row[0]) - 1, int(row[1]) - 1
        got[(min(a, b), max(a, b))] = [int(x) for x in row[2:]]
    checked = 0
    for r in rows:
        if r[0] == r[1]:
            continue
        a, b = ids[r[0]], ids[r[1]]
        assert got[(min(a, b), max(a, b))] == r[2:], r
        checked += 1
    assert checked == 14267


@pytest.mark.parametrize("seed", range(40))
def test_bruteforce_small(seed):
    rng = random.Random(seed)
    n = rng.randint(4, 11)
    edges = brute.random_graph(n, rng.uniform(0.15, 0.85), rng)
    if not edges:
        return
    adj = brute.adjacency(n, edges)
    want = brute.subset_totals(n, adj)
    # reference reader layout: adjacency lists in line order; n as given
    src = [a for a, _ in edges]
    dst = [b for _, b in edges]
    g = OracleGraph.from_edges(src, dst, n)
    res = run_pipeline(g, micro=True)
    assert dict(zip(TOTAL_NAMES, res["totals"])) == want
    assert res["sums"] == brute.lemma_sums(n, adj, edges)
    deg, _, _ = g.degrees()
    mat = micro_matrix(res, deg)
    for row in mat:
        v, u = int(row[0]) - 1, int(row[1]) - 1
        assert [int(x) for x in row[2:]] == brute.micro_columns(adj, v, u)
        assert g.brute_edge(v, u) == list(brute.edge_sets(adj, v, u))


@pytest.mark.parametrize("kind,k", [("clique", 4), ("clique", 7), ("clique", 12), ("star", 5), ("star", 9)])
def test_closed_forms(kind, k):
    if kind == "clique":
        n = k
        edges = [(a, b) for a in range(k) for b in range(a + 1, k)]
    else:
        n = k + 1
        edges = [(0, i) for i in range(1, k + 1)]
    g = OracleGraph.from_edges([a for a, _ in edges], [b for _, b in edges], n)
    res = run_pipeline(g, micro=False)
    assert dict(zip(TOTAL_NAMES, res["totals"])) == brute.closed_forms(kind, k)


@pytest.mark.parametrize("name", ["DD21.mtx", "gnp_2000_0.01", "ba_3000_4", "rmat_10"])
def test_reference_identities(name, tmpdir_mod):
    """The 10 identities of test_graphlet_counts (graphlet_core.cpp:1276-1285)."""
    if name not in GOLDENS:
        pytest.skip("golden not generated")
    path = golden_graph_path(name, GOLDENS[name], tmpdir_mod)
    g = OracleGraph.from_file(path)
    res = run_pipeline(g, micro=False)
    t = dict(zip(TOTAL_NAMES, res["totals"]))
    N = res["sums"][4:]  # N[1..10] at index 1..10
    M = (1 << 64) - 1
    ver = [
        3 * t["total_3_star"] + 3 * t["total_4_tailed_tris"] + 4 * t["total_4_cycle"] + t["total_4_path"]
        + 5 * t["total_4_chordcycle"] + 6 * t["total_4_clique"],
        3 * t["total_3_star"] + t["total_4_tailed_tris"] + 4 * t["total_4_cycle"] + t["total_4_path"],
        t["total_4_chordcycle"] + 6 * t["total_4_clique"],
        2 * t["total_4_tailed_tris"] + 4 * t["total_4_chordcycle"],
        t["total_4_path"] + 4 * t["total_4_cycle"],
        3 * t["total_3_star"] + t["total_4_tailed_tris"],
        2 * t["total_4_tailed_tris"] + 4 * t["total_4_cycle"] + 2 * t["total_4_path"] + 4 * t["total_4_chordcycle"]
        + 6 * t["total_4_clique"] + 2 * t["total_4_2edge"],
        2 * t["total_4_2edge"] + t["total_4_1edge"],
        2 * t["total_4_path"] + 2 * t["total_4_2star"],
        t["total_4_tailed_tris"] + 3 * t["total_4_tri"],
    ]
    for i, v in enumerate(ver, start=1):
        assert (v & M) == N[i], f"identity N{i}"
    n = res["n"]
    from math import comb
    four = sum(t[k] for k in TOTAL_NAMES[6:])
    assert four % (1 << 64) == comb(n, 4) % (1 << 64)


def test_derive_wrap_cases_against_reference_golden(tmpdir_mod):
    """n = 2^24: the n^3 (signed) and n^4 (unsigned) terms wrap and N8 overflows."""
    name = "bigid_2p24"
    if name not in GOLDENS:
        pytest.skip("golden not generated")
    ent = GOLDENS[name]
    path = golden_graph_path(name, ent, tmpdir_mod)
    g = OracleGraph.from_file(path)
    res = run_pipeline(g, micro=False)
    assert res["totals"] == ent["totals"]
    assert res["sums"][12] < res["m"] * (res["n"] ** 2) // 2  # N8 wrapped


# ---------------------------------------------------------------------------
# sort_edges over a sampled edge set E_s (graphlet_core.cpp:836-851): the
# oracle against goldens from the reference library (oracle/ref_subset.cpp)
SUBSET = json.load(open(os.path.join(GOLD, "goldens_subset.json"))) if os.path.exists(
    os.path.join(GOLD, "goldens_subset.json")) else {}


def oracle_subset_run(path: str, flags: list[str], es: np.ndarray):
    kw = flags_to_args(flags)
    g = OracleGraph.from_file(path)
    g.order_neighbors(kw["neigh"], kw["s2l_neigh"])
    ev, eu = g.edge_arrays()
    order = g.sort_edges_subset(ev, eu, es, kw["ordering"], kw["s2l"])
    res = g.decompose(ev, eu, order, micro=True)
    res.update(ev=ev, eu=eu, order=order)
    res["totals"] = derive(g.n, g.m, res["sums"])
    return res


@pytest.mark.parametrize("key", sorted(SUBSET))
def test_oracle_subset_matches_reference(key, tmp_path):
    ent = SUBSET[key]
    path = golden_graph_path(ent["graph"], GOLDENS[ent["graph"]], str(tmp_path))
    data = np.load(os.path.join(GOLD, ent["file"]))
    res = oracle_subset_run(path, ent["flags"], data["E_s"])
    assert res["totals"] == ent["totals"]
    want = data["micro"]
    assert len(res["order"]) == ent["count"] == want.shape[0]
    ev, eu, order = res["ev"], res["eu"], res["order"]
    assert np.array_equal(ev[order] + 1, want[:, 0]) and np.array_equal(eu[order] + 1, want[:, 1])
    assert np.array_equal(res["tri"], want[:, 2])
    assert np.array_equal(res["k4"], want[:, 4])
    assert np.array_equal(res["chordal"], want[:, 5])
    assert np.array_equal(res["c4"], want[:, 7])
    assert np.array_equal(res["star3"], want[:, 8])
    assert np.array_equal(res["path4"], want[:, 9])
