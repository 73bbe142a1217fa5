"""The caller's edge order through the C ABI (MI355X).

* pgd_sort_edges_subset: sort_edges over a sampled edge set E_s
  (graphlet_core.cpp:836-851) and the decomposition over its positions only
  (1380, 1513), against goldens from the reference library
  (tests/golden/make_subset_goldens.py).
* pgd_set_e_ordered: a reference object's own CSR and E_ordered adopted as
  they are, against the reference goldens and the oracle.
* the argument checks of both, and of pgd_graph_from_csr.
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

import pgd_amd
from tests.oracle_lib import OracleGraph, derive
from tests.test_gpu import micro_matrix
from tests.test_oracle import GOLD, GOLDENS, SUBSET, flags_to_args, golden_graph_path

pytestmark = pytest.mark.gpu

ERR_INVALID, ERR_UNSUPPORTED = -1, -5


@pytest.fixture(scope="module")
def tmp():
    d = tempfile.mkdtemp(prefix="pgd_sub_")
    yield d
    shutil.rmtree(d)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("key", sorted(SUBSET))
def test_sampled_edge_set_matches_reference(key, ctx, tmp):
    ent = SUBSET[key]
    say(f"[gpu] E_s {key} count={ent['count']}")
    path = golden_graph_path(ent["graph"], GOLDENS[ent["graph"]], tmp)
    data = np.load(os.path.join(GOLD, ent["file"]))
    kw = flags_to_args(ent["flags"])
    g = pgd_amd.GraphletCore.from_file(path, ctx=ctx, neigh_ordering=kw["neigh"],
                                       neigh_small_to_large=kw["s2l_neigh"])
    g.sort_edges(kw["ordering"], kw["s2l"], E_s=data["E_s"])
    totals = g.graphlet_decomposition_micro()
    assert [totals[k] for k in pgd_amd.TOTAL_NAMES] == ent["totals"]
    mat = micro_matrix(g)
    assert mat.shape[0] == ent["count"]
    assert np.array_equal(mat, data["micro"])
    # positions past |E_s| keep the reference's zeros
    for c in ("tri", "local_4_clique", "local_4_cycle", "local_3_star", "local_4_path", "2_star"):
        assert not g.micro(c)[ent["count"]:].any()
    g.free()


def _oracle_object(path: str, neigh: str, s2l_neigh: bool):
    """A reference object's state after its reader and preprocess, held by
    the oracle: CSR (vertices/edges) and e_v/e_u."""
    og = OracleGraph.from_file(path)
    og.order_neighbors(neigh, s2l_neigh)
    row, col = og.csr()
    ev, eu = og.edge_arrays()
    return og, row, col, ev, eu


@pytest.mark.parametrize("name,flags,golden", [
    ("DD21.mtx", ["-o", "kcore"], "DD21.mtx@o_kcore"),
    ("sample_graph.csv", ["-o", "deg_vol"], "sample_graph.csv@o_deg_vol"),
    ("gnp_2000_0.01", ["--s2l_neigh"], "gnp_2000_0.01@n_s2l"),
    ("DD21.mtx", ["-o", "kcore", "--s2l", "-n", "kcore"], None)])
def test_adopted_csr_and_order_match_reference(name, flags, golden, ctx, tmp):
    """The drop-in binding's route: pgd_graph_from_csr on the object's
    vertices/edges, "natural" preprocessing (keeps its edge ids), then
    pgd_set_e_ordered with the object's own E_ordered."""
    kw = flags_to_args(flags)
    say(f"[gpu] adopt {name} {flags}")
    path = golden_graph_path(name, GOLDENS[name], tmp)
    og, row, col, ev, eu = _oracle_object(path, kw["neigh"], kw["s2l_neigh"])
    order = og.sort_edges(ev, eu, kw["ordering"], kw["s2l"])
    g = pgd_amd.GraphletCore.from_csr(row, col, ctx=ctx)
    g.preprocess("natural")
    gev, geu = g.edge_arrays()
    assert np.array_equal(gev, ev) and np.array_equal(geu, eu)
    g.set_e_ordered(order)
    assert np.array_equal(g.e_ordered(), order)
    totals = g.graphlet_decomposition_micro()
    got = [totals[k] for k in pgd_amd.TOTAL_NAMES]
    res = og.decompose(ev, eu, order, micro=True)
    assert got == derive(og.n, og.m, res["sums"])
    if golden:
        ent = GOLDENS[golden]
        assert ent["flags"] == flags and got == ent["totals"]
        if "micro_file" in ent:
            assert np.array_equal(micro_matrix(g), np.load(os.path.join(GOLD, ent["micro_file"]))["micro"])
    assert np.array_equal(g.micro("tri"), res["tri"])
    assert np.array_equal(g.micro("local_4_clique"), res["k4"])
    assert np.array_equal(g.micro("local_4_cycle"), res["c4"])
    assert np.array_equal(g.micro("local_4_chordal_cycle"), res["chordal"])
    assert np.array_equal(g.micro("local_3_star"), res["star3"])
    assert np.array_equal(g.micro("local_4_path"), res["path4"])
    g.free()


def test_adopted_sampled_order_and_duplicates(ctx, tmp):
    """A 10% E_s adopted as E_ordered, and an E_ordered holding every edge
    twice (what a second sort_edges call leaves in the reference): sums over
    every position, against the oracle."""
    say("[gpu] adopt sampled / doubled E_ordered")
    path = golden_graph_path("DD21.mtx", GOLDENS["DD21.mtx"], tmp)
    og, row, col, ev, eu = _oracle_object(path, "deg", False)
    rng = np.random.default_rng(3)
    es = rng.choice(og.m, size=og.m // 10, replace=False).astype(np.int32)
    g = pgd_amd.GraphletCore.from_csr(row, col, ctx=ctx)
    g.preprocess("natural")
    g.set_e_ordered(es)
    sums = g.decompose(micro=True)
    res = og.decompose(ev, eu, es, micro=True)
    assert sums == res["sums"]
    assert np.array_equal(g.micro("local_4_cycle")[: len(es)], res["c4"])
    assert not g.micro("local_4_cycle")[len(es):].any()
    full = og.sort_edges(ev, eu, "deg", False)
    twice = np.concatenate([full, full])
    g.set_e_ordered(twice)
    sums2 = g.decompose(micro=False)
    assert sums2 == og.decompose(ev, eu, twice, micro=False)["sums"]
    # per-edge columns hold m positions: a longer E_ordered cannot give them
    with pytest.raises(pgd_amd.PGDError) as e:
        g.decompose(micro=True)
    assert e.value.code == ERR_INVALID
    g.free()


def test_edge_order_argument_checks(ctx):
    say("[gpu] edge order argument checks")
    s, d = pgd_amd.gen_gnp(200, 0.05, 4)
    g = pgd_amd.GraphletCore(s, d, n=200, ctx=ctx)
    for bad in ([0, g.m], [-1], [5, 2**31 - 1]):
        with pytest.raises(pgd_amd.PGDError) as e:
            g.set_e_ordered(bad)
        assert e.value.code == ERR_INVALID
        with pytest.raises(pgd_amd.PGDError) as e:
            g.sort_edges("deg", False, E_s=bad)
        assert e.value.code == ERR_INVALID
    # an empty E_s: no positions, zero sums, every micro column 0
    g.sort_edges("deg", False, E_s=[])
    assert len(g.e_ordered()) == 0
    assert g.decompose(micro=True) == [0] * 15
    assert not g.micro("tri").any()
    g.free()


def test_unknown_decompose_flag_is_rejected(ctx):
    """Only PGD_MICRO, PGD_ENGINE_ORIENT and PGD_ENGINE_DENSE are flags."""
    import ctypes as C
    from pgd_amd._lib import lib, ptr
    s, d = pgd_amd.gen_gnp(100, 0.1, 2)
    g = pgd_amd.GraphletCore(s, d, n=100, ctx=ctx)
    g.sort_edges("deg")
    out = np.zeros(15, dtype=np.uint64)
    for flags in (2, 16, 1 << 31):
        assert lib().pgd_decompose(g.h, flags, ptr(out, C.c_uint64)) == ERR_INVALID
    g.free()


def _csr(lists):
    row = np.zeros(len(lists) + 1, dtype=np.int64)
    row[1:] = np.cumsum([len(x) for x in lists])
    col = np.array([y for x in lists for y in x], dtype=np.int32)
    return row, col


@pytest.mark.parametrize("case,code", [
    ("decreasing", ERR_INVALID), ("asymmetric", ERR_INVALID), ("asymmetric_same_count", ERR_INVALID),
    ("repeated", ERR_UNSUPPORTED), ("self_loop", ERR_INVALID), ("out_of_range", ERR_INVALID)])
def test_from_csr_rejects_malformed(case, code, ctx):
    """Checked before the CSR is used, so nothing is written past its arrays."""
    say(f"[gpu] from_csr {case}")
    if case == "decreasing":
        row = np.array([0, 4, 2, 4, 6], dtype=np.int64)   # row[1] > row[2]
        col = np.array([1, 2, 0, 3, 0, 1], dtype=np.int32)
    elif case == "asymmetric":
        row, col = _csr([[1, 2], [0], [3], [2, 1]])       # 6 slots; (2,0) and (1,3) missing
    elif case == "asymmetric_same_count":
        # lists 0:[1] 1:[2] 2:[0,1]: an even slot count, but (0,2) and (1,0) are missing
        row, col = np.array([0, 1, 2, 4], np.int64), np.array([1, 2, 0, 1], np.int32)
    elif case == "repeated":
        row, col = _csr([[1, 1], [0, 0]])                 # the .mtx multi-edge
    elif case == "self_loop":
        row, col = _csr([[0, 1], [0]])
    else:
        row, col = np.array([0, 1, 2, 2], np.int64), np.array([1, 5], np.int32)  # id 5 with n = 3
    with pytest.raises(pgd_amd.PGDError) as e:
        pgd_amd.GraphletCore.from_csr(row, col, ctx=ctx)
    assert e.value.code == code


def test_from_csr_accepts_valid_and_keeps_order(ctx):
    say("[gpu] from_csr valid")
    row, col = _csr([[2, 1], [0, 2], [1, 0, 3], [2]])
    g = pgd_amd.GraphletCore.from_csr(row, col, ctx=ctx)
    r2, c2 = g.csr()
    assert np.array_equal(r2, row) and np.array_equal(c2, col)
    g.preprocess("natural")
    g.sort_edges("deg")
    assert g.graphlet_decomposition()["total_3_tris"] == 1
    g.free()
