"""Parity of the HIP path against the reference goldens and the oracle (MI355X).

Every case runs the full device pipeline: CSR construction from the reader's
COO, neighbour ordering, orientation, k-core, E_ordered, counting, the
reduction and (host, integer) derivation, then compares bit for bit.
"""
from __future__ import annotations

import hashlib
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

import pgd_amd
from tests.oracle_lib import OracleGraph, run_pipeline
from tests.test_oracle import GOLDENS, GOLD, flags_to_args, golden_graph_path

pytestmark = pytest.mark.gpu

# the two engines the library has: degree-oriented (sparse) and dense int8 MFMA
ENGINES = ["orient", "dense"]
DENSE_MAX_N = 8192


@pytest.fixture(scope="module")
def tmp():
    d = tempfile.mkdtemp(prefix="pgd_gpu_")
    yield d
    shutil.rmtree(d)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def product_run(path, flags, ctx, engine, micro=True):
    kw = flags_to_args(flags)
    g = pgd_amd.GraphletCore.from_file(path, ctx=ctx, neigh_ordering=kw["neigh"],
                                       neigh_small_to_large=kw["s2l_neigh"])
    K = g.compute_cores()
    g.sort_edges(kw["ordering"], kw["s2l"])
    totals = g.graphlet_decomposition_micro(engine) if micro else g.graphlet_decomposition(engine)
    return g, [totals[k] for k in pgd_amd.TOTAL_NAMES], K


def micro_matrix(g) -> np.ndarray:
    ev, eu = g.edge_arrays()
    order = g.e_ordered()
    cols = [ev[order].astype(np.uint64) + np.uint64(1), eu[order].astype(np.uint64) + np.uint64(1)]
    k = len(order)  # one row per E_ordered position
    cols += [g.micro(c)[:k] for c in ("tri", "2_star", "local_4_clique", "local_4_chordal_cycle",
                                      "local_4_tailed_tris", "local_4_cycle", "local_3_star", "local_4_path")]
    return np.stack(cols, axis=1) if k else np.zeros((0, 10), np.uint64)


def micro_sha(mat: np.ndarray) -> str:
    head = "% src,dst,triangle,2-star,4-clique,4-chordal-cycle,4-tailed-triangle,4-cycle,3-star,4-path\n"
    body = "".join(",".join(map(str, row)) + "\n" for row in mat.tolist())
    return hashlib.sha256((head + body).encode()).hexdigest()


MULTIGRAPHS = {"rd_mtx_dups.mtx"}
CASES = sorted(k for k, e in GOLDENS.items() if e.get("m", 0) <= 400000)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("name", CASES)
def test_matches_reference_golden(name, engine, ctx, tmp):
    ent = GOLDENS[name]
    say(f"[gpu] golden {name} engine={engine} m={ent.get('m')}")
    path = golden_graph_path(name, ent, tmp)
    if name in MULTIGRAPHS:
        # the reference keeps repeated .mtx pairs as multi-edges; this build
        # accepts simple graphs only and says so
        with pytest.raises(pgd_amd.PGDError) as e:
            product_run(path, ent["flags"], ctx, engine)
        assert e.value.code == -5
        return
    try:
        g, totals, K = product_run(path, ent["flags"], ctx, engine)
    except pgd_amd.PGDError as e:
        if engine == "dense" and e.code == -5:
            pytest.skip(f"more than {DENSE_MAX_N} vertices for the dense engine")
        raise
    assert g.last_engine() == engine
    assert totals == ent["totals"]
    assert K == ent["K"]
    mat = micro_matrix(g)
    if "micro_file" in ent:
        want = np.load(os.path.join(GOLD, ent["micro_file"]))["micro"]
        assert np.array_equal(mat, want)
    assert micro_sha(mat) == ent["micro_sha256"]
    g.free()


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("scale,seed", [(15, 5), (16, 6)])
def test_rmat_per_edge_against_oracle(scale, seed, engine, ctx):
    """Per-edge T/K4/C4 of every edge id against the oracle (R-MAT, n up to 65536,
    which takes the HBM-bitmap path of the edge engine)."""
    say(f"[gpu] rmat scale={scale} seed={seed} engine={engine}")
    s, d = pgd_amd.gen_rmat(scale, 16, seed=seed)
    n = int(max(s.max(), d.max())) + 1
    if engine == "dense" and n > DENSE_MAX_N:
        pytest.skip(f"more than {DENSE_MAX_N} vertices for the dense engine")
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
    g.sort_edges("deg")
    sums = g.decompose(micro=False, engine=engine)
    T, K4, C4 = g.edge_counts()
    og = OracleGraph.from_edges(s, d, n)
    og.order_neighbors("deg")
    ev, eu = og.edge_arrays()
    ids = np.arange(len(ev), dtype=np.int32)
    rng = np.random.default_rng(seed)
    sample = np.sort(rng.choice(ids, size=min(len(ids), 4000), replace=False)).astype(np.int32)
    res = og.decompose(ev, eu, sample, micro=True)
    assert np.array_equal(T[sample], res["tri"])
    assert np.array_equal(K4[sample], res["k4"])
    assert np.array_equal(C4[sample], res["c4"])
    assert sums[0] == int(T.sum()) and sums[3] == int(K4.sum()) and sums[4] == int(C4.sum())
    g.free()


@pytest.mark.parametrize("engine", ENGINES)
def test_deterministic_repeat(engine, ctx):
    s, d = pgd_amd.gen_ba(4000, 6, 21)
    g = pgd_amd.GraphletCore(s, d, ctx=ctx)
    a = g.graphlet_decomposition_micro(engine)
    ma = micro_matrix(g)
    b = g.graphlet_decomposition_micro(engine)
    mb = micro_matrix(g)
    assert a == b
    assert np.array_equal(ma, mb)
    g.free()


TUNINGS = [
    dict(wedge_log2r=8),                                     # many endpoint tiles, items cut at 128 tiles
    dict(wedge_log2r=10, wedge_nt=256, wedge_split=64),      # small items
    dict(wedge_log2r=14, wedge_nt=1024, small_k=32, mid_k=128),
    dict(wedge_log2r=16, wedge_nt=1024),                     # 64K-rank tiles: one block per CU
    dict(wedge_log2r=12, wedge_nt=512, grid_small=4, grid_mid=1, wedge_split=1),
    dict(tri_lists=0),                                       # clique pass searches again
    dict(tri_lists=1),                                       # lists kept until the buffer is full
    dict(tri_search=0),                                      # every out-list probed
    dict(tri_search=1),                                      # table lookups whenever fewer than the list
]


@pytest.mark.parametrize("ti", range(len(TUNINGS)))
def test_tuning_never_changes_counts(ti, ctx):
    """Launch parameters change the schedule, never a count: per-edge T/K4/C4
    under each tuning equal the default's (and the oracle's on a sample)."""
    tun = TUNINGS[ti]
    say(f"[gpu] tuning {tun}")
    s, d = pgd_amd.gen_rmat(14, 16, seed=5)
    n = int(max(s.max(), d.max())) + 1
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
    g.sort_edges("deg")
    base_sums = g.decompose(engine="orient")
    base = g.edge_counts()
    g.set_tuning(**tun)
    assert g.tuning().as_dict() == {**pgd_amd.default_tuning().as_dict(), **tun}
    sums = g.decompose(engine="orient")
    got = g.edge_counts()
    assert sums == base_sums
    for a, b in zip(base, got):
        assert np.array_equal(a, b)
    og = OracleGraph.from_edges(s, d, n)
    og.order_neighbors("deg")
    ev, eu = og.edge_arrays()
    sample = np.sort(np.random.default_rng(ti).choice(len(ev), size=1500, replace=False)).astype(np.int32)
    res = og.decompose(ev, eu, sample, micro=True)
    assert np.array_equal(got[2][sample], res["c4"])
    g.free()


def test_plan_cells_outside_lds(ctx):
    """256-rank endpoint tiles on R-MAT-20 (4096 tiles): an anchor whose tile
    range passes 2048 cells has its cells counted and its segments placed
    with global cursors instead of LDS ones; the counts equal those of the
    default 2^15-rank tiles, edge for edge."""
    say("[gpu] R-MAT-20 with 256-rank tiles")
    s, d = pgd_amd.gen_rmat(20, 16, seed=3)
    n = int(max(s.max(), d.max())) + 1
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
    g.sort_edges("deg")
    base_sums = g.decompose(engine="orient")
    base = g.edge_counts()
    g.set_tuning(wedge_log2r=8)
    assert g.decompose(engine="orient") == base_sums
    for a, b in zip(base, g.edge_counts()):
        assert np.array_equal(a, b)
    g.free()


def test_wedge_counter_overflow(ctx):
    """A hub with 70,000 in-neighbours that all reach one other vertex: the
    16-bit endpoint counter of that vertex overflows and its tile is
    recounted with 32-bit counters.  Closed forms: every wedge a-b-c closes
    69,999 induced 4-cycles a-b-c-b'."""
    k = 70000
    a, c, x = 0, 1, 2            # x: an extra neighbour so that a is the top rank
    bs = np.arange(3, 3 + k, dtype=np.int32)
    src = np.concatenate([np.full(k, a, np.int32), np.full(k, c, np.int32), np.array([a], np.int32)])
    dst = np.concatenate([bs, bs, np.array([x], np.int32)])
    n = 3 + k
    g = pgd_amd.GraphletCore(src, dst, n=n, ctx=ctx)
    g.sort_edges("deg")
    for log2r in (8, 14, 16):
        g.set_tuning(wedge_log2r=log2r)
        g.decompose(engine="orient")
        T, K4, C4 = g.edge_counts()
        ev, eu = g.edge_arrays()
        assert int(T.sum()) == 0 and int(K4.sum()) == 0
        want = np.where((ev == x) | (eu == x), 0, k - 1).astype(np.uint64)
        assert np.array_equal(C4, want), log2r
    g.free()


def test_wide_list_slot_sums(ctx):
    """Per-slot 4-cycle sums of a wide list (d(b) * max degree >= 2^32 takes
    64-bit sums).  Hubs h1 and g (degree K+1) share K common neighbours C;
    h2 (degree K+3: h1, g and K+1 leaves X) is the top rank, so its wedges
    h2-h1-c and h2-g-c land on the slots of the wide lists N(h1), N(g); g's
    K = 65536 in-neighbours also overflow its 16-bit counters.  Closed forms:
    T = K4 = 0 on every edge; C4 = K on h2-h1, h2-g, h1-c, g-c and 0 on h2-x."""
    K = 65536
    h2, h1, g = 0, 1, 2
    C = np.arange(3, 3 + K, dtype=np.int32)
    X = np.arange(3 + K, 4 + 2 * K, dtype=np.int32)
    src = np.concatenate([[h2, h2], np.full(K + 1, h2), np.full(K, h1), np.full(K, g)]).astype(np.int32)
    dst = np.concatenate([[h1, g], X, C, C]).astype(np.int32)
    n = 4 + 2 * K
    say(f"[gpu] wide lists K={K}")
    gr = pgd_amd.GraphletCore(src, dst, n=n, ctx=ctx)
    gr.sort_edges("deg")
    gr.decompose(engine="orient")
    T, K4, C4 = gr.edge_counts()
    ev, eu = gr.edge_arrays()
    assert int(T.sum()) == 0 and int(K4.sum()) == 0
    leaf = (ev >= 3 + K) | (eu >= 3 + K)
    want = np.where(leaf, 0, K).astype(np.uint64)
    assert np.array_equal(C4, want)
    gr.free()


def test_high_degree_local_edge_sums(ctx):
    """Clique-pass sums of an edge whose ends both have degree above the
    packing limit (separate 64-bit K4 and Q sums instead of the packed one;
    for maximum degree 70001 the limit is 61356).  Book graph: hubs x, y
    adjacent, N pages each adjacent to both, pages matched in pairs.  Closed
    forms: (x, y) has T = N, K4 = N/2; a hub-page or page-page edge has T = 2,
    K4 = 1; C4 = 0 on every edge (every 4-cycle has a chord)."""
    N = 70000
    x, y = 0, 1
    w = np.arange(2, 2 + N, dtype=np.int32)
    src = np.concatenate([[x], np.full(N, x), np.full(N, y), w[0::2]]).astype(np.int32)
    dst = np.concatenate([[y], w, w, w[1::2]]).astype(np.int32)
    say(f"[gpu] book graph N={N}")
    g = pgd_amd.GraphletCore(src, dst, n=N + 2, ctx=ctx)
    g.sort_edges("deg")
    g.decompose(engine="orient")
    T, K4, C4 = g.edge_counts()
    ev, eu = g.edge_arrays()
    hub = ((ev == x) & (eu == y)) | ((ev == y) & (eu == x))
    assert hub.sum() == 1
    assert np.array_equal(T, np.where(hub, N, 2).astype(np.uint32))
    assert np.array_equal(K4, np.where(hub, N // 2, 1).astype(np.uint64))
    assert not C4.any()
    g.free()


@pytest.mark.parametrize("n,lists", [(1000, 64), (1000, 0), (2100, 64)])
def test_local_graphs_beyond_lds(n, lists, ctx):
    """K_n: the lowest-ranked vertices have out-degree up to n-1.  At n = 1000
    the clique pass holds their local bit matrices in LDS column blocks (the
    whole matrix exceeds the LDS); at n = 2100 the largest local graphs exceed
    the LDS bin and run from HBM scratch.  With lists=0 the clique pass
    searches the local graphs again instead of reading the triangle pass's
    lists.  Closed forms of K_n per edge: T = n-2, K4 = C(n-2, 2), C4 = 0."""
    say(f"[gpu] K_{n} lists={lists}")
    i, j = np.triu_indices(n, 1)
    g = pgd_amd.GraphletCore(i.astype(np.int32), j.astype(np.int32), n=n, ctx=ctx)
    g.sort_edges("deg")
    g.set_tuning(tri_lists=lists)
    g.decompose(engine="orient")
    T, K4, C4 = g.edge_counts()
    assert np.all(T == n - 2)
    assert np.all(K4 == (n - 2) * (n - 3) // 2)
    assert np.all(C4 == 0)
    g.free()


def test_tuning_rejects_bad_values(ctx):
    s, d = pgd_amd.gen_gnp(100, 0.1, 1)
    g = pgd_amd.GraphletCore(s, d, ctx=ctx)
    for bad in (dict(wedge_log2r=17), dict(wedge_nt=128), dict(small_k=65), dict(mid_k=512), dict(wedge_split=0),
                dict(tri_search=-1),
                dict(tri_lists=-1)):
        with pytest.raises(pgd_amd.PGDError):
            g.set_tuning(**bad)
    g.free()


def test_engines_agree_per_edge(ctx):
    if len(ENGINES) < 2:
        pytest.skip("one engine")
    s, d = pgd_amd.gen_rmat(13, 16, seed=9)
    g = pgd_amd.GraphletCore(s, d, ctx=ctx)
    g.sort_edges("deg")
    out = {}
    for e in ENGINES:
        g.decompose(micro=False, engine=e)
        out[e] = g.edge_counts()
    for e in ENGINES[1:]:
        for a, b in zip(out[ENGINES[0]], out[e]):
            assert np.array_equal(a, b)
    g.free()


def test_invalid_ids_rejected(ctx):
    s = np.array([0, 1, 5], dtype=np.int32)
    d = np.array([1, 2, 1], dtype=np.int32)
    with pytest.raises(pgd_amd.PGDError) as e:
        pgd_amd.GraphletCore(s, d, n=3, ctx=ctx)
    assert e.value.code == -1


def test_memory_cap_enforced():
    with pgd_amd.Context(0) as c:
        c.set_mem_cap(1 << 16)
        s, d = pgd_amd.gen_gnp(3000, 0.01, 2)
        with pytest.raises(pgd_amd.PGDError) as e:
            pgd_amd.GraphletCore(s, d, ctx=c)
        assert e.value.code == -4


def test_empty_and_tiny_graphs(ctx):
    for s, d, n in [([], [], 0), ([0], [0], 1), ([0], [1], 2), ([0, 1], [1, 2], 5)]:
        g = pgd_amd.GraphletCore(np.array(s, np.int32), np.array(d, np.int32), n=n, ctx=ctx)
        tot = g.graphlet_decomposition_micro()
        og = OracleGraph.from_edges(s, d, n)
        want = run_pipeline(og, micro=False)["totals"]
        assert [tot[k] for k in pgd_amd.TOTAL_NAMES] == want
        g.free()


def test_adopt_reference_csr(ctx, tmp):
    """pgd_graph_from_csr keeps the neighbour order of an already ordered CSR
    (the state graphlet_core holds), so edge ids, E_ordered and every micro
    row equal the reference's."""
    name = "DD21.mtx"
    ent = GOLDENS[name]
    path = golden_graph_path(name, ent, tmp)
    og = OracleGraph.from_file(path)
    og.order_neighbors("deg")
    row, col = og.csr()
    g = pgd_amd.GraphletCore.from_csr(row, col, ctx=ctx)
    g.preprocess("natural")
    g.sort_edges("deg")
    totals = g.graphlet_decomposition_micro()
    assert [totals[k] for k in pgd_amd.TOTAL_NAMES] == ent["totals"]
    assert micro_sha(micro_matrix(g)) == ent["micro_sha256"]
    g.free()


def test_atomic_rate_probe(ctx):
    """pgd_atomic_rate: every counter checked to equal the launch count, and
    a positive rate; out-of-range arguments are refused before any launch."""
    r, ms = ctx.atomic_rate(1 << 22, 3)
    assert r > 0 and ms > 0
    for words, reps in ((1 << 10, 1), (1 << 22, 0), (1 << 22, 101), (1 << 31, 1)):
        with pytest.raises(pgd_amd.PGDError) as e:
            ctx.atomic_rate(words, reps)
        assert e.value.code == -1
