"""Parity at the full sizes of the benchmark workloads (MI355X).

* R-MAT scale 22, edge factor 16 (64.2M edges): the device's neighbour
  order, edge arrays, k-core and E_ordered equal the oracle's on the whole
  graph; per-edge T, K4, C4 and the micro columns equal the oracle's on the
  heaviest E_ordered positions (the hub-hub edges) plus a stride sample; the
  15 sums equal the sums of the per-edge counts; the sums give integral
  totals that satisfy the reference's identities (graphlet_core.cpp:1276-1285).
* The batched workload (8192 graphs): every graph equals the plain path,
  every 16th the oracle.
* The served workload: one second of Poisson arrivals at 2000 req/s with 1%
  R-MAT scale-18 graphs: every small request equals the oracle, every large
  one the plain path.
"""
from __future__ import annotations

import sys
import time

import numpy as np
import pytest

import pgd_amd
from tests.oracle_lib import OracleGraph, derive, run_pipeline

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def reader_csr(s: np.ndarray, d: np.ndarray, n: int):
    """CSR in the reference reader's layout: each vertex's list in line order
    (line i appends d to s's list and s to d's list)."""
    L = len(s)
    own = np.empty(2 * L, np.int32)
    nb = np.empty(2 * L, np.int32)
    own[0::2], own[1::2] = s, d
    nb[0::2], nb[1::2] = d, s
    p = np.argsort(own, kind="stable")
    col = nb[p]
    row = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(own, minlength=n), out=row[1:])
    return row, col


def check_sums_and_identities(totals: list[int], sums: list[int]) -> None:
    """The exact graphlet totals from the 15 sums, in integers: every division
    of the derivation (graphlet_core.cpp:1449-1472) must be exact, the
    reference's identities for N1 and N2 (graphlet_core.cpp:1276-1277; those
    for N3..N10 restate the derivation itself) must hold mod 2^64 on them, and
    the derived totals must equal them wherever the reference's binary64
    steps are exact (below 2^53).  Above 2^53 the reference's 4_2star and
    4_tri are rounded, and its own identity check would report the
    difference; their rounding is checked by the derivation tests."""
    sT, sS, sD, sK, sC = sums[:5]
    N = [None] + sums[5:15]
    for v, q, what in ((sT, 3, "sum T"), (sS, 2, "sum S"), (sK, 6, "sum K4"), (sC, 4, "sum C4")):
        assert v % q == 0, what
    tri, s2, k4, c4 = sT // 3, sS // 2, sK // 6, sC // 4
    cc = N[3] - 6 * k4
    p4 = N[5] - 4 * c4
    assert (N[4] - 4 * cc) % 2 == 0
    tt = (N[4] - 4 * cc) // 2
    assert (N[6] - tt) % 3 == 0
    st3 = (N[6] - tt) // 3
    e2 = N[7] - 6 * k4 - 4 * c4 - 4 * cc - 2 * p4 - 2 * tt
    assert e2 % 2 == 0
    e42 = e2 // 2
    assert (N[9] - 2 * p4) % 2 == 0
    st24 = (N[9] - 2 * p4) // 2
    assert (N[10] - tt) % 3 == 0
    tri4 = (N[10] - tt) // 3
    for v in (cc, p4, tt, st3, e42, st24, tri4):
        assert 0 <= v < 1 << 64
    assert (3 * st3 + 3 * tt + 4 * c4 + p4 + 5 * cc + 6 * k4) & M64 == N[1], "identity N1"
    assert (3 * st3 + tt + 4 * c4 + p4) & M64 == N[2], "identity N2"
    exact = {2: tri, 3: s2, 6: k4, 7: cc, 8: tt, 9: c4, 10: st3, 11: p4, 13: e42, 14: st24, 15: tri4}
    for idx, v in exact.items():
        if v < 1 << 53:
            assert totals[idx] == v, idx


@pytest.fixture(scope="module")
def rmat22():
    say("[gpu] generating R-MAT scale 22, ef 16, seed 1")
    s, d = pgd_amd.gen_rmat(22, 16, 0.57, 0.19, 0.19, 1)
    n = int(max(s.max(), d.max())) + 1
    return n, s, d


def test_rmat22_full_size(rmat22, ctx):
    n, s, d = rmat22
    t0 = time.time()
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
    K = g.compute_cores()
    g.sort_edges("deg")
    sums = g.decompose(micro=True)
    totals = derive(n, g.m, sums)
    assert totals == pgd_amd.derive(n, g.m, sums)
    say(f"[gpu] R-MAT-22 counted: m={g.m} max degree {g.max_degree} ({time.time() - t0:.1f}s)")
    check_sums_and_identities(totals, sums)
    T, K4, C4 = g.edge_counts()
    assert int(T.sum()) == sums[0]
    assert int(K4.sum(dtype=np.uint64)) == sums[3]
    assert int(C4.sum(dtype=np.uint64)) == sums[4]
    # preprocessing on the whole graph against the oracle
    row, col = g.csr()
    ev, eu = g.edge_arrays()
    order = g.e_ordered()
    orow, ocol = reader_csr(s, d, n)
    og = OracleGraph.from_csr(orow, ocol)
    del orow, ocol
    og.order_neighbors("deg", False)
    r2, c2 = og.csr()
    assert np.array_equal(r2, row) and np.array_equal(c2, col), "neighbour order"
    del r2, c2
    oev, oeu = og.edge_arrays()
    assert np.array_equal(oev, ev) and np.array_equal(oeu, eu), "edge arrays"
    oK, _ = og.kcore()
    assert oK == K, "k-core"
    assert np.array_equal(og.sort_edges(ev, eu, "deg", False), order), "E_ordered"
    say(f"[gpu] R-MAT-22 preprocessing equals the oracle ({time.time() - t0:.1f}s)")
    # per-edge counts: the 64 heaviest positions and a stride sample
    stride = len(order) // 2000
    pos = np.unique(np.concatenate([np.arange(64), np.arange(64, len(order), stride)]))
    res = og.decompose(ev, eu, np.ascontiguousarray(order[pos]), micro=True)
    ids = order[pos]
    assert np.array_equal(T[ids], res["tri"])
    assert np.array_equal(K4[ids], res["k4"])
    assert np.array_equal(C4[ids], res["c4"])
    for col_name, key in (("tri", "tri"), ("local_4_clique", "k4"), ("local_4_cycle", "c4"),
                          ("local_4_chordal_cycle", "chordal"), ("local_4_path", "path4"),
                          ("local_3_star", "star3")):
        assert np.array_equal(g.micro(col_name)[pos], res[key]), col_name
    say(f"[gpu] R-MAT-22 {len(pos)} per-edge positions equal the oracle ({time.time() - t0:.1f}s)")
    g.free()


def test_batched_full_workload(ctx):
    from pgd_amd.workloads import batched_graphs
    graphs = batched_graphs(8192, seed=1)
    t0 = time.time()
    b = pgd_amd.Batch(graphs, ctx=ctx, micro=True)
    sums, totals = b.run()
    say(f"[gpu] batch of {len(graphs)} graphs, {b.m_total} edges ({time.time() - t0:.1f}s)")
    cols = ("tri", "local_4_clique", "local_4_cycle", "local_4_path", "local_3_star", "local_4_chordal_cycle")
    for i, (n, s, d) in enumerate(graphs):
        g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
        g.sort_edges("deg")
        want = g.decompose(micro=True)
        assert [int(x) for x in sums[i]] == want, i
        assert [int(x) for x in totals[i]] == derive(n, g.m, want), i
        if i % 16 == 0:
            for c in cols:
                assert np.array_equal(b.micro(i, c, g.m), g.micro(c)), (i, c)
            ref = run_pipeline(OracleGraph.from_edges(s, d, n), micro=True)
            assert want == ref["sums"], i
            assert np.array_equal(b.micro(i, "local_4_cycle", g.m), ref["c4"]), i
            assert np.array_equal(b.micro(i, "local_4_clique", g.m), ref["k4"]), i
        g.free()
        if i % 1024 == 1023:
            say(f"[gpu] batch: {i + 1} graphs checked ({time.time() - t0:.1f}s)")
    b.free()


def test_served_stream_against_oracle(ctx):
    from pgd_amd.workloads import batched_graph
    pool = [batched_graph(i, seed=9)[:3] for i in range(400)]
    t0 = time.time()
    pool_sums = [run_pipeline(OracleGraph.from_edges(s, d, n), micro=False)["sums"] for n, s, d in pool]
    say(f"[gpu] served: oracle sums of the {len(pool)}-graph pool ({time.time() - t0:.1f}s)")
    large = []
    for sd in (1, 2):
        s, d = pgd_amd.gen_rmat(18, 16, 0.57, 0.19, 0.19, sd)
        n = int(max(s.max(), d.max())) + 1
        g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
        g.sort_edges("deg")
        large.append(((n, s, d), g.decompose(micro=True), g.micro("local_4_cycle")))
        g.free()
    srv = pgd_amd.Server(ctx=ctx, hbm_budget=96 << 30)
    rng = np.random.default_rng(2)
    want = {}
    start = time.perf_counter()
    t, k = 0.0, 0
    while t < 1.0:
        t += rng.exponential(1.0 / 2000.0)
        now = time.perf_counter() - start
        if t > now:
            time.sleep(t - now)
        if rng.random() < 0.01:
            j = k % 2
            rid = srv.submit(*large[j][0], micro=True)
            want[rid] = ("large", j)
        else:
            j = k % len(pool)
            rid = srv.submit(*pool[j], micro=False)
            want[rid] = ("pool", j)
        k += 1
    srv.drain(600000)
    nlarge = 0
    for rid, (kind, j) in want.items():
        r = srv.result(rid)
        got = r["sums"]
        if kind == "pool":
            assert got == pool_sums[j], (rid, j)
        else:
            nlarge += 1
            assert got == large[j][1], (rid, j)
            assert np.array_equal(srv.micro(rid, "local_4_cycle", r["m"]), large[j][2]), rid
        srv.release(rid)
    st = srv.stats()
    say(f"[gpu] served: {len(want)} requests ({nlarge} R-MAT-18) equal the oracle / plain path; "
        f"{st['finished'] / st['elapsed_s']:.0f} graphs/s, p99 {st['p99_ms']:.1f} ms")
    assert nlarge >= 1
    srv.free()
