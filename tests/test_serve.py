"""Served form: requests that arrive while others are on the GPU are admitted
by estimated HBM footprint and counted in overlapping batches; every result
is bit-identical to the same graph counted alone and to the plain path."""
from __future__ import annotations

import numpy as np
import pytest

import pgd_amd
from pgd_amd.workloads import batched_graph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def plain(ctx, n, s, d, dedup=False):
    g = pgd_amd.GraphletCore(s, d, n=n, dedup=dedup, ctx=ctx)
    g.sort_edges("deg")
    sums = g.decompose(micro=True, engine="orient")
    cols = [g.micro(c) for c in range(8)] if g.m else [np.zeros(0, np.uint64)] * 8
    m = g.m
    g.free()
    return m, sums, pgd_amd.derive(n, m, sums), cols


def test_alone_equals_in_the_mix(ctx):
    probe = batched_graph(3, seed=21, n_hi=2000)[:3]
    rmat = pgd_amd.gen_rmat(12, 16, seed=4)
    rn = int(max(rmat[0].max(), rmat[1].max())) + 1
    # alone
    srv = pgd_amd.Server(ctx=ctx)
    rid = srv.submit(*probe, micro=True)
    srv.drain(60000)
    alone = srv.result(rid)
    alone_cols = [srv.micro(rid, c, alone["m"]) for c in range(8)]
    srv.free()
    # in a mix of 400 other requests (and an R-MAT graph), submitted around it
    srv = pgd_amd.Server(ctx=ctx)
    others = [batched_graph(i, seed=22, n_hi=2500)[:3] for i in range(400)]
    ids = []
    for i, g in enumerate(others):
        ids.append(srv.submit(*g, micro=(i % 3 == 0)))
        if i == 200:
            rid = srv.submit(*probe, micro=True)
            rr = srv.submit(rn, rmat[0], rmat[1], micro=False)
    srv.drain(120000)
    mixed = srv.result(rid)
    assert mixed["sums"] == alone["sums"] and mixed["totals"] == alone["totals"] and mixed["m"] == alone["m"]
    for c in range(8):
        assert np.array_equal(srv.micro(rid, c, mixed["m"]), alone_cols[c]), c
    # the plain path agrees with both, for the probe, the R-MAT graph and a sample of the mix
    m, psums, ptot, pcols = plain(ctx, *probe)
    assert alone["sums"] == psums and alone["totals"] == ptot
    assert all(np.array_equal(alone_cols[c], pcols[c]) for c in range(8))
    r = srv.result(rr)
    _, rs, rt, _ = plain(ctx, rn, rmat[0], rmat[1])
    assert r["sums"] == rs and r["totals"] == rt
    for i in range(0, 400, 37):
        res = srv.result(ids[i])
        m, psums, ptot, pcols = plain(ctx, *others[i])
        assert res["sums"] == psums and res["totals"] == ptot, i
        if i % 3 == 0:
            assert np.array_equal(srv.micro(ids[i], "local_4_cycle", m), pcols[4]), i
    st = srv.stats()
    say(f"[gpu] serve stats {st}")
    assert st["finished"] == 402 and st["batches"] >= 2
    assert st["p99_ms"] >= st["p50_ms"] > 0
    done = srv.poll()
    assert sorted(done) == sorted(ids + [rid, rr])
    srv.free()


def test_admission_by_footprint(ctx):
    graphs = [batched_graph(i, seed=31, n_hi=800)[:3] for i in range(60)]
    fps = [pgd_amd.server_footprint(n, len(s)) for n, s, _ in graphs]
    budget = 3 * max(fps)                      # a handful of requests in flight at a time
    srv = pgd_amd.Server(ctx=ctx, hbm_budget=budget, max_inflight=8)
    ids = [srv.submit(*g) for g in graphs]
    # a request larger than the whole budget is refused outright
    big = pgd_amd.gen_ba(20000, 8, 1)
    with pytest.raises(pgd_amd.PGDError) as e:
        srv.submit(20000, big[0], big[1])
    assert e.value.code == -7  # PGD_ERR_BUSY
    srv.drain(120000)
    st = srv.stats()
    assert st["finished"] == len(graphs)
    assert st["batches"] >= len(graphs) // 8     # never more than 8 requests in flight
    for i in (0, 17, 59):
        res = srv.result(ids[i])
        _, psums, _, _ = plain(ctx, *graphs[i])
        assert res["sums"] == psums
        srv.release(ids[i])
        with pytest.raises(pgd_amd.PGDError):
            srv.result(ids[i])
    srv.free()


def test_flags_and_failures_are_per_request(ctx):
    s = np.array([0, 1, 1, 2, 0], np.int32)
    d = np.array([1, 0, 2, 1, 2], np.int32)
    srv = pgd_amd.Server(ctx=ctx)
    a = srv.submit(3, s, d, dedup=True)                       # repeated pair, deduplicated
    b = srv.submit(3, s, d)                                   # repeated pair, a multigraph: unsupported
    c = srv.submit(*batched_graph(1, seed=5, n_hi=300)[:3])
    srv.drain(60000)
    _, ref, _, _ = plain(ctx, 3, s, d, dedup=True)
    assert srv.result(a)["sums"] == ref
    with pytest.raises(pgd_amd.PGDError):
        srv.result(b)
    assert srv.result(c)["m"] > 0
    srv.free()
