"""Batched form: every graph of a batch gets exactly the plain path's result
(per-graph sums, derived totals and per-edge micro columns), the captured
launch replays identically, and launch parameters (set by hand or by the
autotuner) never change a count."""
from __future__ import annotations

import numpy as np
import pytest

import pgd_amd
from pgd_amd.workloads import batched_graph, batched_graphs
from tests.oracle_lib import OracleGraph, run_pipeline

pytestmark = pytest.mark.gpu

MICRO_KEYS = ["tri", "k4", "chordal", None, "c4", "star3", "path4", None]


def say(*a):
    print(*a, flush=True)


def edge_cases():
    """Graphs at the edges of the domain, mixed into the batches."""
    k5 = [(i, j) for i in range(5) for j in range(i + 1, 5)]
    return [
        (6, np.zeros(0, np.int32), np.zeros(0, np.int32)),                  # no edges
        (2, np.array([0], np.int32), np.array([1], np.int32)),              # one edge
        (5, np.array([a for a, _ in k5], np.int32), np.array([b for _, b in k5], np.int32)),  # K5
        (9, np.array([0, 1, 2, 3, 5], np.int32), np.array([1, 2, 3, 0, 6], np.int32)),  # C4 + edge + isolated
        (0, np.zeros(0, np.int32), np.zeros(0, np.int32)),                  # no vertices
    ]


def plain(ctx, n, s, d):
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
    g.sort_edges("deg")
    sums = g.decompose(micro=True, engine="orient")
    tot = pgd_amd.derive(n, g.m, sums)
    cols = [g.micro(c) for c in range(8)] if g.m else [np.zeros(0, np.uint64)] * 8
    m = g.m
    g.free()
    return m, sums, tot, cols


def test_batch_matches_plain_path_and_oracle(ctx):
    graphs = [batched_graphs(1, seed=7)[0]]
    graphs += [batched_graph(i, seed=7, n_hi=700)[:3] for i in range(1, 60)]
    graphs += edge_cases()
    say(f"[gpu] batch of {len(graphs)} graphs")
    b = pgd_amd.Batch(graphs, ctx=ctx, micro=True)
    assert b.ngraphs == len(graphs)
    sums, tots = b.run()
    for gi, (n, s, d) in enumerate(graphs):
        m, psums, ptot, pcols = plain(ctx, n, s, d)
        assert [int(x) for x in sums[gi]] == psums, gi
        assert [int(x) for x in tots[gi]] == ptot, gi
        for c in range(8):
            assert np.array_equal(b.micro(gi, c, m), pcols[c]), (gi, c)
        if gi % 7 == 0 and n > 0:
            og = OracleGraph.from_edges(s, d, n)
            res = run_pipeline(og)
            assert [int(x) for x in sums[gi]] == [int(x) for x in res["sums"]], gi
            assert [int(x) for x in tots[gi]] == [int(x) for x in res["totals"]], gi
            for c, key in enumerate(MICRO_KEYS):
                if key:
                    assert np.array_equal(b.micro(gi, c, m), np.asarray(res[key], dtype=np.uint64)), (gi, key)
    b.free()


def test_batch_replays_are_identical(ctx):
    """The captured sequence replayed several times, with other device work
    and synchronisations in between, gives the same result every time."""
    graphs = [batched_graph(i, seed=3, n_hi=1500)[:3] for i in range(200)]
    b = pgd_amd.Batch(graphs, ctx=ctx, micro=True)
    s0, t0 = b.run()
    m5 = len(graphs[5][1])  # generated graphs have no repeated pairs or self loops
    m0 = b.micro(5, "local_4_cycle", m5)
    for r in range(4):
        g = pgd_amd.GraphletCore(*pgd_amd.gen_ba(500, 3, r + 1), ctx=ctx)   # eager work between replays
        g.graphlet_decomposition()
        g.free()
        s, t = b.run()
        assert np.array_equal(s, s0) and np.array_equal(t, t0), r
        assert np.array_equal(b.micro(5, "local_4_cycle", m5), m0)
    b.free()


def test_batch_tuning_and_autotune_keep_counts(ctx):
    graphs = [batched_graph(i, seed=11, n_hi=3000)[:3] for i in range(300)]
    b = pgd_amd.Batch(graphs, ctx=ctx, micro=False)
    s0, _ = b.run()
    for tun in (dict(wedge_log2r=8, wedge_nt=256), dict(small_k=32, mid_k=128, wedge_split=16),
                dict(wedge_log2r=15, wedge_nt=1024, grid_small=8)):
        b.set_tuning(**tun)
        s, _ = b.run()
        assert np.array_equal(s, s0), tun
    best, ms = b.autotune(trials=1)
    assert ms > 0
    s, _ = b.run()
    assert np.array_equal(s, s0)
    assert b.tuning().as_dict() == best.as_dict()
    b.free()


def test_batch_dedup_and_invalid(ctx):
    # duplicate lines: with dedup each graph equals its deduplicated plain run
    s = np.array([0, 1, 1, 2, 0, 2, 0], np.int32)
    d = np.array([1, 0, 2, 1, 2, 0, 1], np.int32)
    b = pgd_amd.Batch([(3, s, d), (3, s, d)], ctx=ctx, dedup=True)
    sums, tots = b.run()
    g = pgd_amd.GraphletCore(s, d, n=3, dedup=True, ctx=ctx)
    g.sort_edges("deg")
    ref = g.decompose(micro=True, engine="orient")
    g.free()
    assert [int(x) for x in sums[0]] == ref and [int(x) for x in sums[1]] == ref
    b.free()
    # an id outside its graph's range is rejected, not folded into a neighbour graph
    with pytest.raises(pgd_amd.PGDError):
        pgd_amd.Batch([(3, np.array([0], np.int32), np.array([3], np.int32)),
                       (3, np.array([0], np.int32), np.array([1], np.int32))], ctx=ctx)
