"""Dense tier: per-edge counts from int8 MFMA products equal the sparse
oriented engine's and the oracle's, bit for bit, on the dense configurations
G(500, 0.9) and G(2000, 0.5) and on dense graphs of awkward sizes."""
from __future__ import annotations

import numpy as np
import pytest

import pgd_amd
from tests.oracle_lib import OracleGraph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def both_engines(ctx, s, d, n, micro=True):
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
    g.sort_edges("deg")
    out = {}
    for eng in ("orient", "dense"):
        sums = g.decompose(micro=micro, engine=eng)
        assert g.last_engine() == eng
        out[eng] = (sums, g.edge_counts(), [g.micro(c) for c in range(8)] if micro else None)
    return g, out


def assert_same(out):
    a, b = out["orient"], out["dense"]
    assert a[0] == b[0]
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)
    if a[2] is not None:
        for x, y in zip(a[2], b[2]):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("n,p,seed", [(500, 0.9, 1), (2000, 0.5, 1)])
def test_dense_configurations(n, p, seed, ctx):
    say(f"[gpu] dense G({n},{p}) seed {seed}")
    s, d = pgd_amd.gen_gnp(n, p, seed)
    g, out = both_engines(ctx, s, d, n)
    assert_same(out)
    # left to itself the library picks the faster engine: the dense tier for
    # every graph up to 4096 vertices
    g.decompose(micro=False)
    assert g.last_engine() == "dense"
    # the oracle on a sample of edges
    T, K4, C4 = out["dense"][1]
    og = OracleGraph.from_edges(s, d, n)
    og.order_neighbors("deg")
    ev, eu = og.edge_arrays()
    sample = np.sort(np.random.default_rng(seed).choice(len(ev), size=min(len(ev), 1500), replace=False))
    res = og.decompose(ev, eu, sample.astype(np.int32), micro=True)
    assert np.array_equal(T[sample], res["tri"])
    assert np.array_equal(K4[sample], res["k4"])
    assert np.array_equal(C4[sample], res["c4"])
    g.free()


def complete(n):
    i, j = np.triu_indices(n, 1)
    return i.astype(np.int32), j.astype(np.int32)


# padded sizes N = 128 to 8064; the fan (a hub joined to every vertex of a
# path, n = 4096) gives one vertex a local graph of 4095 vertices, so the
# 4-clique work area holds 256 vertices at a time (16 chunks)
@pytest.mark.parametrize("case", ["k64", "k100", "gnp65", "gnp130_0.7", "gnp777_0.3", "gnp4000_0.3", "gnp6500_0.05",
                                  "gnp8000_0.02", "sparse_ba", "fan4096"])
def test_dense_matches_sparse(case, ctx):
    say(f"[gpu] dense vs sparse {case}")
    if case == "k64":
        n, (s, d) = 64, complete(64)
    elif case == "k100":
        n, (s, d) = 100, complete(100)
    elif case == "gnp65":
        n, (s, d) = 65, pgd_amd.gen_gnp(65, 0.5, 3)
    elif case == "sparse_ba":
        n, (s, d) = 3000, pgd_amd.gen_ba(3000, 3, 8)    # requested explicitly on a sparse graph
    elif case == "fan4096":
        n = 4096
        leaves = np.arange(1, n, dtype=np.int32)
        s = np.concatenate([np.zeros(n - 1, np.int32), leaves[:-1]])
        d = np.concatenate([leaves, leaves[1:]])
    else:
        _, nn, pp = case.split("_")[0][:3], int(case.split("_")[0][3:]), float(case.split("_")[1])
        n, (s, d) = nn, pgd_amd.gen_gnp(nn, pp, 5)
    g, out = both_engines(ctx, s, d, n, micro=(n < 4000))
    assert_same(out)
    if case.startswith("k"):
        # closed forms of the complete graph K_n, per edge
        T, K4, C4 = out["dense"][1]
        assert np.all(T == n - 2)
        assert np.all(K4 == (n - 2) * (n - 3) // 2)
        assert np.all(C4 == 0)
    if case == "fan4096":
        # triangles: hub-leaf edges carry the leaf's path degree, path edges
        # the hub; the hub's chords leave no 4-clique and no induced 4-cycle
        T, K4, C4 = out["dense"][1]
        ev, eu = g.edge_arrays()
        lo, hi = np.minimum(ev, eu), np.maximum(ev, eu)
        want = np.where(lo == 0, np.where((hi == 1) | (hi == n - 1), 1, 2), 1)
        assert np.array_equal(T, want.astype(T.dtype))
        assert not K4.any() and not C4.any()
    g.free()


@pytest.mark.parametrize("p,engine", [(0.02, "orient"), (0.1, "dense")])
def test_automatic_engine_above_4096_vertices(p, engine, ctx):
    # above 4096 vertices the dense tier is picked from 8% of all pairs on
    # (where it is the faster engine), and both engines agree
    say(f"[gpu] automatic engine G(5000,{p})")
    s, d = pgd_amd.gen_gnp(5000, p, 3)
    g = pgd_amd.GraphletCore(s, d, n=5000, ctx=ctx)
    g.sort_edges("deg")
    auto = g.decompose(micro=False)
    assert g.last_engine() == engine
    other = "dense" if engine == "orient" else "orient"
    assert g.decompose(micro=False, engine=other) == auto
    g.free()


def test_automatic_engine_respects_memory_cap():
    # under a cap the dense tier's arrays would pass, the automatic choice
    # takes the oriented engine (same counts)
    s, d = pgd_amd.gen_gnp(3000, 0.01, 4)
    with pgd_amd.Context(0) as c:
        g = pgd_amd.GraphletCore(s, d, n=3000, ctx=c)
        g.sort_edges("deg")
        free = g.decompose(micro=False)
        assert g.last_engine() == "dense"
        c.set_mem_cap(c.mem_in_use() + (64 << 20))    # 64 MiB more: below the tier's 18 N^2 bytes at N = 3072
        capped = g.decompose(micro=False)
        assert g.last_engine() == "orient"
        assert capped == free
        g.free()


def test_dense_rejects_large(ctx):
    s, d = pgd_amd.gen_ba(9000, 2, 1)
    g = pgd_amd.GraphletCore(s, d, ctx=ctx)
    g.sort_edges("deg")
    with pytest.raises(pgd_amd.PGDError) as e:
        g.decompose(engine="dense")
    assert e.value.code == -5
    g.decompose()
    assert g.last_engine() == "orient"
    g.free()
