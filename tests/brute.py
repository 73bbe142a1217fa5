"""Independent checks fixed by the mathematics (pure Python, tiny graphs).

* classify every 4-subset, 3-subset and pair of vertices by its induced
  subgraph: this pins all 17 totals without any of the Lemma algebra;
* per-edge sets by direct set algebra: T, Sv, Su, K4, C4 and the micro
  columns as the reference defines them (graphlet_core.cpp:1559-1575).
"""
from __future__ import annotations

import itertools
from collections import Counter, defaultdict

NAMES = ["total_2_1edge", "total_2_indep", "total_3_tris", "total_2_star", "total_3_1edge", "total_3_indep",
         "total_4_clique", "total_4_chordcycle", "total_4_tailed_tris", "total_4_cycle", "total_3_star",
         "total_4_path", "total_4_1edge", "total_4_2edge", "total_4_2star", "total_4_tri", "total_4_indep"]


def adjacency(n: int, edges) -> list[set]:
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    return adj


def subset_totals(n: int, adj: list[set]) -> dict[str, int]:
    c = Counter()
    for a, b in itertools.combinations(range(n), 2):
        c["total_2_1edge" if b in adj[a] else "total_2_indep"] += 1
    for q in itertools.combinations(range(n), 3):
        k = sum(1 for a, b in itertools.combinations(q, 2) if b in adj[a])
        c[{3: "total_3_tris", 2: "total_2_star", 1: "total_3_1edge", 0: "total_3_indep"}[k]] += 1
    for q in itertools.combinations(range(n), 4):
        es = [(a, b) for a, b in itertools.combinations(q, 2) if b in adj[a]]
        deg = Counter()
        for a, b in es:
            deg[a] += 1
            deg[b] += 1
        ds = tuple(sorted((deg[x] for x in q), reverse=True))
        k = len(es)
        if k == 6:
            name = "total_4_clique"
        elif k == 5:
            name = "total_4_chordcycle"
        elif k == 4:
            name = "total_4_cycle" if ds == (2, 2, 2, 2) else "total_4_tailed_tris"
        elif k == 3:
            name = {(3, 1, 1, 1): "total_3_star", (2, 2, 1, 1): "total_4_path", (2, 2, 2, 0): "total_4_tri"}[ds]
        elif k == 2:
            name = "total_4_2star" if ds[0] == 2 else "total_4_2edge"
        elif k == 1:
            name = "total_4_1edge"
        else:
            name = "total_4_indep"
        c[name] += 1
    return {k: c.get(k, 0) for k in NAMES}


def edge_sets(adj: list[set], v: int, u: int):
    Nv, Nu = adj[v] - {u}, adj[u] - {v}
    T = Nv & Nu
    Sv, Su = Nv - Nu, Nu - Nv
    K4 = sum(1 for a, b in itertools.combinations(sorted(T), 2) if b in adj[a])
    C4 = sum(1 for w in Su for x in adj[w] if x in Sv)
    return len(T), len(Sv), len(Su), K4, C4


def micro_columns(adj: list[set], v: int, u: int) -> list[int]:
    """tri, 2-star, 4-clique, chordal, tailed (0), 4-cycle, 3-star, 4-path."""
    t, sv, su, k4, c4 = edge_sets(adj, v, u)
    return [t, sv + su, k4, t * (t - 1) // 2 - k4, 0, c4, sv * (sv - 1) // 2 + su * (su - 1) // 2, sv * su - c4]


def lemma_sums(n: int, adj: list[set], edges) -> list[int]:
    """The 15 sums of the decomposition loop, computed from edge_sets."""
    m = len(edges)
    s = [0] * 15
    for a, b in edges:
        t, sv, su, k4, c4 = edge_sets(adj, a, b)
        S = sv + su
        D = n - S - t - 2
        s[0] += t; s[1] += S; s[2] += D; s[3] += k4; s[4] += c4
        s[5] += (t + S) * (t + S - 1) // 2
        s[6] += S * (S - 1) // 2
        s[7] += t * (t - 1) // 2
        s[8] += t * S
        s[9] += sv * su
        s[10] += sv * (sv - 1) // 2 + su * (su - 1) // 2
        s[11] += m - (sv + t + 1) - (su + t + 1) + 1
        s[12] += D * (D - 1) // 2
        s[13] += S * D
        s[14] += t * D
    return [x % (1 << 64) for x in s]


def random_graph(n: int, p: float, rng) -> list[tuple[int, int]]:
    return [(a, b) for a, b in itertools.combinations(range(n), 2) if rng.random() < p]


def closed_forms(kind: str, k: int) -> dict[str, int]:
    """Known totals for the clique K_k and the star with k leaves."""
    from math import comb
    z = {name: 0 for name in NAMES}
    if kind == "clique":
        n = k
        z.update(total_2_1edge=comb(n, 2), total_3_tris=comb(n, 3), total_4_clique=comb(n, 4))
    elif kind == "star":
        n = k + 1
        z.update(total_2_1edge=k, total_2_indep=comb(k, 2), total_2_star=comb(k, 2), total_3_indep=comb(k, 3),
                 total_3_1edge=0, total_3_star=comb(k, 3), total_4_indep=comb(k, 4))
    return z
