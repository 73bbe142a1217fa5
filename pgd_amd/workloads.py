"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

batched_graphs(): the batched configuration -- graphs drawn from a seed as
Barabasi-Albert graphs (attachments m in [2, 8]) or G(n, p) graphs with
average degree in [4, 32], n log-uniform on [20, 5000], the two families in
equal proportion (alternating by index).
"""
from __future__ import annotations

import numpy as np

from . import gen_ba, gen_gnp


def batched_graph(i: int, seed: int = 1, n_lo: int = 20, n_hi: int = 5000):
    """Graph i of the batched workload: (n, src, dst, family, parameter)."""
    rng = np.random.default_rng([seed, i])
    n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
    gseed = int(rng.integers(1, 2**31 - 1))
    if i % 2 == 0:
        k = int(rng.integers(2, 9))
        k = min(k, max(1, n - 1))
        s, d = gen_ba(n, k, gseed)
        return n, s, d, "ba", k
    deg = float(rng.uniform(4.0, 32.0))
    p = min(1.0, deg / max(n - 1, 1))
    s, d = gen_gnp(n, p, gseed)
    return n, s, d, "gnp", p


def batched_graphs(count: int = 8192, seed: int = 1, n_lo: int = 20, n_hi: int = 5000):
    """List of (n, src, dst) for the batched workload."""
    out = []
    for i in range(count):
        n, s, d, _, _ = batched_graph(i, seed, n_lo, n_hi)
        out.append((n, s, d))
    return out
