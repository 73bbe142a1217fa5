"""Seeded graph inputs of the golden cases (shared by the generator script and the tests)."""
from __future__ import annotations

import hashlib

import numpy as np

import pgd_amd


def gen_cases():
    """Seeded synthetic inputs: (name, kind, params)."""
    cases = [
        ("gnp_2000_0.01", "gnp", dict(n=2000, p=0.01, seed=1)),
        ("gnp_500_0.9", "gnp", dict(n=500, p=0.9, seed=3)),
        ("gnp_300_0.3", "gnp", dict(n=300, p=0.3, seed=7)),
        ("ba_3000_4", "ba", dict(n=3000, m=4, seed=11)),
        ("ba_1000_8", "ba", dict(n=1000, m=8, seed=12)),
        ("rmat_10", "rmat", dict(scale=10, ef=16, seed=1)),
        ("rmat_12", "rmat", dict(scale=12, ef=16, seed=2)),
        ("rmat_14", "rmat", dict(scale=14, ef=16, seed=3)),
    ]
    for i in range(6):
        cases.append((f"gnp_small_{i}", "gnp", dict(n=30 + 5 * i, p=0.15 + 0.1 * i, seed=100 + i)))
    return cases


def make_graph(kind: str, prm: dict):
    if kind == "gnp":
        s, d = pgd_amd.gen_gnp(prm["n"], prm["p"], prm["seed"])
        return prm["n"], s, d
    if kind == "ba":
        s, d = pgd_amd.gen_ba(prm["n"], prm["m"], prm["seed"])
        return prm["n"], s, d
    if kind == "rmat":
        s, d = pgd_amd.gen_rmat(prm["scale"], prm["ef"], 0.57, 0.19, 0.19, prm["seed"])
        return 1 << prm["scale"], s, d
    if kind == "bigid":
        # star with `leaves` leaves, `pairs` disjoint edges, and one edge to id n-1:
        # n ~ 2^24 makes n^3 (signed) and n^4 (unsigned) wrap and N8 overflow
        n, leaves, pairs = prm["n"], prm["leaves"], prm["pairs"]
        s = [0] * leaves + [leaves + 1 + 2 * i for i in range(pairs)] + [leaves + 2 * pairs + 2]
        d = list(range(1, leaves + 1)) + [leaves + 2 + 2 * i for i in range(pairs)] + [n - 1]
        return n, np.array(s, dtype=np.int32), np.array(d, dtype=np.int32)
    raise ValueError(kind)


def edge_hash(src, dst) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(src, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(dst, dtype=np.int32).tobytes())
    return h.hexdigest()
