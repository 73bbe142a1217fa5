"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP path against it, and it is
itself pinned against golden vectors, brute force and the compiled reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_BIN = os.path.join(ORACLE_DIR, "_ref", "pgd")

NSUMS, NTOTALS = 15, 17


class CSR(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("row", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int32))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        L = C.CDLL(LIB)
        P32, P64, PU = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)
        L.or_read_graph.argtypes = [C.c_char_p, C.POINTER(CSR), P64]
        L.or_read_graph.restype = C.c_int
        L.or_free.argtypes = [C.POINTER(CSR)]
        L.or_csr_from_arrays.argtypes = [C.c_int64, C.c_int64, P64, P32, C.POINTER(CSR)]
        L.or_degrees.argtypes = [C.POINTER(CSR), P32, P32, P32]
        L.or_kcore.argtypes = [C.POINTER(CSR), P32]
        L.or_kcore.restype = C.c_int32
        L.or_order_neighbors.argtypes = [C.POINTER(CSR), C.c_char_p, C.c_int]
        L.or_edge_arrays.argtypes = [C.POINTER(CSR), P32, P32]
        L.or_edge_arrays.restype = C.c_int64
        L.or_sort_edges.argtypes = [C.POINTER(CSR), P32, P32, C.c_int64, C.c_char_p, C.c_int, P32]
        L.or_sort_edges.restype = C.c_int
        L.or_sort_edges_subset.argtypes = [C.POINTER(CSR), P32, P32, P32, C.c_int64, C.c_char_p, C.c_int, P32]
        L.or_sort_edges_subset.restype = C.c_int
        L.or_decompose.argtypes = [C.POINTER(CSR), P32, P32, P32, C.c_int64, C.c_int64, C.c_int, PU,
                                   PU, PU, PU, PU, PU, PU]
        L.or_derive.argtypes = [C.c_int64, C.c_int64, PU, PU]
        L.or_brute_edge.argtypes = [C.POINTER(CSR), C.c_int32, C.c_int32, PU]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class OracleGraph:
    """A CSR owned by the oracle library."""

    def __init__(self, csr: CSR):
        self.c = csr

    @classmethod
    def from_file(cls, path: str) -> "OracleGraph":
        g = CSR()
        loops = C.c_int64()
        rc = lib().or_read_graph(path.encode(), C.byref(g), C.byref(loops))
        if rc != 0:
            raise ValueError(f"oracle reader rejected {path} ({rc})")
        return cls(g)

    @classmethod
    def from_csr(cls, row: np.ndarray, col: np.ndarray) -> "OracleGraph":
        row = np.ascontiguousarray(row, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        g = CSR()
        lib().or_csr_from_arrays(len(row) - 1, len(col), _p(row, C.c_int64), _p(col, C.c_int32), C.byref(g))
        return cls(g)

    @classmethod
    def from_edges(cls, src, dst, n: int) -> "OracleGraph":
        """Adjacency lists in line order (the reference reader's layout)."""
        lists = [[] for _ in range(n)]
        for a, b in zip(src, dst):
            if a == b:
                continue
            lists[a].append(b)
            lists[b].append(a)
        row = np.zeros(n + 1, dtype=np.int64)
        row[1:] = np.cumsum([len(x) for x in lists])
        col = np.array([y for x in lists for y in x], dtype=np.int32)
        return cls.from_csr(row, col)

    def __del__(self):
        try:
            lib().or_free(C.byref(self.c))
        except Exception:
            pass

    @property
    def n(self) -> int:
        return self.c.n

    @property
    def m(self) -> int:
        return self.c.nnz // 2

    def csr(self):
        row = np.ctypeslib.as_array(self.c.row, shape=(self.c.n + 1,)).copy()
        col = np.ctypeslib.as_array(self.c.col, shape=(max(self.c.nnz, 1),))[: self.c.nnz].copy()
        return row, col

    def degrees(self):
        deg = np.zeros(max(self.n, 1), dtype=np.int32)
        md, ad = C.c_int32(), C.c_int32()
        lib().or_degrees(C.byref(self.c), _p(deg, C.c_int32), C.byref(md), C.byref(ad))
        return deg[: self.n], md.value, ad.value

    def kcore(self):
        core = np.zeros(max(self.n, 1), dtype=np.int32)
        K = lib().or_kcore(C.byref(self.c), _p(core, C.c_int32))
        return K, core[: self.n]

    def order_neighbors(self, method: str = "deg", s2l: bool = False) -> None:
        lib().or_order_neighbors(C.byref(self.c), method.encode(), int(s2l))

    def edge_arrays(self):
        m = self.m
        ev = np.zeros(max(m, 1), dtype=np.int32)
        eu = np.zeros(max(m, 1), dtype=np.int32)
        k = lib().or_edge_arrays(C.byref(self.c), _p(ev, C.c_int32), _p(eu, C.c_int32))
        return ev[:k], eu[:k]

    def sort_edges(self, ev, eu, ordering: str = "deg", s2l: bool = False):
        m = len(ev)
        out = np.zeros(max(m, 1), dtype=np.int32)
        rc = lib().or_sort_edges(C.byref(self.c), _p(ev, C.c_int32), _p(eu, C.c_int32), m, ordering.encode(),
                                 int(s2l), _p(out, C.c_int32))
        if rc != 0:
            raise ValueError(f"ordering {ordering} has no reproducible key")
        return out[:m]

    def sort_edges_subset(self, ev, eu, es, ordering: str = "deg", s2l: bool = False):
        """sort_edges over a sampled edge set E_s (graphlet_core.cpp:836-851)."""
        es = np.ascontiguousarray(es, dtype=np.int32)
        out = np.zeros(max(len(es), 1), dtype=np.int32)
        rc = lib().or_sort_edges_subset(C.byref(self.c), _p(ev, C.c_int32), _p(eu, C.c_int32), _p(es, C.c_int32),
                                        len(es), ordering.encode(), int(s2l), _p(out, C.c_int32))
        if rc != 0:
            raise ValueError(f"ordering {ordering} has no reproducible key")
        return out[: len(es)]

    def decompose(self, ev, eu, order, micro: bool = True, threads: int = 0, count: int | None = None):
        cnt = len(order) if count is None else count
        sums = np.zeros(NSUMS, dtype=np.uint64)
        cols = [np.zeros(max(cnt, 1), dtype=np.uint64) for _ in range(6)] if micro else None
        ptrs = [_p(c, C.c_uint64) for c in cols] if micro else [None] * 6
        order = np.ascontiguousarray(order, dtype=np.int32)
        lib().or_decompose(C.byref(self.c), _p(ev, C.c_int32), _p(eu, C.c_int32), _p(order, C.c_int32), cnt,
                           self.m, threads, _p(sums, C.c_uint64), *ptrs)
        res = {"sums": [int(x) for x in sums]}
        if micro:
            for name, c in zip(["tri", "k4", "c4", "chordal", "path4", "star3"], cols):
                res[name] = c[:cnt]
        return res

    def brute_edge(self, v: int, u: int):
        out = np.zeros(5, dtype=np.uint64)
        lib().or_brute_edge(C.byref(self.c), v, u, _p(out, C.c_uint64))
        return [int(x) for x in out]


def derive(n: int, m: int, sums) -> list[int]:
    s = np.ascontiguousarray(np.asarray(sums, dtype=np.uint64))
    t = np.zeros(NTOTALS, dtype=np.uint64)
    lib().or_derive(n, m, _p(s, C.c_uint64), _p(t, C.c_uint64))
    return [int(x) for x in t]


def run_pipeline(g: OracleGraph, neigh: str = "deg", s2l_neigh: bool = False, ordering: str = "deg",
                 s2l: bool = False, micro: bool = True, threads: int = 0) -> dict:
    """The reference driver's sequence: neighbour order, edge arrays, k-core,
    sort_edges, decomposition, derivation (graphlet_driver.cpp:53-92)."""
    g.order_neighbors(neigh, s2l_neigh)
    ev, eu = g.edge_arrays()
    K, _ = g.kcore()
    order = g.sort_edges(ev, eu, ordering, s2l)
    res = g.decompose(ev, eu, order, micro=micro, threads=threads)
    res.update(ev=ev, eu=eu, order=order, K=K, n=g.n, m=g.m)
    res["totals"] = derive(g.n, g.m, res["sums"])
    return res


TOTAL_NAMES = [
    "total_2_1edge", "total_2_indep", "total_3_tris", "total_2_star", "total_3_1edge", "total_3_indep",
    "total_4_clique", "total_4_chordcycle", "total_4_tailed_tris", "total_4_cycle", "total_3_star",
    "total_4_path", "total_4_1edge", "total_4_2edge", "total_4_2star", "total_4_tri", "total_4_indep",
]


def micro_rows(res: dict) -> list[tuple]:
    """Rows of the .micro file (graphlet_core.cpp:1668-1761) in E_ordered order."""
    ev, eu, order = res["ev"], res["eu"], res["order"]
    rows = []
    for e in range(len(order)):
        i = order[e]
        t = int(res["tri"][e])
        s2 = None  # filled below
        rows.append([int(ev[i]) + 1, int(eu[i]) + 1, t, s2, int(res["k4"][e]), int(res["chordal"][e]), 0,
                     int(res["c4"][e]), int(res["star3"][e]), int(res["path4"][e])])
    return rows
