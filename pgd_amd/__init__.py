"""pgd_amd -- MI355X implementation of PGD's exact 2/3/4-node graphlet decomposition.

Python mirror of the reference's library interface (graphlet::graphlet_core,
graphlet_core.h:54-703) over the C ABI in include/pgd_hip.h.  Every count is
computed by the HIP kernels in pgd_amd/csrc; there is no CPU fallback: if the
library or the GPU is missing, the calls raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from ._lib import (DEDUP, ENGINE_DENSE, ENGINE_ORIENT, MICRO, NMICRO, NSUMS, NTOTALS,
                   TOTAL_NAMES, PGDError, Tuning, check, lib, ptr)

__all__ = [
    "Context", "GraphletCore", "derive", "gen_gnp", "gen_ba", "gen_rmat", "write_mtx", "read_graph",
    "TOTAL_NAMES", "PGDError", "MICRO_COLUMNS", "Tuning", "default_tuning",
]

MICRO_COLUMNS = ["tri", "local_4_clique", "local_4_chordal_cycle", "local_4_tailed_tris", "local_4_cycle",
                 "local_3_star", "local_4_path", "2_star"]
_ENGINE = {None: 0, "orient": ENGINE_ORIENT, "dense": ENGINE_DENSE}


def default_tuning(**overrides) -> Tuning:
    """pgd_tuning_default, with fields replaced by keyword."""
    t = Tuning()
    check(lib().pgd_tuning_default(C.byref(t)), what="pgd_tuning_default")
    for k, v in overrides.items():
        if not hasattr(t, k):
            raise KeyError(k)
        setattr(t, k, int(v))
    return t


class Context:
    """A device context (pgd_open): one HIP stream and the device memory it owns."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().pgd_open(device, C.byref(h)), what="pgd_open")
        self.h = h
        self._graphs = weakref.WeakSet()   # graphs whose device memory this context owns
        _live_contexts.add(self)

    def close(self) -> None:
        """Free every graph of this context, then the context itself."""
        if self.h:
            for g in list(self._graphs):
                g.free()
            lib().pgd_close(self.h)
            self.h = None

    def set_mem_cap(self, nbytes: int) -> None:
        check(lib().pgd_set_mem_cap(self.h, nbytes), self.h, "pgd_set_mem_cap")

    def mem_in_use(self) -> int:
        v = C.c_int64()
        check(lib().pgd_mem_in_use(self.h, C.byref(v)), self.h)
        return v.value

    def atomic_rate(self, words: int = 1 << 27, reps: int = 10) -> tuple[float, float]:
        """(requests/s, ms per launch) of global atomic adds (pgd_atomic_rate)."""
        r, ms = C.c_double(), C.c_double()
        check(lib().pgd_atomic_rate(self.h, words, reps, C.byref(r), C.byref(ms)), self.h, "pgd_atomic_rate")
        return r.value, ms.value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


_default_ctx: Context | None = None
_live_contexts: "weakref.WeakSet[Context]" = weakref.WeakSet()


@atexit.register
def _release_device_objects() -> None:
    """Free device objects before the HIP runtime is torn down at exit."""
    for c in list(_live_contexts):
        c.close()


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ---------------------------------------------------------------------------
# host utilities (no GPU needed)
# ---------------------------------------------------------------------------
def derive(n: int, m: int, sums) -> list[int]:
    """The 17 totals from the 15 sums, bit-identical to the reference build."""
    s = np.ascontiguousarray(np.asarray(sums, dtype=np.uint64))
    t = np.zeros(NTOTALS, dtype=np.uint64)
    check(lib().pgd_derive(n, m, ptr(s, C.c_uint64), ptr(t, C.c_uint64)), what="pgd_derive")
    return [int(x) for x in t]


def _gen(fn, *args):
    m = fn(*args, None, None, 0)
    if m < 0:
        raise PGDError(-1, "generator arguments")
    src = np.empty(max(m, 1), dtype=np.int32)
    dst = np.empty(max(m, 1), dtype=np.int32)
    fn(*args, ptr(src, C.c_int32), ptr(dst, C.c_int32), m)
    return src[:m], dst[:m]


def gen_gnp(n: int, p: float, seed: int):
    """Erdos-Renyi G(n,p) edge list (src<dst, row-major pair order)."""
    return _gen(lib().pgd_gen_gnp, n, p, seed)


def gen_ba(n: int, m: int, seed: int):
    """Barabasi-Albert graph with m attachments per new vertex."""
    return _gen(lib().pgd_gen_ba, n, m, seed)


def gen_rmat(scale: int, ef: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1):
    """R-MAT graph, symmetrised, self loops and duplicates removed, sorted."""
    return _gen(lib().pgd_gen_rmat, scale, ef, a, b, c, seed)


def write_mtx(path: str, n: int, src, dst) -> None:
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    check(lib().pgd_write_mtx(path.encode(), n, len(src), ptr(src, C.c_int32), ptr(dst, C.c_int32)),
          what="pgd_write_mtx")


def read_graph(path: str):
    """Parse a file with the reference reader's rules (read_graph).

    Returns (n, src, dst, dedup, self_loops): the COO in line order."""
    n, nl, loops, dd = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
    ps, pd = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    check(lib().pgd_read_graph(path.encode(), C.byref(n), C.byref(nl), C.byref(ps), C.byref(pd), C.byref(dd),
                               C.byref(loops)), what=f"read_graph({path})")
    k = nl.value
    src = np.ctypeslib.as_array(ps, shape=(max(k, 1),))[:k].copy()
    dst = np.ctypeslib.as_array(pd, shape=(max(k, 1),))[:k].copy()
    lib().pgd_free(C.cast(ps, C.c_void_p))
    lib().pgd_free(C.cast(pd, C.c_void_p))
    return n.value, src, dst, bool(dd.value), loops.value


# ---------------------------------------------------------------------------
class GraphletCore:
    """Mirror of graphlet::graphlet_core for one graph on one device.

    Construction follows the reference constructor (graphlet_core.cpp:59-67):
    read (or take) the edge list, build the CSR, then preprocess (neighbour
    order + edge arrays).  graphlet_decomposition() runs sort_edges first when
    no edge order exists (graphlet_core.cpp:1376-1378)."""

    def __init__(self, src, dst, n: int | None = None, dedup: bool = False, ctx: Context | None = None,
                 neigh_ordering: str = "deg", neigh_small_to_large: bool = False, preprocess: bool = True):
        self.ctx = ctx or default_context()
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        if n is None:
            keep = src != dst
            n = int(max(src[keep].max(initial=-1), dst[keep].max(initial=-1))) + 1
        h = C.c_void_p()
        check(lib().pgd_graph_from_coo(self.ctx.h, n, len(src), ptr(src, C.c_int32), ptr(dst, C.c_int32),
                                       DEDUP if dedup else 0, C.byref(h)), self.ctx.h, "pgd_graph_from_coo")
        self.h = h
        self.ctx._graphs.add(self)
        nn, mm, md, ad = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        lib().pgd_graph_info(h, C.byref(nn), C.byref(mm), C.byref(md), C.byref(ad))
        self.n, self.m, self.max_degree, self.avg_degree = nn.value, mm.value, md.value, ad.value
        self.sums: list[int] | None = None
        self.totals: dict[str, int] | None = None
        self._ordered = False
        if preprocess:
            self.preprocess(neigh_ordering, neigh_small_to_large)

    @classmethod
    def from_csr(cls, row, col, ctx: Context | None = None) -> "GraphletCore":
        """Adopt a CSR as-is (graphlet_core's vertices/edges), keeping its
        neighbour order; preprocess with "natural" to keep the edge ids."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        row = np.ascontiguousarray(row, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        h = C.c_void_p()
        check(lib().pgd_graph_from_csr(self.ctx.h, len(row) - 1, ptr(row, C.c_int64), ptr(col, C.c_int32),
                                       C.byref(h)), self.ctx.h, "pgd_graph_from_csr")
        self.h = h
        self.ctx._graphs.add(self)
        nn, mm, md, ad = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        lib().pgd_graph_info(h, C.byref(nn), C.byref(mm), C.byref(md), C.byref(ad))
        self.n, self.m, self.max_degree, self.avg_degree = nn.value, mm.value, md.value, ad.value
        self.sums = None
        self.totals = None
        self._ordered = False
        return self

    @classmethod
    def from_file(cls, path: str, ctx: Context | None = None, neigh_ordering: str = "deg",
                  neigh_small_to_large: bool = False, preprocess: bool = True, threads: int = 0) -> "GraphletCore":
        """read_graph straight to the device (pgd_graph_from_file): the text
        parsed by `threads` host threads, renumbering and CSR on the device."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        loops = C.c_int64()
        check(lib().pgd_graph_from_file(self.ctx.h, path.encode(), threads, C.byref(h), C.byref(loops)),
              self.ctx.h, f"pgd_graph_from_file({path})")
        self.h = h
        self.ctx._graphs.add(self)
        nn, mm, md, ad = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        lib().pgd_graph_info(h, C.byref(nn), C.byref(mm), C.byref(md), C.byref(ad))
        self.n, self.m, self.max_degree, self.avg_degree = nn.value, mm.value, md.value, ad.value
        self.self_loops = loops.value
        self.sums = None
        self.totals = None
        self._ordered = False
        if preprocess:
            self.preprocess(neigh_ordering, neigh_small_to_large)
        return self

    def free(self) -> None:
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().pgd_graph_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # -- preprocessing ---------------------------------------------------
    def preprocess(self, neigh_ordering: str = "deg", small_to_large: bool = False) -> None:
        check(lib().pgd_preprocess(self.h, neigh_ordering.encode(), int(small_to_large)), self.ctx.h,
              "pgd_preprocess")
        self._ordered = False

    def compute_cores(self, want_cores: bool = False):
        K = C.c_int32()
        core = np.zeros(max(self.n, 1), dtype=np.int32) if want_cores else None
        check(lib().pgd_compute_cores(self.h, ptr(core, C.c_int32) if want_cores else None, C.byref(K)),
              self.ctx.h, "pgd_compute_cores")
        return (K.value, core[: self.n]) if want_cores else K.value

    def sort_edges(self, ordering: str = "deg", small_to_large: bool = False, E_s=None) -> None:
        """sort_edges (graphlet_core.cpp:854-867), or with E_s the overload
        over a sampled edge set (836-851): E_ordered holds only those ids."""
        if E_s is None:
            check(lib().pgd_sort_edges(self.h, ordering.encode(), int(small_to_large)), self.ctx.h,
                  "pgd_sort_edges")
        else:
            ids = np.ascontiguousarray(E_s, dtype=np.int32)
            check(lib().pgd_sort_edges_subset(self.h, ordering.encode(), int(small_to_large), ptr(ids, C.c_int32),
                                              len(ids)), self.ctx.h, "pgd_sort_edges_subset")
        self._ordered = True

    def set_e_ordered(self, ids) -> None:
        """Adopt a caller's E_ordered (edge ids in order, at most m)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        check(lib().pgd_set_e_ordered(self.h, ptr(ids, C.c_int32), len(ids)), self.ctx.h, "pgd_set_e_ordered")
        self._ordered = True

    # -- state getters ----------------------------------------------------
    def csr(self):
        row = np.zeros(self.n + 1, dtype=np.int64)
        col = np.zeros(max(2 * self.m, 1), dtype=np.int32)
        check(lib().pgd_graph_csr(self.h, ptr(row, C.c_int64), ptr(col, C.c_int32)), self.ctx.h)
        return row, col[: 2 * self.m]

    def edge_arrays(self):
        ev = np.zeros(max(self.m, 1), dtype=np.int32)
        eu = np.zeros(max(self.m, 1), dtype=np.int32)
        check(lib().pgd_edge_arrays(self.h, ptr(ev, C.c_int32), ptr(eu, C.c_int32)), self.ctx.h)
        return ev[: self.m], eu[: self.m]

    def e_ordered(self):
        k = C.c_int64()
        check(lib().pgd_e_ordered_count(self.h, C.byref(k)), self.ctx.h)
        ids = np.zeros(max(k.value, 1), dtype=np.int32)
        check(lib().pgd_e_ordered(self.h, ptr(ids, C.c_int32)), self.ctx.h)
        return ids[: k.value]

    # -- decomposition ----------------------------------------------------
    def decompose(self, micro: bool = False, engine: str | None = None) -> list[int]:
        """Device counting + reduction; returns the 15 sums."""
        if not self._ordered:
            self.sort_edges()
        s = np.zeros(NSUMS, dtype=np.uint64)
        flags = (MICRO if micro else 0) | _ENGINE[engine]
        check(lib().pgd_decompose(self.h, flags, ptr(s, C.c_uint64)), self.ctx.h, "pgd_decompose")
        self.sums = [int(x) for x in s]
        return self.sums

    def graphlet_decomposition(self, engine: str | None = None) -> dict[str, int]:
        """graphlet_decomposition (graphlet_core.cpp:1366-1474)."""
        self.decompose(False, engine)
        self.totals = dict(zip(TOTAL_NAMES, derive(self.n, self.m, self.sums)))
        return self.totals

    def graphlet_decomposition_micro(self, engine: str | None = None) -> dict[str, int]:
        """graphlet_decomposition_micro (graphlet_core.cpp:1483-1625)."""
        self.decompose(True, engine)
        self.totals = dict(zip(TOTAL_NAMES, derive(self.n, self.m, self.sums)))
        return self.totals

    def micro(self, column: int | str) -> np.ndarray:
        if isinstance(column, str):
            column = MICRO_COLUMNS.index(column)
        out = np.zeros(max(self.m, 1), dtype=np.uint64)
        check(lib().pgd_micro(self.h, column, ptr(out, C.c_uint64)), self.ctx.h, "pgd_micro")
        return out[: self.m]

    def edge_counts(self):
        """(T, K4, C4) per edge id."""
        t = np.zeros(max(self.m, 1), dtype=np.uint64)
        k = np.zeros(max(self.m, 1), dtype=np.uint64)
        c = np.zeros(max(self.m, 1), dtype=np.uint64)
        check(lib().pgd_edge_counts(self.h, ptr(t, C.c_uint64), ptr(k, C.c_uint64), ptr(c, C.c_uint64)),
              self.ctx.h)
        return t[: self.m], k[: self.m], c[: self.m]

    def set_tuning(self, tuning: Tuning | None = None, **overrides) -> None:
        """Launch parameters of later decompositions (pgd_graph_set_tuning)."""
        t = tuning if tuning is not None else default_tuning()
        for k, v in overrides.items():
            setattr(t, k, int(v))
        check(lib().pgd_graph_set_tuning(self.h, C.byref(t)), self.ctx.h, "pgd_graph_set_tuning")

    def tuning(self) -> Tuning:
        t = Tuning()
        check(lib().pgd_graph_get_tuning(self.h, C.byref(t)), self.ctx.h)
        return t

    def last_engine(self) -> str:
        return {2: "orient", 3: "dense"}.get(lib().pgd_last_engine(self.h), "none")

    def last_phases(self) -> list[dict]:
        """Per-pass device time and work model of the last decomposition."""
        out = []
        buf = C.create_string_buffer(128)
        ms, by, un = C.c_double(), C.c_double(), C.c_double()
        i = 0
        while lib().pgd_last_phase(self.h, i, buf, 128, C.byref(ms), C.byref(by), C.byref(un)) == 0:
            out.append({"name": buf.value.decode(), "ms": ms.value, "bytes": by.value, "units": un.value})
            i += 1
        return out

    def last_timing(self):
        a, b, k = C.c_double(), C.c_double(), C.c_int32()
        lib().pgd_last_timing(self.h, C.byref(a), C.byref(b), C.byref(k))
        return a.value, b.value, k.value

    def write_micro_fd(self, fd: int, output_id: bool = True, prefix: str = "% ") -> int:
        """The micro table as text, formatted on the device, to a file
        descriptor (pgd_write_micro); returns the bytes written."""
        n = C.c_int64()
        check(lib().pgd_write_micro(self.h, fd, int(output_id), prefix.encode(), C.byref(n)), self.ctx.h,
              "pgd_write_micro")
        return n.value

    def write_micro_stats(self, filename: str, output_id: bool = True) -> int:
        """write_micro_stats (graphlet_core.cpp:1668-1761), device formatted."""
        import os
        fd = os.open(filename, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        try:
            return self.write_micro_fd(fd, output_id, "% ")
        finally:
            os.close(fd)

    # -- output formats (graphlet_core.cpp:2061-2084, 2138-2162, 1668-1761) --
    def macro_text(self) -> str:
        return "".join(f"{k} = {v}\n" for k, v in self.totals.items())

    def micro_text(self) -> str:
        ev, eu = self.edge_arrays()
        order = self.e_ordered()
        k = len(order)  # one row per E_ordered position
        cols = [self.micro(c)[:k] for c in ("tri", "2_star", "local_4_clique", "local_4_chordal_cycle",
                                            "local_4_tailed_tris", "local_4_cycle", "local_3_star", "local_4_path")]
        v1 = ev[order].astype(np.uint64) + np.uint64(1)
        u1 = eu[order].astype(np.uint64) + np.uint64(1)
        head = "% src,dst,triangle,2-star,4-clique,4-chordal-cycle,4-tailed-triangle,4-cycle,3-star,4-path\n"
        rows = np.stack([v1, u1] + [c.astype(np.uint64) for c in cols], axis=1) if k else None
        if rows is None:
            return head
        return head + "\n".join(",".join(str(int(x)) for x in r) for r in rows) + "\n"


# ---------------------------------------------------------------------------
# batched form
# ---------------------------------------------------------------------------
class Batch:
    """Many independent graphs counted together (pgd_batch_*).

    graphs: sequence of (n, src, dst) with ids local to each graph.  Each
    graph gets the plain path's preprocessing (deg/deg defaults) and E_ordered;
    run() replays one captured launch sequence for all of them."""

    def __init__(self, graphs, ctx: Context | None = None, micro: bool = True, dedup: bool = False):
        self.ctx = ctx or default_context()
        ng = len(graphs)
        n = np.ascontiguousarray([g[0] for g in graphs], dtype=np.int32)
        lens = np.array([len(g[1]) for g in graphs], dtype=np.int64)
        line_off = np.zeros(ng + 1, dtype=np.int64)
        np.cumsum(lens, out=line_off[1:])
        L = int(line_off[-1])
        src = np.empty(max(L, 1), dtype=np.int32)
        dst = np.empty(max(L, 1), dtype=np.int32)
        for i, g in enumerate(graphs):
            src[line_off[i]:line_off[i + 1]] = g[1]
            dst[line_off[i]:line_off[i + 1]] = g[2]
        h = C.c_void_p()
        flags = (MICRO if micro else 0) | (DEDUP if dedup else 0)
        check(lib().pgd_batch_create(self.ctx.h, ng, ptr(n, C.c_int32), ptr(line_off, C.c_int64),
                                     ptr(src, C.c_int32), ptr(dst, C.c_int32), flags, C.byref(h)),
              self.ctx.h, "pgd_batch_create")
        self.h = h
        self.ctx._graphs.add(self)
        self.micro_enabled = micro
        self.ngraphs, self.n_total, self.m_total = self.info()
        self.n = n.astype(np.int64)
        self.m = None

    def info(self):
        g, nv, ne = C.c_int32(), C.c_int64(), C.c_int64()
        check(lib().pgd_batch_info(self.h, C.byref(g), C.byref(nv), C.byref(ne)), self.ctx.h)
        return g.value, nv.value, ne.value

    def run(self):
        """Replay the counting; returns (sums[ng,15], totals[ng,17]) as uint64."""
        s = np.zeros((self.ngraphs, NSUMS), dtype=np.uint64)
        t = np.zeros((self.ngraphs, NTOTALS), dtype=np.uint64)
        check(lib().pgd_batch_run(self.h, ptr(s, C.c_uint64), ptr(t, C.c_uint64)), self.ctx.h, "pgd_batch_run")
        return s, t

    def last_ms(self) -> float:
        v = C.c_double()
        check(lib().pgd_batch_last_ms(self.h, C.byref(v)), self.ctx.h)
        return v.value

    def micro(self, gi: int, column: int | str, m: int) -> np.ndarray:
        """Micro column of graph gi (m = its edge count), E_ordered order."""
        if isinstance(column, str):
            column = MICRO_COLUMNS.index(column)
        out = np.zeros(max(m, 1), dtype=np.uint64)
        check(lib().pgd_batch_micro(self.h, gi, column, ptr(out, C.c_uint64)), self.ctx.h, "pgd_batch_micro")
        return out[:m]

    def set_tuning(self, tuning: Tuning | None = None, **overrides) -> None:
        t = tuning if tuning is not None else self.tuning()
        for k, v in overrides.items():
            setattr(t, k, int(v))
        check(lib().pgd_batch_set_tuning(self.h, C.byref(t)), self.ctx.h, "pgd_batch_set_tuning")

    def tuning(self) -> Tuning:
        t = Tuning()
        check(lib().pgd_batch_get_tuning(self.h, C.byref(t)), self.ctx.h)
        return t

    def autotune(self, trials: int = 3):
        """Built-in coordinate search; returns (best Tuning, its median ms)."""
        t, ms = Tuning(), C.c_double()
        check(lib().pgd_batch_autotune(self.h, trials, C.byref(t), C.byref(ms)), self.ctx.h, "pgd_batch_autotune")
        return t, ms.value

    def free(self) -> None:
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().pgd_batch_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# served form
# ---------------------------------------------------------------------------
def server_footprint(n: int, nlines: int) -> int:
    """Estimated HBM bytes the server charges a request (pgd_server_footprint)."""
    return int(lib().pgd_server_footprint(n, nlines))


class Server:
    """Request loop with continuous batching (pgd_server_*).

    submit() queues a graph and returns at once; a worker thread admits
    queued requests by estimated HBM footprint, counts them in batches on two
    overlapping streams, and keeps each result until release()."""

    def __init__(self, ctx: Context | None = None, hbm_budget: int = 32 << 30, max_inflight: int = 16384):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(lib().pgd_server_create(self.ctx.h, hbm_budget, max_inflight, C.byref(h)), self.ctx.h,
              "pgd_server_create")
        self.h = h
        self.ctx._graphs.add(self)

    def submit(self, n: int, src, dst, micro: bool = False, dedup: bool = False) -> int:
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        rid = C.c_int64()
        flags = (MICRO if micro else 0) | (DEDUP if dedup else 0)
        check(lib().pgd_server_submit(self.h, n, len(src), ptr(src, C.c_int32), ptr(dst, C.c_int32), flags,
                                      C.byref(rid)), self.ctx.h, "pgd_server_submit")
        return rid.value

    def poll(self, cap: int = 65536) -> list[int]:
        buf = np.zeros(cap, dtype=np.int64)
        k = C.c_int32()
        check(lib().pgd_server_poll(self.h, ptr(buf, C.c_int64), cap, C.byref(k)), self.ctx.h)
        return [int(x) for x in buf[: k.value]]

    def drain(self, timeout_ms: int = -1) -> None:
        check(lib().pgd_server_drain(self.h, timeout_ms), self.ctx.h, "pgd_server_drain")

    def result(self, rid: int) -> dict:
        s = np.zeros(NSUMS, dtype=np.uint64)
        t = np.zeros(NTOTALS, dtype=np.uint64)
        m, lat = C.c_int64(), C.c_double()
        check(lib().pgd_server_result(self.h, rid, ptr(s, C.c_uint64), ptr(t, C.c_uint64), C.byref(m),
                                      C.byref(lat)), self.ctx.h, "pgd_server_result")
        return {"sums": [int(x) for x in s], "totals": [int(x) for x in t], "m": m.value, "latency_ms": lat.value}

    def micro(self, rid: int, column: int | str, m: int) -> np.ndarray:
        if isinstance(column, str):
            column = MICRO_COLUMNS.index(column)
        out = np.zeros(max(m, 1), dtype=np.uint64)
        check(lib().pgd_server_micro(self.h, rid, column, ptr(out, C.c_uint64)), self.ctx.h, "pgd_server_micro")
        return out[:m]

    def release(self, rid: int) -> None:
        check(lib().pgd_server_release(self.h, rid), self.ctx.h, "pgd_server_release")

    def stats(self) -> dict:
        f, e, b = C.c_int64(), C.c_int64(), C.c_int64()
        el, p50, p99 = C.c_double(), C.c_double(), C.c_double()
        check(lib().pgd_server_stats(self.h, C.byref(f), C.byref(e), C.byref(el), C.byref(p50), C.byref(p99),
                                     C.byref(b)), self.ctx.h)
        return {"finished": f.value, "edges": e.value, "elapsed_s": el.value, "p50_ms": p50.value,
                "p99_ms": p99.value, "batches": b.value}

    def free(self) -> None:
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().pgd_server_free(self.h)
        self.h = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
