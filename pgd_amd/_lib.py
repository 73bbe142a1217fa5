"""ctypes binding of the C ABI declared in include/pgd_hip.h.

The library is built in-tree (pgd_amd/lib/libpgd_hip.so).  Loading fails
loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpgd_hip.so")

NSUMS = 15
NTOTALS = 17
NMICRO = 8
MICRO_TRI, MICRO_CLIQUE, MICRO_CHORDAL, MICRO_TAILED, MICRO_CYCLE, MICRO_STAR3, MICRO_PATH, MICRO_STAR2 = range(8)
DEDUP = 1
MICRO = 1
ENGINE_ORIENT = 4
ENGINE_DENSE = 8

TOTAL_NAMES = [
    "total_2_1edge", "total_2_indep", "total_3_tris", "total_2_star", "total_3_1edge", "total_3_indep",
    "total_4_clique", "total_4_chordcycle", "total_4_tailed_tris", "total_4_cycle", "total_3_star",
    "total_4_path", "total_4_1edge", "total_4_2edge", "total_4_2star", "total_4_tri", "total_4_indep",
]

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = [
    "pgd_open", "pgd_close", "pgd_status_str", "pgd_set_mem_cap", "pgd_mem_in_use", "pgd_last_error",
    "pgd_atomic_rate", "pgd_graph_from_coo", "pgd_graph_from_csr", "pgd_graph_from_file", "pgd_graph_free", "pgd_graph_info", "pgd_graph_csr", "pgd_preprocess",
    "pgd_edge_arrays", "pgd_compute_cores", "pgd_sort_edges", "pgd_sort_edges_subset",
    "pgd_set_e_ordered", "pgd_e_ordered_count", "pgd_e_ordered", "pgd_decompose", "pgd_decompose_noalloc", "pgd_micro", "pgd_write_micro",
    "pgd_edge_counts", "pgd_last_engine", "pgd_last_timing", "pgd_last_phase", "pgd_derive", "pgd_batch_create", "pgd_batch_run",
    "pgd_batch_micro", "pgd_batch_info", "pgd_batch_free", "pgd_batch_set_tuning", "pgd_batch_get_tuning",
    "pgd_batch_autotune", "pgd_batch_last_ms", "pgd_tuning_default", "pgd_graph_set_tuning", "pgd_graph_get_tuning",
    "pgd_server_create", "pgd_server_submit", "pgd_server_poll", "pgd_server_result", "pgd_server_micro",
    "pgd_server_release", "pgd_server_free", "pgd_server_footprint", "pgd_server_drain", "pgd_server_stats", "pgd_gen_gnp", "pgd_gen_ba", "pgd_gen_rmat", "pgd_write_mtx",
    "pgd_read_graph", "pgd_free",
]


class Tuning(C.Structure):
    """pgd_tuning: launch parameters of the oriented engine."""
    _fields_ = [(f, C.c_int32) for f in (
        "small_k", "mid_k", "grid_small", "grid_mid", "wedge_log2r", "wedge_nt", "wedge_split", "tri_lists",
        "tri_search")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PGDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


_lib = None

P = C.c_void_p
I32, I64, U32, U64, D = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
PI32, PI64, PU64 = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or python pgd_amd/build.py")
    L = C.CDLL(LIB_PATH)
    sig = {
        "pgd_open": (I32, [I32, C.POINTER(P)]),
        "pgd_close": (I32, [P]),
        "pgd_status_str": (C.c_char_p, [I32]),
        "pgd_set_mem_cap": (I32, [P, I64]),
        "pgd_mem_in_use": (I32, [P, PI64]),
        "pgd_last_error": (C.c_char_p, [P]),
        "pgd_atomic_rate": (I32, [P, I64, I32, C.POINTER(D), C.POINTER(D)]),
        "pgd_graph_from_coo": (I32, [P, I64, I64, PI32, PI32, U32, C.POINTER(P)]),
        "pgd_graph_from_csr": (I32, [P, I64, PI64, PI32, C.POINTER(P)]),
        "pgd_graph_from_file": (I32, [P, C.c_char_p, I32, C.POINTER(P), PI64]),
        "pgd_graph_free": (I32, [P]),
        "pgd_graph_info": (I32, [P, PI64, PI64, PI32, PI32]),
        "pgd_graph_csr": (I32, [P, PI64, PI32]),
        "pgd_preprocess": (I32, [P, C.c_char_p, I32]),
        "pgd_edge_arrays": (I32, [P, PI32, PI32]),
        "pgd_compute_cores": (I32, [P, PI32, PI32]),
        "pgd_sort_edges": (I32, [P, C.c_char_p, I32]),
        "pgd_sort_edges_subset": (I32, [P, C.c_char_p, I32, PI32, I64]),
        "pgd_set_e_ordered": (I32, [P, PI32, I64]),
        "pgd_e_ordered_count": (I32, [P, PI64]),
        "pgd_e_ordered": (I32, [P, PI32]),
        "pgd_decompose": (I32, [P, U32, PU64]),
        "pgd_decompose_noalloc": (I32, [P, U32, C.c_void_p, C.c_void_p]),
        "pgd_micro": (I32, [P, I32, PU64]),
        "pgd_write_micro": (I32, [P, I32, I32, C.c_char_p, PI64]),
        "pgd_edge_counts": (I32, [P, PU64, PU64, PU64]),
        "pgd_last_engine": (I32, [P]),
        "pgd_last_timing": (I32, [P, C.POINTER(D), C.POINTER(D), PI32]),
        "pgd_last_phase": (I32, [P, I32, C.c_char_p, I32, C.POINTER(D), C.POINTER(D), C.POINTER(D)]),
        "pgd_derive": (I32, [I64, I64, PU64, PU64]),
        "pgd_batch_create": (I32, [P, I32, PI32, PI64, PI32, PI32, U32, C.POINTER(P)]),
        "pgd_batch_run": (I32, [P, PU64, PU64]),
        "pgd_batch_micro": (I32, [P, I32, I32, PU64]),
        "pgd_batch_info": (I32, [P, PI32, PI64, PI64]),
        "pgd_batch_free": (I32, [P]),
        "pgd_batch_set_tuning": (I32, [P, C.POINTER(Tuning)]),
        "pgd_batch_get_tuning": (I32, [P, C.POINTER(Tuning)]),
        "pgd_batch_autotune": (I32, [P, I32, C.POINTER(Tuning), C.POINTER(D)]),
        "pgd_batch_last_ms": (I32, [P, C.POINTER(D)]),
        "pgd_tuning_default": (I32, [C.POINTER(Tuning)]),
        "pgd_graph_set_tuning": (I32, [P, C.POINTER(Tuning)]),
        "pgd_graph_get_tuning": (I32, [P, C.POINTER(Tuning)]),
        "pgd_server_create": (I32, [P, I64, I32, C.POINTER(P)]),
        "pgd_server_submit": (I32, [P, I32, I64, PI32, PI32, U32, PI64]),
        "pgd_server_poll": (I32, [P, PI64, I32, PI32]),
        "pgd_server_result": (I32, [P, I64, PU64, PU64, PI64, C.POINTER(D)]),
        "pgd_server_footprint": (I64, [I32, I64]),
        "pgd_server_drain": (I32, [P, I64]),
        "pgd_server_stats": (I32, [P, PI64, PI64, C.POINTER(D), C.POINTER(D), C.POINTER(D), PI64]),
        "pgd_server_micro": (I32, [P, I64, I32, PU64]),
        "pgd_server_release": (I32, [P, I64]),
        "pgd_server_free": (I32, [P]),
        "pgd_gen_gnp": (I64, [I32, D, U64, PI32, PI32, I64]),
        "pgd_gen_ba": (I64, [I32, I32, U64, PI32, PI32, I64]),
        "pgd_gen_rmat": (I64, [I32, I32, D, D, D, U64, PI32, PI32, I64]),
        "pgd_write_mtx": (I32, [C.c_char_p, I32, I64, PI32, PI32]),
        "pgd_read_graph": (I32, [C.c_char_p, PI64, PI64, C.POINTER(PI32), C.POINTER(PI32), PI32, PI64]),
        "pgd_free": (None, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


def check(st: int, ctx=None, what: str = "") -> None:
    if st != 0:
        L = lib()
        msg = L.pgd_status_str(st).decode()
        if ctx is not None:
            detail = L.pgd_last_error(ctx)
            if detail:
                msg += ": " + detail.decode()
        raise PGDError(st, f"{what}: {msg}" if what else msg)
