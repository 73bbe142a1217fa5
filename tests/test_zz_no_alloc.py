"""pgd_decompose_noalloc (the allocation-free entry point) through the C ABI,
checked bit for bit against pgd_decompose with the dense engine on the same
graph: two calls in a row, the second on a smaller graph (the scratch array
keeps the first call's contents), then a third without micro columns."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import pgd_amd
from pgd_amd._lib import ENGINE_DENSE, MICRO, NMICRO, NSUMS, lib, ptr

pytestmark = pytest.mark.gpu

ERR_INVALID, ERR_UNSUPPORTED = -1, -5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def reference_run(g, micro: bool):
    """The existing entry point: pgd_decompose with the dense engine."""
    out = np.zeros(NSUMS, dtype=np.uint64)
    st = lib().pgd_decompose(g.h, (MICRO if micro else 0) | ENGINE_DENSE, ptr(out, C.c_uint64))
    assert st == 0, st
    cols = np.stack([g.micro(c) for c in range(NMICRO)]) if micro else None
    return out, cols


class DeviceArray:
    """Test-side device memory from the HIP runtime the library links (the
    caller's buffers: the entry point under test allocates nothing)."""
    _hip = None

    def __init__(self, count: int, fill: int = 0xFF):
        if DeviceArray._hip is None:
            lib()  # the library's HIP runtime is loaded first
            DeviceArray._hip = C.CDLL("libamdhip64.so.7")
        self.nbytes = 8 * max(count, 1)
        self.p = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(self.p), C.c_size_t(self.nbytes)) == 0
        # stale contents the entry point must overwrite
        assert self._hip.hipMemset(self.p, C.c_int(fill), C.c_size_t(self.nbytes)) == 0
        assert self._hip.hipDeviceSynchronize() == 0

    def host(self, count: int) -> np.ndarray:
        out = np.empty(max(count, 1), dtype=np.uint64)
        assert self._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * max(count, 1)), 2) == 0
        return out[:count]

    def free(self):
        if self.p:
            self._hip.hipFree(self.p)
            self.p = None


def noalloc_run(g, micro: bool):
    sums = DeviceArray(NSUMS)
    mic = DeviceArray(NMICRO * g.m) if micro else None
    try:
        st = lib().pgd_decompose_noalloc(g.h, MICRO if micro else 0, sums.p, mic.p if micro else None)
        assert st == 0, (st, lib().pgd_last_error(g.ctx.h))
        out = sums.host(NSUMS)
        cols = mic.host(NMICRO * g.m).reshape(NMICRO, g.m) if micro else None
    finally:
        sums.free()
        if mic:
            mic.free()
    return out, cols


def make(kind: str, ctx):
    if kind == "gnp_1000_0.85":
        s, d = pgd_amd.gen_gnp(1000, 0.85, 21)
        n = 1000
    elif kind == "ba_600_3":
        s, d = pgd_amd.gen_ba(600, 3, 22)
        n = 600
    else:
        s, d = pgd_amd.gen_gnp(1100, 0.1, 23)
        n = 1100
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
    g.sort_edges("deg")
    return g


def test_noalloc_matches_dense_decompose(ctx):
    big = make("gnp_1000_0.85", ctx)
    small = make("ba_600_3", ctx)
    for g, micro, name in ((big, True, "gnp_1000_0.85"), (small, True, "ba_600_3"), (big, False, "gnp_1000_0.85")):
        say(f"[gpu] noalloc {name} m={g.m} micro={micro}")
        got_s, got_m = noalloc_run(g, micro)
        want_s, want_m = reference_run(g, micro)
        assert np.array_equal(got_s, want_s), name
        if micro:
            assert np.array_equal(got_m, want_m), name
    big.free()
    small.free()


def test_noalloc_rejects_beyond_its_limits(ctx):
    say("[gpu] noalloc limits")
    g = make("gnp_1100", ctx)
    sums = DeviceArray(NSUMS)
    try:
        assert lib().pgd_decompose_noalloc(g.h, 0, sums.p, None) == ERR_UNSUPPORTED
        assert lib().pgd_decompose_noalloc(g.h, MICRO, sums.p, None) == ERR_INVALID
        assert lib().pgd_decompose_noalloc(g.h, ENGINE_DENSE, sums.p, None) == ERR_INVALID
    finally:
        sums.free()
    g.free()
