"""Host-side checks of the product library (CPU only, no device calls)."""
from __future__ import annotations

import ctypes as C
import json
import os
import re
import shutil
import tempfile

import numpy as np
import pytest

import pgd_amd
from pgd_amd import _lib
from tests.oracle_lib import OracleGraph
from tests.oracle_lib import derive as oracle_derive
from tests.test_oracle import GOLDENS, golden_graph_path

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols() -> list[str]:
    text = open(os.path.join(ROOT, "include", "pgd_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pgd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == syms


def test_status_strings():
    L = _lib.lib()
    for code in range(0, -8, -1):
        assert L.pgd_status_str(code)


@pytest.fixture(scope="module")
def tmp():
    d = tempfile.mkdtemp(prefix="pgd_host_")
    yield d
    shutil.rmtree(d)


def coo_to_csr(n, src, dst, dedup):
    """Adjacency lists in line order; dedup keeps each pair's first line."""
    lists = [[] for _ in range(n)]
    seen = set()
    for a, b in zip(src.tolist(), dst.tolist()):
        if a == b:
            continue
        if dedup:
            k = (min(a, b), max(a, b))
            if k in seen:
                continue
            seen.add(k)
        lists[a].append(b)
        lists[b].append(a)
    row = np.zeros(n + 1, dtype=np.int64)
    row[1:] = np.cumsum([len(x) for x in lists])
    col = np.array([y for x in lists for y in x], dtype=np.int32)
    return row, col


@pytest.mark.parametrize("name", sorted(k for k in GOLDENS if "@" not in k))
def test_reader_matches_oracle_reader(name, tmp):
    ent = GOLDENS[name]
    if ent.get("m", 0) > 200000:
        pytest.skip("large input")
    path = golden_graph_path(name, ent, tmp)
    n, src, dst, dedup, _ = pgd_amd.read_graph(path)
    row, col = coo_to_csr(n, src, dst, dedup)
    orow, ocol = OracleGraph.from_file(path).csr()
    assert np.array_equal(row, orow)
    assert np.array_equal(col, ocol)


def test_reader_rejects_missing_file():
    with pytest.raises(pgd_amd.PGDError):
        pgd_amd.read_graph("/nonexistent/graph.mtx")


def _rand_sums(rng, big: bool):
    if big:
        return [int(x) for x in rng.integers(0, 2**64, size=15, dtype=np.uint64)]
    scale = 2 ** int(rng.integers(1, 63))
    return [int(x) for x in rng.integers(0, scale, size=15, dtype=np.uint64)]


def test_derive_emulation_matches_binary64():
    """Integer derivation == the oracle's native binary64 derivation."""
    rng = np.random.default_rng(7)
    for i in range(20000):
        n = int(rng.choice([int(rng.integers(0, 5000)), int(rng.integers(0, 2**22)), int(rng.integers(0, 2**31))]))
        m = int(rng.integers(0, max(1, n * 20)))
        s = _rand_sums(rng, big=i % 3 == 0)
        assert pgd_amd.derive(n, m, s) == oracle_derive(n, m, s), (n, m, s)


def test_derive_rounding_edge_values():
    """Values around 2^53 and 2^64 where round-to-even and fl(1/3) matter."""
    vals = []
    for b in (52, 53, 54, 55, 60, 62, 63, 64):
        for d in (-3, -2, -1, 0, 1, 2, 3, 5, 7):
            vals.append(((1 << b) + d) % (1 << 64))
    vals += [3 * (1 << 53) + 1, 3 * (1 << 53) + 2, 6 * (1 << 55) - 1, (1 << 64) - 1, (1 << 64) - 2048]
    for a in vals:
        for n in (5, 2**21 + 7, 2**24, 2**31 - 1):
            s = [a] * 15
            assert pgd_amd.derive(n, 1000, s) == oracle_derive(n, 1000, s)
            s2 = [a if k % 2 else (a >> 3) for k in range(15)]
            assert pgd_amd.derive(n, 3, s2) == oracle_derive(n, 3, s2)


def test_generators_deterministic():
    a = pgd_amd.gen_gnp(500, 0.05, 9)
    b = pgd_amd.gen_gnp(500, 0.05, 9)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    s, d = pgd_amd.gen_rmat(12, 16, seed=4)
    assert np.all(s < d)
    key = s.astype(np.int64) * (1 << 32) + d
    assert np.all(np.diff(key) > 0)  # sorted, no duplicates
    s, d = pgd_amd.gen_ba(2000, 5, 3)
    assert np.all(s != d)
    assert len(set(zip(np.minimum(s, d).tolist(), np.maximum(s, d).tolist()))) == len(s)
    m = len(pgd_amd.gen_gnp(2000, 0.01, 1)[0])
    assert 18000 < m < 22000


def test_tuning_defaults():
    """pgd_tuning_default (host only) returns the documented defaults."""
    t = pgd_amd.default_tuning().as_dict()
    assert t == {"small_k": 64, "mid_k": 256, "grid_small": 64, "grid_mid": 32, "wedge_log2r": 15,
                 "wedge_nt": 512, "wedge_split": 4, "tri_lists": 64, "tri_search": 4}
    with pytest.raises(KeyError):
        pgd_amd.default_tuning(no_such_field=1)


def test_server_footprint_formula():
    """Admission estimate: CSR 8m + 8(n+1), per-edge scratch and outputs 64m, 4 KiB."""
    for n, m in ((0, 0), (20, 50), (5000, 80000), (262144, 3805169)):
        assert pgd_amd.server_footprint(n, m) == 8 * m + 8 * (n + 1) + 64 * m + 4096
    assert pgd_amd.server_footprint(-1, 5) < 0


def test_batched_workload_is_seeded():
    """The batched workload: same seed, same graphs; the size law and families."""
    from pgd_amd.workloads import batched_graph, batched_graphs
    a = batched_graphs(24, seed=3)
    b = batched_graphs(24, seed=3)
    for (n1, s1, d1), (n2, s2, d2) in zip(a, b):
        assert n1 == n2 and np.array_equal(s1, s2) and np.array_equal(d1, d2)
    fams = [batched_graph(i, seed=3)[3] for i in range(24)]
    assert fams[::2] == ["ba"] * 12 and fams[1::2] == ["gnp"] * 12
    ns = np.array([g[0] for g in batched_graphs(400, seed=5)])
    assert ns.min() >= 20 and ns.max() <= 5000
    # log-uniform on [20, 5000]: the median is near sqrt(20 * 5000) = 316
    assert 200 < np.median(ns) < 500
    for n, s, d in a:
        assert np.all(s != d) and (len(s) == 0 or (s.max() < n and d.max() < n))


REF_ARITH = os.path.join(ROOT, "oracle", "_ref", "ref_arith")


def _ref_arith(lines: list[str]) -> list[list[int]]:
    import subprocess
    r = subprocess.run([REF_ARITH], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=120,
                       check=True)
    return [[int(x) for x in ln.split()] for ln in r.stdout.splitlines()]


def _edge_sums(rng, n_inputs: int):
    vals = []
    for b in (52, 53, 54, 55, 60, 62, 63, 64):
        for d in (-3, -2, -1, 0, 1, 2, 3, 5, 7):
            vals.append(((1 << b) + d) % (1 << 64))
    vals += [3 * (1 << 53) + 1, 3 * (1 << 53) + 2, 6 * (1 << 55) - 1, (1 << 64) - 1, (1 << 64) - 2048]
    cases = []
    for a in vals:
        for n in (5, 2**21 + 7, 2**24, 2**31 - 1):
            cases.append((n, 1000, [a] * 15))
            cases.append((n, 3, [a if k % 2 else (a >> 3) for k in range(15)]))
    for i in range(n_inputs):
        n = int(rng.choice([int(rng.integers(0, 5000)), int(rng.integers(0, 2**22)), int(rng.integers(0, 2**31))]))
        cases.append((n, int(rng.integers(0, max(1, n * 20))), _rand_sums(rng, big=i % 3 == 0)))
    return cases


@pytest.mark.skipif(not os.path.exists(REF_ARITH), reason="needs the reference compiler harness (build container)")
def test_derive_matches_reference_build_flags():
    """pgd_derive == the reference's own derivation expressions compiled with
    the reference's -ffast-math flags (oracle/ref_arith.cpp), including the
    values above 2^53 where x/3.0 and x*fl(1/3) could differ."""
    cases = _edge_sums(np.random.default_rng(11), 5000)
    ref = _ref_arith([f"D {n} {m} " + " ".join(str(x) for x in s) for n, m, s in cases])
    for (n, m, s), want in zip(cases, ref):
        assert pgd_amd.derive(n, m, s) == want, (n, m, s)


def _rne53(p: int) -> int:
    """Nearest binary64 value to p (ties to even), as an integer."""
    if p < 1 << 53:
        return p
    sh = p.bit_length() - 53
    q, rem = p >> sh, p & ((1 << sh) - 1)
    hf = 1 << (sh - 1)
    if rem > hf or (rem == hf and q & 1):
        q += 1
    return q << sh


def _micro_emulation(T: int, K: int, C: int, dv: int, du: int):
    """The device epilogue's integer steps for the chordal, 3-star and 4-path
    columns (pgd_epilogue.hpp), restated for this check."""
    M = (1 << 64) - 1
    sv, su = dv - T - 1, du - T - 1
    chordal = _rne53((_rne53(T * (T - 1)) >> 1) - _rne53(K))
    star3 = _rne53((_rne53(sv * (sv - 1)) >> 1) + (_rne53(su * (su - 1)) >> 1))
    path4 = (sv * su - C) & M
    return [chordal, star3, path4]


@pytest.mark.skipif(not os.path.exists(REF_ARITH), reason="needs the reference compiler harness (build container)")
def test_micro_emulation_matches_reference_build_flags():
    """The epilogue's emulation of the micro columns' binary64 steps ==
    graphlet_core.cpp:1566-1575 compiled with the reference's flags, on
    per-edge values up to and past 2^53 (T, Sv, Su up to 2^31)."""
    rng = np.random.default_rng(5)
    cases = []
    for T in (0, 1, 2, 94906265, 94906266, 94906267, 94906268, 134217729, 2**30 + 3, 2**31 - 2):
        for extra in (0, 1, 5, 94906267, 2**30):
            dv = min(T + 1 + extra, 2**31 - 1)
            du = T + 1 + int(rng.integers(0, max(1, 2**31 - T - 2)))
            kmax = T * (T - 1) // 2
            for K in sorted({0, min(1, kmax), kmax // 3, max(kmax - 7, 0), kmax}):  # K4 <= C(T, 2)
                cases.append((T, K, int(rng.integers(0, 2**40)), dv, du))
    for _ in range(3000):
        T = int(rng.integers(0, 2**31 - 2))
        dv = T + 1 + int(rng.integers(0, max(1, 2**31 - T - 2)))
        du = T + 1 + int(rng.integers(0, max(1, 2**31 - T - 2)))
        K = int(rng.integers(0, T * (T - 1) // 2 + 1))
        cases.append((T, K, int(rng.integers(0, 2**50)), dv, du))
    ref = _ref_arith([f"M {T} {K} {C} {dv} {du}" for T, K, C, dv, du in cases])
    for c, want in zip(cases, ref):
        assert _micro_emulation(*c) == want, c


@pytest.mark.parametrize("chunk", ["1", "7", "64"])
def test_parallel_reader_with_tiny_chunks(chunk, tmp, monkeypatch):
    """The reader cut into chunks of a few bytes, so that carried values,
    comments and the first-line rules meet chunk boundaries: the same COO as
    one chunk, and as the oracle reader (pinned against the reference)."""
    names = [k for k in GOLDENS if "@" not in k and GOLDENS[k].get("m", 0) <= 20000 and "input" in GOLDENS[k]]
    assert any(k.startswith("rd_carry") for k in names)
    for name in names:
        path = golden_graph_path(name, GOLDENS[name], tmp)
        monkeypatch.delenv("PGD_READER_CHUNK", raising=False)
        one = pgd_amd.read_graph(path)
        monkeypatch.setenv("PGD_READER_CHUNK", chunk)
        many = pgd_amd.read_graph(path)
        assert one[0] == many[0] and one[3:] == many[3:], name
        assert np.array_equal(one[1], many[1]) and np.array_equal(one[2], many[2]), name
        n, src, dst, dedup, _ = many
        row, col = coo_to_csr(n, src, dst, dedup)
        orow, ocol = OracleGraph.from_file(path).csr()
        assert np.array_equal(row, orow) and np.array_equal(col, ocol), name
