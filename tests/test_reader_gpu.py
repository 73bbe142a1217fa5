"""pgd_graph_from_file (text parsed by host threads, renumbering and CSR on
the device) against the host route pgd_read_graph + pgd_graph_from_coo, on
every golden input, with the text cut into chunks of a few bytes so that
carried values and the first-line rules meet chunk boundaries.  The golden
tests (tests/test_gpu.py) run the same entry point end to end against the
reference's outputs."""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

import pgd_amd
from tests.test_oracle import GOLDENS, golden_graph_path

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tmp():
    d = tempfile.mkdtemp(prefix="pgd_rd_")
    yield d
    shutil.rmtree(d)


INPUTS = sorted(k for k, e in GOLDENS.items() if "@" not in k)


@pytest.mark.parametrize("chunk", [None, "5", "61"])
@pytest.mark.parametrize("name", INPUTS)
def test_from_file_matches_host_route(name, chunk, ctx, tmp, monkeypatch):
    print(f"[gpu] from_file {name} chunk={chunk}", flush=True)
    sys.stdout.flush()
    path = golden_graph_path(name, GOLDENS[name], tmp)
    if chunk:
        monkeypatch.setenv("PGD_READER_CHUNK", chunk)
    n, src, dst, dedup, loops = pgd_amd.read_graph(path)
    try:
        a = pgd_amd.GraphletCore(src, dst, n=n, dedup=dedup, ctx=ctx, preprocess=False)
    except pgd_amd.PGDError as e:
        with pytest.raises(pgd_amd.PGDError) as e2:
            pgd_amd.GraphletCore.from_file(path, ctx=ctx, preprocess=False, threads=3)
        assert e2.value.code == e.code
        return
    b = pgd_amd.GraphletCore.from_file(path, ctx=ctx, preprocess=False, threads=3)
    assert (a.n, a.m, a.max_degree) == (b.n, b.m, b.max_degree)
    assert b.self_loops == loops
    ra, ca = a.csr()
    rb, cb = b.csr()
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
    a.free()
    b.free()


def test_from_file_rejects_missing_file(ctx):
    with pytest.raises(pgd_amd.PGDError) as e:
        pgd_amd.GraphletCore.from_file("/nonexistent/graph.txt", ctx=ctx)
    assert e.value.code == -1
