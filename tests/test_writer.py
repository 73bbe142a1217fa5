"""The device .micro writer (pgd_write_micro) against the reference's own
text: the SHA-256 of the .micro file of every golden case (written by the
reference binary in the build container), a sampled edge set, the table
without ids, and a table of several chunks against host formatting."""
from __future__ import annotations

import hashlib
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

import pgd_amd
from tests.test_gpu import micro_matrix, product_run
from tests.test_oracle import GOLD, GOLDENS, SUBSET, flags_to_args, golden_graph_path

pytestmark = pytest.mark.gpu

MULTIGRAPHS = {"rd_mtx_dups.mtx"}
HEAD = "triangle,2-star,4-clique,4-chordal-cycle,4-tailed-triangle,4-cycle,3-star,4-path\n"


@pytest.fixture(scope="module")
def tmp():
    d = tempfile.mkdtemp(prefix="pgd_wr_")
    yield d
    shutil.rmtree(d)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sha_file(path: str) -> str:
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def host_text(mat: np.ndarray, prefix: str, output_id: bool) -> bytes:
    head = prefix + ("src,dst," if output_id else "") + HEAD
    body = mat if output_id else mat[:, 2:]
    return (head + "".join(",".join(map(str, r)) + "\n" for r in body.tolist())).encode()


@pytest.mark.parametrize("name", sorted(k for k in GOLDENS if k.split("@")[0] not in MULTIGRAPHS))
def test_device_micro_file_matches_reference(name, ctx, tmp):
    ent = GOLDENS[name]
    say(f"[gpu] writer {name} m={ent.get('m')}")
    path = golden_graph_path(name, ent, tmp)
    g, _, _ = product_run(path, ent["flags"], ctx, None)
    out = os.path.join(tmp, "x.micro")
    n = g.write_micro_stats(out)
    assert n == os.path.getsize(out)
    assert sha_file(out) == ent["micro_sha256"]
    g.free()


def test_device_micro_sampled_and_without_ids(ctx, tmp):
    key = "DD21.mtx@es10_o_kcore_s2l"
    ent = SUBSET[key]
    say(f"[gpu] writer {key}")
    path = golden_graph_path(ent["graph"], GOLDENS[ent["graph"]], tmp)
    data = np.load(os.path.join(GOLD, ent["file"]))
    kw = flags_to_args(ent["flags"])
    g = pgd_amd.GraphletCore.from_file(path, ctx=ctx, neigh_ordering=kw["neigh"],
                                       neigh_small_to_large=kw["s2l_neigh"])
    g.sort_edges(kw["ordering"], kw["s2l"], E_s=data["E_s"])
    g.graphlet_decomposition_micro()
    for output_id, prefix in ((True, "% "), (False, ""), (True, "")):
        out = os.path.join(tmp, "s.micro")
        fd = os.open(out, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        try:
            g.write_micro_fd(fd, output_id, prefix)
        finally:
            os.close(fd)
        with open(out, "rb") as f:
            assert f.read() == host_text(data["micro"], prefix, output_id), (output_id, prefix)
    g.free()


def test_device_micro_several_chunks(ctx, tmp):
    """A table of more than one chunk (2^20 rows) against host formatting."""
    s, d = pgd_amd.gen_rmat(17, 16, 0.57, 0.19, 0.19, 3)
    n = int(max(s.max(), d.max())) + 1
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
    g.sort_edges("deg")
    g.decompose(micro=True)
    say(f"[gpu] writer R-MAT-17 m={g.m}")
    assert g.m > 1 << 20
    out = os.path.join(tmp, "r.micro")
    g.write_micro_stats(out)
    with open(out, "rb") as f:
        assert f.read() == host_text(micro_matrix(g), "% ", True)
    g.free()


def test_device_micro_needs_micro_columns(ctx):
    s, d = pgd_amd.gen_gnp(100, 0.1, 3)
    g = pgd_amd.GraphletCore(s, d, n=100, ctx=ctx)
    g.sort_edges("deg")
    g.decompose(micro=False)
    with pytest.raises(pgd_amd.PGDError) as e:
        g.write_micro_fd(1)
    assert e.value.code == -6
    g.free()
