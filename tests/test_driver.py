"""The `pgd` driver is a drop-in for the reference's: same flags, same files,
same stdout apart from the timing line.  Compared against the reference
binary compiled from its sources (oracle/Makefile, target `ref`)."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import pytest

from tests.oracle_lib import REF_BIN
from tests.test_oracle import GOLDENS, golden_graph_path

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PGD = os.path.join(ROOT, "pgd_amd", "bin", "pgd")

CASES = [
    ("4-clique-and-edge.txt", []), ("4-chordal-cycle.txt", []), ("DD21.mtx", []), ("DD21.mtx", ["-o", "kcore"]),
    ("DD21.mtx", ["--s2l", "-n", "kcore"]), ("sample_graph.csv", []), ("rd_csv_header.csv", []),
    ("rd_weighted.txt", []), ("rd_triple_space.txt", []), ("rd_gap_ids.mtx", []), ("rd_tab.edges", []),
    ("C250-9.mtx", []), ("gnp_2000_0.01", ["--s2l_neigh"]), ("ba_3000_4", ["-o", "natural"]),
]


def run(exe, args, cwd):
    r = subprocess.run([exe, *args], capture_output=True, text=True, cwd=cwd, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return [ln for ln in r.stdout.splitlines() if not ln.startswith("graphlet decomposition time:")]


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference binary not built")
@pytest.mark.parametrize("name,flags", CASES, ids=[f"{n}{''.join(f)}" for n, f in CASES])
def test_driver_matches_reference(name, flags):
    print(f"[gpu] driver {name} {flags}", flush=True)
    tmp = tempfile.mkdtemp(prefix="pgd_drv_")
    try:
        path = golden_graph_path(name, GOLDENS[name], tmp)
        a, b = os.path.join(tmp, "ref"), os.path.join(tmp, "gpu")
        os.makedirs(a)
        os.makedirs(b)
        out_a = run(REF_BIN, ["-f", path, "-c", "x.macro", "-m", "x.micro", "-w", "4", *flags], a)
        out_b = run(PGD, ["-f", path, "-c", "x.macro", "-m", "x.micro", *flags], b)
        for f in ("x.macro", "x.micro"):
            with open(os.path.join(a, f), "rb") as fa, open(os.path.join(b, f), "rb") as fb:
                assert fa.read() == fb.read(), f
        assert out_a == out_b
    finally:
        shutil.rmtree(tmp)


def test_driver_rejects_missing_file():
    r = subprocess.run([PGD, "-f", "/nonexistent.mtx"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
