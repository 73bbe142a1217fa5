"""The `pgd` driver is a drop-in for the reference's: same flags, same files,
same stdout apart from the timing line.  Compared against fixtures written by
tests/golden/make_driver_goldens.py from the reference binary compiled from
its sources in the build container: SHA-256 of the .macro and .micro files
and of the stdout without the timing line.  No reference binary runs here."""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import tempfile

import pytest

from tests.driver_cases import CASES, TIMING_PREFIX, case_id
from tests.test_oracle import GOLDENS, golden_graph_path

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PGD = os.path.join(ROOT, "pgd_amd", "bin", "pgd")
with open(os.path.join(ROOT, "tests", "golden", "driver.json")) as _f:
    DRIVER = json.load(_f)


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("name,flags", CASES, ids=[case_id(n, f) for n, f in CASES])
def test_driver_matches_reference(name, flags):
    print(f"[gpu] driver {name} {flags}", flush=True)
    want = DRIVER[case_id(name, flags)]
    tmp = tempfile.mkdtemp(prefix="pgd_drv_")
    try:
        path = golden_graph_path(name, GOLDENS[name], tmp)
        r = subprocess.run([PGD, "-f", os.path.basename(path), "-c", "x.macro", "-m", "x.micro", *flags],
                           capture_output=True, cwd=tmp, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        lines = [ln for ln in r.stdout.split(b"\n") if not ln.startswith(TIMING_PREFIX.encode())]
        for f in ("x.macro", "x.micro"):
            with open(os.path.join(tmp, f), "rb") as fh:
                assert sha(fh.read()) == want[f.split(".")[1] + "_sha256"], f
        assert sha(b"\n".join(lines)) == want["stdout_sha256"], "stdout differs"
    finally:
        shutil.rmtree(tmp)


def test_driver_fixtures_cover_every_case():
    assert sorted(DRIVER) == sorted(case_id(n, f) for n, f in CASES)


def test_driver_rejects_missing_file():
    r = subprocess.run([PGD, "-f", "/nonexistent.mtx"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
