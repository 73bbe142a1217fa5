"""Write the `pgd` driver fixtures: for each driver case, the SHA-256 of the
.macro file, the .micro file and the stdout (without its timing line) that
the reference `pgd` produces.  The reference binary is compiled from the
reference sources by oracle/Makefile (`make -C oracle ref`) in this container
only; the GPU tests compare the drop-in driver with these hashes.

Usage: python tests/golden/make_driver_goldens.py
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.driver_cases import CASES, TIMING_PREFIX, case_id  # noqa: E402
from tests.oracle_lib import REF_BIN  # noqa: E402
from tests.test_oracle import GOLDENS, golden_graph_path  # noqa: E402


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def main() -> None:
    if not os.path.exists(REF_BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = {}
    for name, flags in CASES:
        tmp = tempfile.mkdtemp(prefix="pgd_drvg_")
        try:
            path = golden_graph_path(name, GOLDENS[name], tmp)
            local = os.path.basename(path)
            r = subprocess.run([REF_BIN, "-f", local, "-c", "x.macro", "-m", "x.micro", "-w", "4", *flags],
                               capture_output=True, cwd=tmp, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"reference failed on {name} {flags}: {r.stderr[-500:]!r}")
            lines = [ln for ln in r.stdout.split(b"\n") if not ln.startswith(TIMING_PREFIX.encode())]
            ent = {"stdout_sha256": sha(b"\n".join(lines))}
            for f in ("x.macro", "x.micro"):
                with open(os.path.join(tmp, f), "rb") as fh:
                    ent[f.split(".")[1] + "_sha256"] = sha(fh.read())
            out[case_id(name, flags)] = ent
            print(case_id(name, flags), ent["stdout_sha256"][:16], flush=True)
        finally:
            shutil.rmtree(tmp)
    with open(os.path.join(HERE, "driver.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
