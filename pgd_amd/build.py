"""Build the MI355X library and the `pgd` driver in-tree.

Outputs:
  pgd_amd/lib/libpgd_hip.so   C-ABI library (HIP kernels for gfx950 + host code)
  pgd_amd/bin/pgd             drop-in driver with the reference's flags

Every call recompiles (the build check must exercise the compiler).  Object
files go to pgd_amd/_build/ (git-ignored).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")

ARCH = os.environ.get("PGD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", shutil.which("g++") or "g++")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", f"-I{INC}", f"-I{CSRC}",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", f"-I{INC}", f"-I{CSRC}", "-Wall", "-Wno-unused-function"]

HIP_SOURCES = [
    "pgd_core.hip",
    "pgd_engine_orient.hip",
    "pgd_engine_wedge.hip",
    "pgd_engine_dense.hip",
    "pgd_batch.hip",
    "pgd_writer.hip",
]
HOST_SOURCES = ["pgd_reader.cpp", "pgd_gen.cpp", "pgd_derive.cpp"]
DRIVER_SOURCES = ["graphlet_core.cpp", "pgd_main.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        _run([HIPCC, *HIP_FLAGS, "-c", "-o", obj, path])
    else:
        _run([CXX, *HOST_FLAGS, "-c", "-o", obj, path])
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    sources = HIP_SOURCES + HOST_SOURCES
    jobs = min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    lib = os.path.join(LIB, "libpgd_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-fopenmp"])
    os.replace(tmp, lib)
    drv = [os.path.join(CSRC, s) for s in DRIVER_SOURCES]
    if all(os.path.exists(p) for p in drv):
        exe = os.path.join(BIN, "pgd")
        _run([CXX, *HOST_FLAGS, "-o", exe + ".tmp", *drv, f"-L{LIB}", "-lpgd_hip",
              "-Wl,-rpath,$ORIGIN/../lib"])
        os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build()
