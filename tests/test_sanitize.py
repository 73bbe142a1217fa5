"""The library's host code (reader, derivation, generators, .mtx writer)
under AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer
(CPU only).  tests/sanitize/host_checks.cpp drives it on every reader test
file with several thread counts and tiny parse chunks (so the reader's
threads split even small files), and checks its answers as it goes."""
from __future__ import annotations

import gzip
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pgd_amd", "csrc")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")
CXX = shutil.which("g++")

VARIANTS = {
    # OpenMP's runtime is not instrumented for threads, so the ThreadSanitizer
    # build compiles the generators without it (serial loops)
    "asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fopenmp"],
    "tsan": ["-fsanitize=thread"],
}
SIGNS = ("ERROR: AddressSanitizer", "WARNING: ThreadSanitizer", "runtime error:", "ERROR: LeakSanitizer")


@pytest.fixture(scope="module")
def workdir():
    d = tempfile.mkdtemp(prefix="pgd_san_")
    yield d
    shutil.rmtree(d)


@pytest.fixture(scope="module")
def inputs(workdir):
    out = []
    for name in sorted(os.listdir(INPUTS)):
        src = os.path.join(INPUTS, name)
        if name.endswith(".gz"):
            dst = os.path.join(workdir, name[:-3])
            with gzip.open(src, "rb") as f, open(dst, "wb") as g:
                shutil.copyfileobj(f, g)
            out.append(dst)
        else:
            out.append(src)
    return out


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_host_code_clean_under_sanitizer(variant, workdir, inputs):
    exe = os.path.join(workdir, f"host_checks_{variant}")
    cmd = [CXX, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", *VARIANTS[variant],
           f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
           os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"),
           os.path.join(CSRC, "pgd_reader.cpp"), os.path.join(CSRC, "pgd_derive.cpp"),
           os.path.join(CSRC, "pgd_gen.cpp"), "-lpthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, TMPDIR=workdir, PGD_READER_CHUNK="61", TSAN_OPTIONS="halt_on_error=1",
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1")
    r = subprocess.run([exe, *inputs], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert not any(s in out for s in SIGNS), out[-4000:]
    assert "failures=0" in r.stdout
    # every input parsed, the derivation and the generators ran
    assert sum(line.startswith(workdir) or line.startswith(INPUTS) for line in r.stdout.splitlines()) == len(inputs)
    assert "derive runs=" in r.stdout and "generators " in r.stdout
