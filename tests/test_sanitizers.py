"""The host set-up (C++ under dg-on-cuda_amd/csrc/host) and the oracle
(oracle/dg_oracle.cpp) under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer: tests/sanitize/host_sanitize.cpp drives meshes
(structured, periodic, the reference's gmsh fixtures, from arrays), geometry,
reference operators N = 1..7, the mesh and plot-file writers, the C ABI's
error paths and every oracle function; any report fails the run. CPU only."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with the sanitizer runtimes")
def test_host_code_and_oracle_are_clean_under_asan_and_ubsan(tmp_path):
    pkg = os.path.join(ROOT, "dg-on-cuda_amd")
    subprocess.run(["make", "-C", pkg, "sanitize"], check=True, capture_output=True, timeout=600)
    exe = os.path.join(pkg, "_build", "san", "host_sanitize")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, GOLDEN, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 check failures" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
