"""Host-code sanitizers on the CPU: the oracle (plain C, OpenMP) built with AddressSanitizer and
UndefinedBehaviorSanitizer and cross-checked against its brute-force definitions on random
inputs (tests/native/oracle_sanitize.c)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_sanitize"
    build = subprocess.run(
        ["gcc", "-O1", "-g", "-std=c11", "-fopenmp", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-o", str(exe),
         os.path.join(ROOT, "tests", "native", "oracle_sanitize.c"),
         os.path.join(ROOT, "oracle", "hm_oracle.c")],
        capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", OMP_NUM_THREADS="2")
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "agree" in run.stdout
