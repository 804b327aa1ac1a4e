"""The library's host-side planning code (minijclsquant_amd/csrc/plan.cpp:
tile-plan renumbering and validation, realization-batch chunk tables) built
with g++ under AddressSanitizer and UndefinedBehaviorSanitizer and run over
random, lattice-shaped and degenerate inputs (tests/sanitize/plan_check.cpp).
CPU only."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "minijclsquant_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_plan_code_under_asan_ubsan(tmp_path):
    exe = tmp_path / "plan_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", f"-I{CSRC}",
           str(ROOT / "tests" / "sanitize" / "plan_check.cpp"), str(CSRC / "plan.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert res.returncode == 0, res.stderr[-4000:]
    assert "plan checks passed" in res.stdout
