"""The persistent scheduler's host emulation under ThreadSanitizer.

tests/sched_tsan_main.cpp runs mu_sched.hpp's worker loop (the code the
gfx950 kernel runs) on one host thread per wave, four per workgroup, over
tiled LU and Cholesky DAGs: with stalled producers and with CU budgets.  It is
built here with g++ -fsanitize=thread; any data race TSan finds, or any item
lost, repeated or run early, fails the test.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mustard_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_scheduler_emulation_under_tsan(tmp_path):
    exe = str(tmp_path / "sched_tsan")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread",
                    "-Wno-tsan", "-I", CSRC, os.path.join(ROOT, "tests", "sched_tsan_main.cpp"),
                    os.path.join(CSRC, "mu_sched_host.cpp"), os.path.join(CSRC, "mu_dag.cpp"),
                    "-o", exe], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ThreadSanitizer" not in p.stderr, p.stderr
    lines = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(lines) == 6 and all(l.endswith(" ok") for l in lines), p.stdout
