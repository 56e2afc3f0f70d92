"""The persistent scheduler's host emulation under ThreadSanitizer, and the
host recorder / lowering / emulation under AddressSanitizer + UBSan.

tests/sched_tsan_main.cpp runs mu_sched.hpp's worker loop (the code the
gfx950 kernel runs) on one host thread per wave, four per workgroup, over
tiled LU and Cholesky DAGs: with stalled producers, with CU budgets and with
diagonal tiles as inserted sub-DAGs.  It is built here with g++
-fsanitize=thread, and again with -fsanitize=address,undefined; any data
race, memory error or undefined behaviour the sanitizers find, or any item
lost, repeated or run early, fails the test.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mustard_amd", "csrc")


def _build_and_run(tmp_path, name, flags, env):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", *flags, "-pthread", "-I", CSRC,
                    os.path.join(ROOT, "tests", "sched_tsan_main.cpp"),
                    os.path.join(CSRC, "mu_sched_host.cpp"), os.path.join(CSRC, "mu_dag.cpp"),
                    "-o", exe], check=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, **env))
    assert p.returncode == 0, p.stdout + p.stderr
    lines = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(lines) == 7 and all(l.endswith(" ok") for l in lines), p.stdout
    return p


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_scheduler_emulation_under_tsan(tmp_path):
    p = _build_and_run(tmp_path, "sched_tsan", ["-fsanitize=thread", "-Wno-tsan"],
                       {"TSAN_OPTIONS": "halt_on_error=1 exitcode=66"})
    assert "ThreadSanitizer" not in p.stderr, p.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_ubsan(tmp_path):
    p = _build_and_run(tmp_path, "sched_asan",
                       ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                        "-fno-omit-frame-pointer"],
                       {"ASAN_OPTIONS": "halt_on_error=1 exitcode=67 detect_leaks=1"})
    assert "AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr, p.stderr
