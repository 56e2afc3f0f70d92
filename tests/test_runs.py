"""The benchmark runner's log layout and reader (mustard_amd/runs.py), pinned
on the reference's own result logs and on a stub CLI.

tests/golden/calc_logs/lu/24000/ holds the banner and `device` lines of four
of the reference's published logs (scripts/results/lu/24000/, the lines
scripts/calc.py reads); the means below are the reference's published A100
figures for N = 24000, T = 24 computed from those logs.  The stub stands in for lu_mustard on the CPU: it prints the CLI's
`device` lines, so the runner's file names and argument forms are checked
without a GPU (the GPU test in test_gpu_cli.py runs the real CLIs).
"""
import os
import stat
import sys

import pytest

from mustard_amd import runs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "calc_logs", "lu",
                    "24000")


def test_reads_the_reference_logs_like_calc_py():
    pts = runs.read_logs(GOLD, 24000)
    by = {}
    for name, g, secs, tf in pts:
        by.setdefault((name, g), []).append(secs)
    # one point per run; the 2-GPU log pairs device lines (the pair's last)
    assert {k: len(v) for k, v in by.items()} == {
        ("single-kernel", 1): 11, ("tiled", 1): 11, ("mustard", 1): 11, ("mustard", 2): 11}
    mean = {k: sum(v) / len(v) for k, v in by.items()}
    assert abs(mean["single-kernel", 1] - 0.5356) < 5e-5
    assert abs(mean["tiled", 1] - 0.5752) < 5e-5
    assert abs(mean["mustard", 1] - 0.5982) < 5e-5
    # the rate convention: 2N^3/3 flops for every method
    name, g, secs, tf = pts[0]
    assert abs(tf - 2 * 24000 ** 3 / 3 / secs * 1e-12) < 1e-9


def test_runner_writes_the_reference_layout(tmp_path):
    stub = tmp_path / "bin" / "lu_mustard"
    stub.parent.mkdir()
    stub.write_text(
        "#!" + sys.executable + "\n"
        "import sys\n"
        "a = dict((x.lstrip('-').split('=', 1) + [''])[:2] for x in sys.argv[1:])\n"
        "mode = 'SUBGRAPH' if 'subgraph' in a else 'TILED' if 'tiled' in a else 'Single-kernel'\n"
        "print(f\"{mode} with N={a['N']} ({a.get('T', 1)} tiles)\")\n"
        "for r in range(int(a['run'])):\n"
        "    print(f'device 0 | {r} run | time (s): {0.25 + 0.001 * r:4.4f}')\n"
        "print('Total time used (s): 1.0000')\n")
    stub.chmod(stub.stat().st_mode | stat.S_IEXEC)
    out = tmp_path / "logs"
    paths = runs.run("lu", 600, 2, 3, str(out), bin_dir=str(stub.parent))
    assert [os.path.basename(p) for p in paths] == \
        ["log0_600.log", "log1_600_2.log", "log2_600_2_1GPU.log"]
    text = open(paths[2]).read()
    assert text.startswith("SUBGRAPH with N=600 (2 tiles)")
    pts = runs.read_logs(str(out / "lu" / "600"), 600)
    assert sorted({(p[0], p[1]) for p in pts}) == [("mustard", 1), ("single-kernel", 1),
                                                   ("tiled", 1)]
    assert len(pts) == 9 and min(p[2] for p in pts) == 0.25
    # a second invocation appends, as the reference scripts do
    runs.run("lu", 600, 2, 1, str(out), methods=(2,), bin_dir=str(stub.parent))
    assert len(runs.read_logs(str(out / "lu" / "600"), 600)) == 10
