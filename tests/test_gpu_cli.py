"""The lu_mustard / cholesky_mustard command lines on the GPU: the reference's
output lines (lu_mustard.cu:622, 656, 661; verify.h:68), the printed
dependencies (mustard.h:183-195) and argh's "-n 600" behaviour."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mustard_amd", "bin")
DEVICE_LINE = re.compile(r"^device 0 \| (\d+) run \| time \(s\): (\d+\.\d{4})$")


def run(prog, *args, cwd=None):
    p = subprocess.run([os.path.join(BIN, prog), *args], capture_output=True, text=True,
                       timeout=240, cwd=cwd)
    return p.returncode, p.stdout, p.stderr


@pytest.mark.parametrize("prog,name", [("lu_mustard", "lu"), ("cholesky_mustard", "chol")])
def test_config0_verify(prog, name, tmp_path):
    rc, out, err = run(prog, "-n=600", "-t=2", "--tiled", "--verify", "-r=2", "--dot",
                       cwd=str(tmp_path))
    assert rc == 0, err
    lines = out.splitlines()
    assert lines[0] == "TILED with N=600 (2 of 300x300 tiles)"
    assert lines[1] == "SM Limit per kernel = 20"
    assert lines[2] == "cuBLAS workspace = 256 kB"
    dev = [l for l in lines if l.startswith("device")]
    assert len(dev) == 2 and all(DEVICE_LINE.match(l) for l in dev)
    # scripts/calc.py:128-147 reads the last token of each device line
    assert all(float(l.split()[-1]) >= 0 for l in dev)
    e = [l for l in lines if l.startswith("error = ")]
    assert len(e) == 1 and float(e[0].split()[-1]) <= 1e-6
    assert "Result passes verification: 1" in lines
    assert any(l.startswith("Total time used (s): ") for l in lines)
    sr = [l for l in lines if l.startswith("scaled residual")]
    assert sr and "pass" in sr[0]
    dot = (tmp_path / "graph.dot").read_text()
    assert dot.startswith("digraph") and ("GETRF" in dot or "POTRF" in dot)


def test_single_kernel_mode_and_argh_space_form():
    # "-n 600" is a flag plus a positional word: N keeps its default 15 and,
    # without --tiled/--subgraph, T = 1 (lu_mustard.cu:694-696)
    rc, out, err = run("lu_mustard", "-n", "600", "--verify")
    assert rc == 0, err
    assert out.splitlines()[0] == "Single-kernel with N=15 (1 of 15x15 tiles)"
    assert "Result passes verification: 1" in out


@pytest.mark.parametrize("prog", ["lu_mustard", "cholesky_mustard"])
@pytest.mark.parametrize("sm", ["20", "100"])
def test_subgraph_mode_runs_and_verifies(prog, sm):
    """--subgraph: one persistent scheduler launch per run, the reference's
    output lines, and --verify passes."""
    rc, out, err = run(prog, "-n=600", "-t=2", "--subgraph", "--verify", "-r=2", f"--sm={sm}")
    assert rc == 0, err
    lines = out.splitlines()
    assert lines[0] == "SUBGRAPH with N=600 (2 of 300x300 tiles)"
    assert f"SM Limit per kernel = {sm}" in lines
    assert sum(1 for l in lines if re.match(r"device 0 \| \d+ run \| time \(s\): ", l)) == 2
    assert "Result passes verification: 1" in out


def test_verbose_prints_totalnodes(oracle):
    rc, out, err = run("lu_mustard", "-n=1200", "-t=4", "--tiled", "-v")
    assert rc == 0, err
    assert "totalNodes=30" in out
    assert re.search(r"lowered: \d+ units, \d+ items", out)


def test_zero_pivot_exits_nonzero():
    # B = N/T = 1 tiles of a matrix whose first pivot is nonzero: fine; the
    # reference generator never has a zero pivot, so only exit 0 is checked
    rc, out, err = run("cholesky_mustard", "-n=64", "-t=64", "--tiled", "--verify")
    assert rc == 0, err
    assert "Result passes verification: 1" in out
