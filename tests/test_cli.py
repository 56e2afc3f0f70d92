"""CLI argument handling of lu_mustard / cholesky_mustard (CPU: every case
here exits before the GPU is touched).  Messages and exit codes follow
include/cli.h:67-97 and lu_mustard.cu:686-689."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mustard_amd", "bin")


def run(prog, *args):
    p = subprocess.run([os.path.join(BIN, prog), *args], capture_output=True, text=True,
                       timeout=60)
    return p.returncode, p.stdout, p.stderr


@pytest.mark.parametrize("prog,name", [("lu_mustard", "LU"), ("cholesky_mustard", "Cholesky")])
def test_help(prog, name):
    rc, out, err = run(prog, "--help")
    assert rc == 1
    assert f"{name} decomposition" in err and "--subgraph" in err and "--verify" in err
    rc2, _, err2 = run(prog, "-h")
    assert rc2 == 1 and err2 == err


@pytest.mark.parametrize("args,msg", [
    (["-n=600", "-t=7"], "Error: N must be divisible by T! Got 'N=600 & T=7'"),
    (["-n=abc"], "Error: Must provide a valid N value! Got 'abc'"),
    (["--T=x"], "Error: Must provide a valid T value! Got 'x'"),
    (["--sm=0"], "Error: Must provide a valid SM Limit value! Got '0'"),
    (["--smLimit=300"], "Error: Must provide a valid SM Limit value! Got '300'"),
    (["--ws=0"], "Error: Must provide a valid workspace (in kBytes) value! Got '0'"),
    (["-r=0"], "Error: Must provide a valid number of runs! Got '0'"),
    # the reference echoes the value found under run/r/R only (cli.h:93)
    (["--runs=-2"], "Error: Must provide a valid number of runs! Got ''"),
    (["-r=-2"], "Error: Must provide a valid number of runs! Got '-2'"),
])
@pytest.mark.parametrize("prog", ["lu_mustard", "cholesky_mustard"])
def test_validation(prog, args, msg):
    rc, out, err = run(prog, *args)
    assert rc == 1
    assert msg in err
    assert "Usage:" in err
