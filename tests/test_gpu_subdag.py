"""insertSubgraph on the GPU: a tiled LU whose diagonal GETRF tasks carry a
hand-recorded nested sub-DAG (the reference's recordSubgraph +
insertSubgraph for B > MAX_TILE, lu_mustard.cu:133-254, 381-388;
mustard.h:214-248), run in the level-batched graph and in the persistent
scheduler, against the oracle's tiled LU; and lu_mustard --subgraph with
B > MAX_TILE, which records its diagonal tiles that way."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _record(mu, dA, n, t, subt, info, key):
    """Tiled LU tasks (lu_mustard.cu:343-535) with each diagonal GETRF task
    replaced by an inserted sub-DAG: the tiled LU of that tile in subt x subt
    sub-tiles, leading dimension n."""
    b = n // t
    base = dA.data_ptr()

    def blk(i, j):
        return ctypes.c_void_p(base + (i * b + j * b * n) * 8)

    d = mu.Dag()
    kids = []
    d.capture(key)
    for k in range(t):
        d.begin_task((k, k), [(k, k)])
        d.end_task()
        child = mu.Dag()
        child.record_lu(blk(k, k), b, subt, lda=n, info=info)
        d.insert_subdag(child)
        kids.append(child)
        for i in range(k + 1, t):
            d.begin_task((k, i), [(k, k), (k, i)])
            mu.call("mu_dtrsm", key, b"L", b"L", b"N", b"U", b, b, 1.0, blk(k, k), n, blk(k, i), n)
            d.end_task()
        for i in range(k + 1, t):
            d.begin_task((i, k), [(k, k), (i, k)])
            mu.call("mu_dtrsm", key, b"R", b"U", b"N", b"N", b, b, 1.0, blk(k, k), n, blk(i, k), n)
            d.end_task()
            for j in range(k + 1, t):
                d.begin_task((i, j), [(i, k), (k, j), (i, j)])
                mu.call("mu_dgemm", key, b"N", b"N", b, b, b, -1.0, blk(i, k), n, blk(k, j), n,
                        1.0, blk(i, j), n)
                d.end_task()
    d.end_capture(key)
    for c in kids:  # the parent holds copies of the inserted tasks
        c.close()
    return d


@pytest.mark.parametrize("mode", ["tiled", "subgraph"])
@pytest.mark.parametrize("n,t,subt", [(1200, 2, 3), (2000, 4, 2)])
def test_inserted_subdag_lu_matches_oracle(torch_gpu, oracle, mode, n, t, subt):
    import mustard_amd as mu
    print("insertSubgraph LU", mode, n, t, subt, flush=True)
    A = mu.gen_ref_host(n)
    ref = A.copy(order="F")
    assert oracle.tiled_lu(ref, t * subt) == 0  # the same block structure, t*subt tiles
    dA = mu.to_device(A)
    info = mu.info_cell(dA.device)
    d = _record(mu, dA, n, t, subt, mu.ptr(info), ctypes.c_void_p(0x5ab))
    ex = mu.Exec(d, mode)
    ex.launch()
    torch_gpu.cuda.synchronize()
    ex.check()
    assert int(info.item()) in (0, 2 ** 31 - 1)
    F = np.asfortranarray(mu.to_host(dA))
    assert mu.scaled_residual(mu.to_device(A), dA, n) <= 10.0
    scale = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    assert np.max(np.abs(F - ref) / scale) <= 1e-9


def test_cli_subgraph_nests_tiles_above_max_tile():
    """lu_mustard --subgraph with B = 4500 > MAX_TILE = 4000: each diagonal
    tile is recorded as a 2 x 2 sub-graph and inserted (the reference's
    "Create subgraph" line per tile), and the factors verify."""
    exe = os.path.join(ROOT, "mustard_amd", "bin", "lu_mustard")
    p = subprocess.run([exe, "-n=9000", "-t=2", "--subgraph", "--verify", "-r=1"],
                       capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    assert lines[0] == "SUBGRAPH with N=9000 (2 of 4500x4500 tiles)"
    assert lines.count("PE 0. Create subgraph with subN=2 (2 of 2250x2250 tiles)") == 2
    sr = [l for l in lines if l.startswith("scaled residual")]
    assert sr and sr[0].endswith("(pass)"), sr
    assert sum(1 for l in lines if re.match(r"device 0 \| \d+ run \| time \(s\): ", l)) == 1
