"""The benchmark configuration itself on the GPU: tiled LU and Cholesky of
N = 24000 in 24 x 24 tiles of 1000, leading dimension 24000, in the level-
batched graph (--tiled) and as one persistent-scheduler launch (--subgraph).

Oracle: the same tiled algorithm on the host (OpenMP task DAG over the host
BLAS, reference task order, lu_mustard.cu:343-535 / cholesky_mustard.cu:
221-400).  Checks, as the reference's --verify asks and tighter: every entry
of the factors within 1e-10 relative of the oracle's (floored at 1e-3 of the
largest entry), the scaled residual ||A - LU||_F / (||A||_F N eps) <= 10 on
the device, and info == 0.  LU runs on the reference generator (include/
gen.h); Cholesky on A = M M^T / N + N I.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, T = 24000, 24


def _mu():
    import mustard_amd as mu
    return mu


def _threads():
    return max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1))


def _worst_ratio(F, R, lower_only=False, rtol=1e-10, floor_frac=1e-3, chunk=1000):
    """max |F - R| / (rtol * max(|R|, floor_frac * max|R|)) over column chunks
    (keeps the host memory of the comparison small)."""
    n = R.shape[0]
    rmax = 0.0
    for j in range(0, n, chunk):
        rmax = max(rmax, float(np.max(np.abs(R[j:, j:j + chunk] if lower_only
                                             else R[:, j:j + chunk]))))
    floor = floor_frac * rmax + 1e-300
    worst = 0.0
    for j in range(0, n, chunk):
        r0 = j if lower_only else 0
        f, r = F[r0:, j:j + chunk], R[r0:, j:j + chunk]
        if lower_only:  # rows below the diagonal of each column of the chunk
            m = np.tril(np.ones(f.shape, dtype=bool))
            f, r = f[m], r[m]
        worst = max(worst, float(np.max(np.abs(f - r) / (rtol * np.maximum(np.abs(r), floor)))))
    return worst


@pytest.fixture(scope="module")
def lu_case(oracle):
    mu = _mu()
    A = mu.gen_ref_host(N)
    ref = A.copy(order="F")
    assert oracle.tiled_lu(ref, T, threads=_threads()) == 0
    return A, ref


@pytest.fixture(scope="module")
def chol_case(oracle, torch_gpu):
    mu = _mu()
    A = mu.to_host(mu.gen_spd(N, 420))
    ref = A.copy(order="F")
    assert oracle.tiled_cholesky(ref, T, threads=_threads()) == 0
    return A, ref


@pytest.mark.parametrize("mode", ["tiled", "subgraph"])
def test_lu_24000_t24(torch_gpu, lu_case, mode):
    mu = _mu()
    A, ref = lu_case
    dA = mu.to_device(A)
    info, st = mu.factorize("lu", dA, N, T, mode=mode)
    torch_gpu.cuda.synchronize()
    assert info == 0
    assert mu.scaled_residual(mu.to_device(A), dA, N) <= 10.0
    assert _worst_ratio(np.asfortranarray(mu.to_host(dA)), ref) <= 1.0


@pytest.mark.parametrize("mode", ["tiled", "subgraph"])
def test_cholesky_24000_t24(torch_gpu, chol_case, mode):
    mu = _mu()
    A, ref = chol_case
    dA = mu.to_device(A)
    info, st = mu.factorize("chol", dA, N, T, mode=mode)
    torch_gpu.cuda.synchronize()
    assert info == 0
    assert mu.scaled_residual(mu.to_device(A), dA, N, chol=True) <= 10.0
    F = np.asfortranarray(mu.to_host(dA))
    assert _worst_ratio(F, ref, lower_only=True) <= 1.0
    # the strict upper triangle is left as it was (rows above each chunk)
    for j in range(1000, N, 1000):
        assert np.array_equal(F[:j, j:j + 1000], A[:j, j:j + 1000])
