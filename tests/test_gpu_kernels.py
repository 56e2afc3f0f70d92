"""Parity of the gfx950 tile kernels against LAPACK semantics.

GEMM/SYRK are checked against numpy's float64 product (a dgemm), TRSM against
scipy's solve_triangular, and GETRF/POTRF against the oracle (LAPACK dpotrf,
and the no-pivot Doolittle restatement, itself pinned to dgetrf with identity
pivots in test_oracle.py).  Every call goes through the C ABI.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from tests.util import EPS, close_rel, diag_dominant, rand, spd

pytestmark = pytest.mark.gpu


def _mu():
    import mustard_amd as mu
    return mu


def _dev(A):
    return _mu().to_device(np.asfortranarray(A))


def _host(T, shape):
    return _mu().to_host(T)[: shape[0], : shape[1]]


GEMM_CASES = [
    # m, n, k, transa, transb, pad, alpha, beta
    (1, 1, 1, "N", "N", 0, -1.0, 1.0),
    (16, 16, 4, "N", "N", 0, -1.0, 1.0),
    (17, 33, 5, "N", "T", 0, -1.0, 1.0),
    (128, 128, 16, "N", "N", 0, 1.0, 0.0),
    (129, 130, 131, "T", "N", 0, 0.5, 2.0),
    (300, 257, 1000, "N", "N", 3, -1.0, 1.0),
    (257, 300, 999, "T", "T", 1, -1.0, 1.0),
    (1000, 1000, 1000, "N", "N", 24, -1.0, 1.0),   # LU trailing update shape
    (1000, 1000, 1000, "N", "T", 24, -1.0, 1.0),   # Cholesky trailing update shape
    (125, 125, 125, "N", "N", 875, -1.0, 1.0),     # nested sub-tile, odd offsets
    (64, 16, 1000, "N", "N", 0, -1.0, 1.0),
    # n <= 16 (substitution steps of a solve): the narrow-output path
    (1000, 16, 1000, "T", "N", 0, -1.0, 1.0),
    (1000, 16, 1000, "N", "N", 5, 1.0, 0.0),
    (100, 3, 37, "N", "N", 1, 0.5, 2.0),
    (65, 1, 1, "T", "N", 0, -1.0, 1.0),
    (130, 15, 67, "T", "N", 2, 1.0, 0.0),
]


@pytest.mark.parametrize("m,n,k,ta,tb,pad,alpha,beta", GEMM_CASES)
def test_dgemm(torch_gpu, m, n, k, ta, tb, pad, alpha, beta):
    mu = _mu()
    rng = np.random.default_rng(m * 7 + n * 13 + k)
    ar, ac = (k, m) if ta == "T" else (m, k)
    br, bc = (n, k) if tb == "T" else (k, n)
    lda, ldb, ldc = ar + pad, br + pad, m + pad
    A = rand(rng, lda, ac)
    B = rand(rng, ldb, bc)
    C = rand(rng, ldc, n)
    if beta == 0.0:
        C[:m, :n] = np.nan  # beta = 0 must not read C
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    mu.dgemm(ta, tb, m, n, k, alpha, mu.ptr(dA), lda, mu.ptr(dB), ldb, beta, mu.ptr(dC), ldc)
    torch_gpu.cuda.synchronize()
    opA = A[:ar, :ac].T if ta == "T" else A[:ar, :ac]
    opB = B[:br, :bc].T if tb == "T" else B[:br, :bc]
    ref = alpha * (opA @ opB) + (0.0 if beta == 0.0 else beta * C[:m, :n])
    got = mu.to_host(dC)
    # elementwise floating-point bound of a length-k dot product, for both sides
    tol = 2 * (k + 2) * EPS * (abs(alpha) * (np.abs(opA) @ np.abs(opB)) + (
        0.0 if beta == 0.0 else abs(beta) * np.abs(C[:m, :n]))) + 1e-300
    assert np.all(np.abs(got[:m, :n] - ref) <= tol)
    # nothing outside the m x n block may change
    if pad:
        assert np.array_equal(got[m:, :], C[m:, :])


LOW_CASES = [(1, 1, 1, "N", "N", 0), (17, 5, 3, "N", "T", 0), (64, 64, 16, "T", "N", 0),
             (129, 130, 131, "T", "T", 1), (300, 257, 1000, "N", "N", 3), (1000, 1000, 1000, "N", "T", 24)]


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("m,n,k,ta,tb,pad", LOW_CASES)
def test_dgemm_reduced_precision(torch_gpu, prec, m, n, k, ta, tb, pad):
    """FP32 / BF16 trailing-update GEMM: operands rounded, FP32 accumulation,
    FP64 C.  Bound per element: operand rounding (2u per product) plus an
    FP32 length-k sum, times (|A||B|)_ij."""
    mu = _mu()
    rng = np.random.default_rng(m * 5 + n * 3 + k)
    ar, ac = (k, m) if ta == "T" else (m, k)
    br, bc = (n, k) if tb == "T" else (k, n)
    lda, ldb, ldc = ar + pad, br + pad, m + pad
    A, B, C = rand(rng, lda, ac), rand(rng, ldb, bc), rand(rng, ldc, n)
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    mu.set_update_precision(prec)
    try:
        mu.dgemm(ta, tb, m, n, k, -1.0, mu.ptr(dA), lda, mu.ptr(dB), ldb, 1.0, mu.ptr(dC), ldc)
    finally:
        mu.set_update_precision("fp64")
    torch_gpu.cuda.synchronize()
    opA = A[:ar, :ac].T if ta == "T" else A[:ar, :ac]
    opB = B[:br, :bc].T if tb == "T" else B[:br, :bc]
    ref = C[:m, :n] - opA @ opB
    got = mu.to_host(dC)
    u_op = 2.0 ** -24 if prec == "fp32" else 2.0 ** -8
    tol = (2.5 * u_op + (k + 2) * 2.0 ** -24) * (np.abs(opA) @ np.abs(opB)) + 2 * EPS * np.abs(C[:m, :n])
    assert np.all(np.abs(got[:m, :n] - ref) <= tol)
    if prec == "fp32" and k >= 64:
        # and it really is reduced: not FP64 accurate
        assert np.max(np.abs(got[:m, :n] - ref)) > 1e3 * EPS
    if pad:
        assert np.array_equal(got[m:, :], C[m:, :])


A32_CASES = [
    # m, n, k, pad, offset of A (elements), alpha, beta
    (1, 1, 1, 0, 0, -1.0, 1.0),
    (64, 16, 32, 0, 0, 1.0, 0.0),
    (1000, 16, 2000, 0, 0, -1.0, 1.0),   # update of a two-tile solve step
    (4000, 16, 2000, 0, 0, 1.0, 0.0),    # two-tile diagonal-block inverse times R
    (8192, 16, 2000, 0, 0, -1.0, 1.0),   # 128 row blocks: k not split
    (130, 15, 67, 2, 0, 0.5, 2.0),       # ragged rows and k, lda % 4 != 0
    (101, 3, 37, 0, 1, -1.0, 1.0),       # m % 4 != 0, misaligned A
    (260, 16, 129, 4, 1, -1.0, 1.0),     # lda % 4 == 0 but A not 16-byte aligned
    (300, 16, 513, 0, 0, -1.0, 1.0),     # two k-slices, the second one k long
    (77, 7, 2049, 3, 1, 1.0, 0.0),       # five k-slices, scalar loads
]


@pytest.mark.parametrize("rows", [0, 1])
@pytest.mark.parametrize("m,n,k,pad,off,alpha,beta", A32_CASES)
def test_gemm_a32(torch_gpu, m, n, k, pad, off, alpha, beta, rows):
    """FP32-A narrow GEMM (the refinement solve's substitution step): FP32
    products of A and B rounded to FP32, summed in FP32 per k-slice of 512
    (the slices' sums added in FP64) and added into the FP64 C.  Bound per
    element: an FP32 length-k sum times (|A| |fl32(B)|)_ij, plus the FP64
    update of C.  B and C column-major, or stored by rows.  Run eagerly: the
    split's arrival counters start from a freshly zeroed workspace (graph
    launches that reuse them run in the refinement tests)."""
    mu = _mu()
    rng = np.random.default_rng(m * 11 + n * 3 + k + off)
    lda = m + pad
    A = rand(rng, lda, k).astype(np.float32)
    # B (k x n) and C (m x n) with `pad` spare rows (column-major) or spare
    # columns (stored by rows: the by-rows matrix is the transpose in memory)
    B = rand(rng, k + pad, n) if not rows else rand(rng, k, n + pad)
    C = rand(rng, m + pad, n) if not rows else rand(rng, m, n + pad)
    if beta == 0.0:
        C[:m, :n] = np.nan  # beta = 0 must not read C
    flat = np.zeros(off + lda * k, dtype=np.float32)
    flat[off:] = A.ravel(order="F")
    dA = torch_gpu.from_numpy(flat).cuda()
    if rows:
        ldb, ldc = n + pad, n + pad
        dB = torch_gpu.from_numpy(np.ascontiguousarray(B)).cuda()
        dC = torch_gpu.from_numpy(np.ascontiguousarray(C)).cuda()
    else:
        ldb, ldc = k + pad, m + pad
        dB, dC = _dev(B), _dev(C)
    mu.call("mu_gemm_a32", None, m, n, k, alpha, mu.ptr(dA, off), lda, mu.ptr(dB), ldb, beta,
            mu.ptr(dC), ldc, rows)
    torch_gpu.cuda.synchronize()
    opA = A[:m, :k].astype(np.float64)
    B32 = B[:k, :n].astype(np.float32).astype(np.float64)
    ref = alpha * (opA @ B32) + (0.0 if beta == 0.0 else beta * C[:m, :n])
    got = dC.cpu().numpy() if rows else mu.to_host(dC)
    tol = (k + 2) * 2.0 ** -24 * abs(alpha) * (np.abs(opA) @ np.abs(B32)) + 2 * EPS * (
        0.0 if beta == 0.0 else abs(beta) * np.abs(C[:m, :n])) + 1e-300
    assert np.all(np.abs(got[:m, :n] - ref) <= tol)
    if pad:  # nothing outside the m x n block may change
        if rows:
            assert np.array_equal(got[:, n:], C[:, n:])
        else:
            assert np.array_equal(got[m:, :], C[m:, :])


def test_gemm_a32_rejects_wide_outputs(torch_gpu):
    mu = _mu()
    with pytest.raises(Exception, match="n <= 16"):
        mu.call("mu_gemm_a32", None, 4, 17, 4, 1.0, None, 4, None, 4, 0.0, None, 4, 0)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_dsyrk_reduced_precision(torch_gpu, prec):
    mu = _mu()
    n, k = 300, 1000
    rng = np.random.default_rng(7)
    A, C = rand(rng, n, k), rand(rng, n, n)
    dA, dC = _dev(A), _dev(C)
    mu.set_update_precision(prec)
    try:
        mu.dsyrk("L", "N", n, k, -1.0, mu.ptr(dA), n, 1.0, mu.ptr(dC), n)
    finally:
        mu.set_update_precision("fp64")
    torch_gpu.cuda.synchronize()
    got = mu.to_host(dC)
    ref = C - A @ A.T
    u_op = 2.0 ** -24 if prec == "fp32" else 2.0 ** -8
    tol = (2.5 * u_op + (k + 2) * 2.0 ** -24) * (np.abs(A) @ np.abs(A.T)) + 2 * EPS * np.abs(C)
    lo = np.tril(np.ones((n, n), dtype=bool))
    assert np.all(np.abs(got - ref)[lo] <= tol[lo])
    assert np.array_equal(got[~lo], C[~lo])


@pytest.mark.parametrize("n,k,pad", [(1, 1, 0), (17, 3, 0), (128, 64, 0), (129, 129, 1),
                                     (300, 1000, 2), (1000, 1000, 24)])
def test_dsyrk_lower(torch_gpu, n, k, pad):
    mu = _mu()
    rng = np.random.default_rng(n + 31 * k)
    lda, ldc = n + pad, n + pad
    A = rand(rng, lda, k)
    C = rand(rng, ldc, n)
    dA, dC = _dev(A), _dev(C)
    mu.dsyrk("L", "N", n, k, -1.0, mu.ptr(dA), lda, 1.0, mu.ptr(dC), ldc)
    torch_gpu.cuda.synchronize()
    got = mu.to_host(dC)
    ref = C[:n, :n] - A[:n, :] @ A[:n, :].T
    lo = np.tril(np.ones((n, n), dtype=bool))
    tol = 2 * (k + 2) * EPS * (np.abs(A[:n, :]) @ np.abs(A[:n, :]).T + np.abs(C[:n, :n]))
    assert np.all(np.abs(got[:n, :n][lo] - ref[lo]) <= tol[lo])
    # strict upper triangle untouched (the reference's dsyrk writes LOWER only)
    assert np.array_equal(got[:n, :n][~lo], C[:n, :n][~lo])


def _tri(rng, n, lower, unit):
    T = rand(rng, n, n, -0.5, 0.5) / max(n, 1) ** 0.5
    T[np.arange(n), np.arange(n)] = 1.0 + rng.uniform(0.5, 1.5, n)
    mask = np.tril(np.ones((n, n), dtype=bool)) if lower else np.triu(np.ones((n, n), dtype=bool))
    T = np.where(mask, T, np.nan)  # the other triangle must never be read
    if unit:
        T[np.arange(n), np.arange(n)] = np.nan  # nor the diagonal of a unit triangle
    return np.asfortranarray(T)


# Triangles of order <= 128 take the serial diagonal path; larger ones go
# through a TRINV unit (inverted 16 x 16 diagonal blocks); above 1024 they nest.
TRSM_SIZES = [(1, 1), (16, 16), (17, 5), (100, 37), (37, 100), (128, 9), (129, 7), (7, 129),
              (300, 50), (50, 300), (1000, 64), (64, 1000), (1100, 20), (20, 1100)]


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["U", "N"])
@pytest.mark.parametrize("m,n", TRSM_SIZES)
def test_dtrsm(torch_gpu, side, uplo, trans, diag, m, n):
    mu = _mu()
    rng = np.random.default_rng(m * 1000 + n + ord(side) + 2 * ord(uplo) + 3 * ord(trans))
    na = m if side == "L" else n
    T = _tri(rng, na, uplo == "L", diag == "U")
    B = rand(rng, m, n)
    alpha = 1.0 if (m + n) % 2 else 0.75
    dT, dB = _dev(T), _dev(B)
    mu.dtrsm(side, uplo, trans, diag, m, n, alpha, mu.ptr(dT), na, mu.ptr(dB), m)
    torch_gpu.cuda.synchronize()
    got = mu.to_host(dB)
    Tc = np.where(np.isnan(T), 0.0, T)
    if diag == "U":
        Tc[np.arange(na), np.arange(na)] = 1.0
    opT = Tc.T if trans == "T" else Tc
    if side == "L":
        ref = sl.solve_triangular(opT, alpha * B, lower=(uplo == "L") != (trans == "T"))
    else:
        ref = sl.solve_triangular(opT.T, alpha * B.T, lower=(uplo == "L") == (trans == "T")).T
    assert np.all(np.isfinite(got))
    assert close_rel(got, ref, rtol=1e-11) <= 1.0


@pytest.mark.parametrize("n", [1, 2, 17, 64, 128, 129, 250, 1000])
def test_dgetrf_nopiv(torch_gpu, oracle, n):
    mu = _mu()
    rng = np.random.default_rng(n)
    A = diag_dominant(rng, n)
    ref = A.copy(order="F")
    assert oracle.getrf_nopiv(ref) == 0
    lu, piv = sl.lu_factor(A)
    assert np.array_equal(piv, np.arange(n))
    dA = _dev(A)
    info = mu.info_cell()
    mu.dgetrf_nopiv(n, mu.ptr(dA), n, mu.ptr(info))
    torch_gpu.cuda.synchronize()
    got = mu.to_host(dA)
    assert mu.info_value(info) == 0
    assert close_rel(got, ref, rtol=1e-11) <= 1.0
    assert close_rel(got, lu, rtol=1e-11) <= 1.0


@pytest.mark.parametrize("n", [1, 2, 17, 64, 128, 129, 250, 1000])
def test_dpotrf(torch_gpu, oracle, n):
    mu = _mu()
    rng = np.random.default_rng(n + 5)
    A = spd(rng, n)
    A[np.triu_indices(n, 1)] = 7.0  # upper triangle must be left as it is
    ref = A.copy(order="F")
    assert oracle.potrf_lower(ref) == 0
    dA = _dev(A)
    info = mu.info_cell()
    mu.dpotrf("L", n, mu.ptr(dA), n, mu.ptr(info))
    torch_gpu.cuda.synchronize()
    got = mu.to_host(dA)
    assert mu.info_value(info) == 0
    lo = np.tril(np.ones((n, n), dtype=bool))
    assert close_rel(got[lo], ref[lo], rtol=1e-11) <= 1.0
    assert np.array_equal(got[~lo], A[~lo])


@pytest.mark.parametrize("n,col", [(16, 0), (64, 40), (300, 200)])
def test_getrf_zero_pivot_info(torch_gpu, n, col):
    """A zero pivot is reported as LAPACK info = first failing column (1-based)."""
    mu = _mu()
    A = np.asfortranarray(np.eye(n) * 2.0)
    A[col, col] = 0.0
    dA = _dev(A)
    info = mu.info_cell()
    mu.dgetrf_nopiv(n, mu.ptr(dA), n, mu.ptr(info))
    torch_gpu.cuda.synchronize()
    assert mu.info_value(info) == col + 1


@pytest.mark.parametrize("n,col", [(16, 3), (64, 0), (300, 150)])
def test_potrf_not_spd_info(torch_gpu, n, col):
    mu = _mu()
    A = np.asfortranarray(np.eye(n) * 4.0)
    A[col, col] = -1.0
    dA = _dev(A)
    info = mu.info_cell()
    mu.dpotrf("L", n, mu.ptr(dA), n, mu.ptr(info))
    torch_gpu.cuda.synchronize()
    assert mu.info_value(info) == col + 1
