"""mu_dtrsm_noalloc against mu_dtrsm: the same triangular solves, bit for bit,
with the inverted diagonal blocks in the library's fixed device scratch
array instead of a per-call allocation.

Two calls in a row (the second smaller, so it runs over the first call's
stale scratch contents), a call at the largest accepted order (8192), the
small-order path that needs no scratch, and the refusals (order above 8192,
a stream being recorded), which return before any device work.
"""
import ctypes

import numpy as np
import pytest

import mustard_amd as mu
from mustard_amd import _lib


def _call(name, *args):
    return getattr(mu.lib(), name)(*args)


def _tri_case(torch, order, nrhs_or_rows, side, seed):
    """A well-conditioned triangle (random entries in [-1, 1), diagonal
    order + 1 in both triangles, so either uplo reads a good one) and a
    right-hand side, column-major on the device (tensor rows = columns)."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.rand((order, order), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    A.diagonal().add_(order + 1.0)
    rows, cols = (order, nrhs_or_rows) if side == "L" else (nrhs_or_rows, order)
    B = torch.rand((cols, rows), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    return A, B, rows, cols


def _both(torch, side, uplo, trans, diag, order, other, alpha, seed):
    A, B, m, n = _tri_case(torch, order, other, side, seed)
    X_ref, X_na = B.clone(), B.clone()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = (side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, alpha,
            mu.ptr(A), order)
    assert _call("mu_dtrsm", s, *args, mu.ptr(X_ref), m) == _lib.MU_OK
    assert _call("mu_dtrsm_noalloc", s, *args, mu.ptr(X_na), m) == _lib.MU_OK, \
        mu.lib().mu_last_error()
    torch.cuda.synchronize()
    return A, B, X_ref, X_na, m, n


def _residual_ok(torch, A, B, X, side, uplo, trans, diag, alpha):
    """max |op(A) X - alpha B| (or X op(A)) relative to the operands' scale."""
    Acm = A.T  # column-major storage -> the matrix
    T = torch.tril(Acm) if uplo == "L" else torch.triu(Acm)
    if diag == "U":
        T = T - torch.diag(torch.diagonal(T)) + torch.eye(T.shape[0], dtype=T.dtype,
                                                          device=T.device)
    if trans == "T":
        T = T.T
    Xm, Bm = X.T, B.T
    R = (T @ Xm if side == "L" else Xm @ T) - alpha * Bm
    scale = float(T.abs().max()) * float(Xm.abs().max()) * T.shape[0]
    return float(R.abs().max()) <= 1e-14 * scale


@pytest.mark.gpu
def test_two_calls_in_a_row_match_mu_dtrsm(torch_gpu):
    torch = torch_gpu
    cases = [
        # order 1000 nests into four blocks of 250, each with its own inverses
        ("L", "L", "N", "N", 1000, 300, 1.0, 1),
        # smaller: one triangle of order 200 with inverses, over the stale
        # contents the first call left in the scratch array
        ("R", "U", "T", "U", 200, 150, -0.5, 2),
    ]
    for side, uplo, trans, diag, order, other, alpha, seed in cases:
        print("mu_dtrsm_noalloc case", side, uplo, trans, diag, order, other, flush=True)
        A, B, X_ref, X_na, m, n = _both(torch, side, uplo, trans, diag, order, other, alpha,
                                        seed)
        assert torch.equal(X_ref, X_na)
        assert _residual_ok(torch, A, B, X_na, side, uplo, trans, diag, alpha)


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans,diag,order,other", [
    ("L", "U", "T", "N", 8192, 8),    # the largest order accepted
    ("L", "L", "N", "N", 8100, 8),    # the order whose nesting needs the most scratch
    ("R", "L", "N", "N", 100, 40),    # order <= 128: no inverses, no scratch
    ("L", "L", "T", "U", 4000, 33),   # ragged blocks of 250, transposed
])
def test_orders_match_mu_dtrsm(torch_gpu, side, uplo, trans, diag, order, other):
    torch = torch_gpu
    print("mu_dtrsm_noalloc case", side, uplo, trans, diag, order, other, flush=True)
    A, B, X_ref, X_na, m, n = _both(torch, side, uplo, trans, diag, order, other, 2.0, 3)
    assert torch.equal(X_ref, X_na)
    assert _residual_ok(torch, A, B, X_na, side, uplo, trans, diag, 2.0)


def test_refusals_before_any_device_work():
    """Order 8193 is above the order the array is sized for; a stream being
    recorded into a DAG is refused (nothing runs: null pointers are fine)."""
    s = ctypes.c_void_p(0)
    L = mu.lib()
    assert L.mu_dtrsm_noalloc(s, b"L", b"L", b"N", b"N", 8193, 4, 1.0, None, 8193, None,
                              8193) == _lib.MU_EINVAL
    assert b"8192" in L.mu_last_error()
    assert L.mu_dtrsm_noalloc(s, b"Q", b"L", b"N", b"N", 4, 4, 1.0, None, 4, None,
                              4) == _lib.MU_EINVAL
    d = mu.Dag()
    key = ctypes.c_void_p(4242)
    d.capture(key)
    d.begin_task((0, 0), [(0, 0)])
    assert L.mu_dtrsm_noalloc(key, b"L", b"L", b"N", b"N", 300, 4, 1.0, None, 300, None,
                              300) == _lib.MU_ESTATE
    d.end_task()
