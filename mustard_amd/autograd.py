"""torch.autograd binding of the tiled Cholesky and its blocked reverse mode.

`cholesky(A, tiles, prec)` returns the lower factor L of the symmetric
positive definite CUDA float64 matrix A; its backward runs mu_chol_backward
(three tiled DAGs of GEMM and TRSM tasks on the device) on the saved factor:
the forward result is kept and nothing is recomputed.  The gradient is the
symmetric G = (Abar + Abar^T)/2, Abar = L^-T Phi(L^T Lbar) L^-1, the same
convention as torch.linalg.cholesky's backward.
"""
from __future__ import annotations

import torch

from . import PREC, call, factorize, ptr, stream_handle


class TiledCholesky(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, tiles, prec):
        n = A.shape[0]
        # A is symmetric, so its row-major storage is also its column-major
        # storage; F keeps the factor exactly as the tiled Cholesky leaves it
        F = A.detach().contiguous().clone()
        info, _ = factorize("chol", F, n, tiles, prec=prec)
        if info != 0:
            raise RuntimeError(f"tiled Cholesky: leading minor {info} is not positive definite")
        ctx.save_for_backward(F)
        ctx.tiles, ctx.prec = tiles, prec
        # column-major lower factor = row-major upper: transpose back to L
        return torch.tril(F.t())

    @staticmethod
    def backward(ctx, gL):
        (F,) = ctx.saved_tensors
        n = F.shape[0]
        Lb = gL.t().contiguous()          # column-major copy of gL
        G = torch.empty_like(F)
        call("mu_chol_backward", ptr(F), n, ptr(Lb), n, ptr(G), n, n, ctx.tiles, ctx.prec,
             stream_handle(torch.cuda.current_stream()))
        return G, None, None              # G is symmetric: row- and column-major agree


def cholesky(A, tiles=1, prec="fp64"):
    """Lower Cholesky factor of A (n x n CUDA float64, n % tiles == 0), with a
    tiled backward.  prec: precision of the trailing GEMM/SYRK tasks of the
    forward and of the GEMM tasks of the backward ("fp64", "fp32", "bf16")."""
    return TiledCholesky.apply(A, tiles, PREC[prec] if isinstance(prec, str) else prec)
