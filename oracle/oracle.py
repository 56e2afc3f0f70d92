"""CPU oracle for the single-GPU tiled LU / Cholesky path (numpy + ctypes).

TEST INFRASTRUCTURE ONLY.  Importable from tests/, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- always as the checker or the
reported CPU baseline, never as the measured or shipped path.  The product
package ``mustard_amd`` never imports this module.

What lives here:

* thin numpy wrappers over ``liboracle.so`` (``oracle.c``), which restates the
  reference's generator (include/gen.h:23-38), its tile library calls
  (mustard/lu_mustard.cu:360-522, mustard/cholesky_mustard.cu:238-387), its
  task order and its verifier (include/verify.h:17-129);
* :func:`task_dag`, a restatement of the reference's task recording and
  dependency rule (mustard/mustard.h:148-181, 279-296) over the task order of
  lu_mustard.cu:343-535 / cholesky_mustard.cu:221-400.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags["F_CONTIGUOUS"] or a.flags["C_CONTIGUOUS"])
    return a.ctypes.data_as(_dp)


def blas_candidates():
    """Host LAPACK libraries in preference order: (path, symbol prefix)."""
    out = []
    for p in ("/usr/lib/x86_64-linux-gnu/libopenblas.so.0",
              "/usr/lib/x86_64-linux-gnu/liblapack.so.3"):
        if os.path.exists(p):
            out.append((p, ""))
    try:
        import scipy  # noqa: F401  (only to find its bundled OpenBLAS)
        root = os.path.dirname(os.path.dirname(scipy.__file__))
        for p in sorted(glob.glob(os.path.join(root, "scipy.libs", "libscipy_openblas*.so"))):
            out.append((p, "scipy_"))
    except ImportError:
        pass
    return out


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    """Load liboracle.so (building it if it is missing) and a host LAPACK."""
    global _LIB, BLAS_PATH
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    i, d, sz, u64 = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_uint64
    sig = {
        "or_gen_ref": (None, [_dp, sz]),
        "or_u01": (d, [u64, u64]),
        "or_gen_uniform": (None, [_dp, sz, u64, d, d]),
        "or_gen_spd": (None, [_dp, sz, u64]),
        "or_blas_load": (i, [ctypes.c_char_p, ctypes.c_char_p]),
        "or_blas_loaded": (i, []),
        "or_blas_set_threads": (None, [i]),
        "or_getrf_nopiv": (i, [i, _dp, i]),
        "or_getrf_nopiv_unblocked": (i, [i, _dp, i]),
        "or_potrf_lower": (i, [i, _dp, i]),
        "or_potrf_lower_unblocked": (i, [i, _dp, i]),
        "or_trsm_llnu": (None, [i, i, _dp, i, _dp, i]),
        "or_trsm_runn": (None, [i, i, _dp, i, _dp, i]),
        "or_trsm_rltn": (None, [i, i, _dp, i, _dp, i]),
        "or_gemm_nn_sub": (None, [i, i, i, _dp, i, _dp, i, _dp, i]),
        "or_gemm_nt_sub": (None, [i, i, i, _dp, i, _dp, i, _dp, i]),
        "or_syrk_ln_sub": (None, [i, i, _dp, i, _dp, i]),
        "or_tiled_lu": (i, [i, i, _dp]),
        "or_tiled_cholesky": (i, [i, i, _dp]),
        "or_tiled_lu_omp": (i, [i, i, _dp, i]),
        "or_tiled_cholesky_omp": (i, [i, i, _dp, i]),
        "or_verify_lu_l1": (d, [_dp, _dp, i]),
        "or_verify_chol_l1": (d, [_dp, _dp, i]),
        "or_scaled_residual_lu": (d, [_dp, _dp, i]),
        "or_scaled_residual_chol": (d, [_dp, _dp, i]),
        "or_lu_solve": (None, [i, i, _dp, _dp]),
        "or_chol_solve": (None, [i, i, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    BLAS_PATH = None
    for p, prefix in blas_candidates():
        if L.or_blas_load(p.encode(), prefix.encode()) == 0:
            BLAS_PATH = p
            break
    _LIB = L
    return L


BLAS_PATH = None


# --------------------------------------------------------------------------
# generators
# --------------------------------------------------------------------------
def gen_ref(n: int) -> np.ndarray:
    """include/gen.h:23-38 (glibc rand(), seed 420). Column-major n x n."""
    A = np.empty((n, n), dtype=np.float64, order="F")
    lib().or_gen_ref(_ptr(A), n)
    return A


def gen_uniform(shape, seed: int, lo: float, hi: float) -> np.ndarray:
    """Counter-based uniform stream, element k (column-major) = stream value k."""
    M = np.empty(shape, dtype=np.float64, order="F")
    lib().or_gen_uniform(_ptr(M), M.size, seed, lo, hi)
    return M


def u01(seed: int, k: int) -> float:
    return lib().or_u01(seed, k)


def gen_spd(n: int, seed: int) -> np.ndarray:
    """A = M M^T / n + n I with M uniform [-1, 1) from `seed`."""
    A = np.empty((n, n), dtype=np.float64, order="F")
    lib().or_gen_spd(_ptr(A), n, seed)
    return A


# --------------------------------------------------------------------------
# tile operations (all in place on Fortran-ordered arrays)
# --------------------------------------------------------------------------
def _f(a):
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"], "need a Fortran-ordered float64 array"
    return _ptr(a)


def getrf_nopiv(A, unblocked=False):
    n = A.shape[0]
    f = lib().or_getrf_nopiv_unblocked if unblocked else lib().or_getrf_nopiv
    return f(n, _f(A), n)


def potrf_lower(A, unblocked=False):
    n = A.shape[0]
    f = lib().or_potrf_lower_unblocked if unblocked else lib().or_potrf_lower
    return f(n, _f(A), n)


def trsm_llnu(L, B):
    lib().or_trsm_llnu(B.shape[0], B.shape[1], _f(L), L.shape[0], _f(B), B.shape[0])


def trsm_runn(U, B):
    lib().or_trsm_runn(B.shape[0], B.shape[1], _f(U), U.shape[0], _f(B), B.shape[0])


def trsm_rltn(L, B):
    lib().or_trsm_rltn(B.shape[0], B.shape[1], _f(L), L.shape[0], _f(B), B.shape[0])


def gemm_nn_sub(A, B, C):
    lib().or_gemm_nn_sub(C.shape[0], C.shape[1], A.shape[1], _f(A), A.shape[0], _f(B),
                         B.shape[0], _f(C), C.shape[0])


def gemm_nt_sub(A, B, C):
    lib().or_gemm_nt_sub(C.shape[0], C.shape[1], A.shape[1], _f(A), A.shape[0], _f(B),
                         B.shape[0], _f(C), C.shape[0])


def syrk_ln_sub(A, C):
    lib().or_syrk_ln_sub(C.shape[0], A.shape[1], _f(A), A.shape[0], _f(C), C.shape[0])


# --------------------------------------------------------------------------
# tiled factorizations
# --------------------------------------------------------------------------
def tiled_lu(A, t, threads=0):
    """In place; reference task order. threads > 0 runs the OpenMP task DAG."""
    n = A.shape[0]
    if threads > 0:
        return lib().or_tiled_lu_omp(n, t, _f(A), threads)
    return lib().or_tiled_lu(n, t, _f(A))


def tiled_cholesky(A, t, threads=0):
    n = A.shape[0]
    if threads > 0:
        return lib().or_tiled_cholesky_omp(n, t, _f(A), threads)
    return lib().or_tiled_cholesky(n, t, _f(A))


def verify_lu_l1(A, F):
    return lib().or_verify_lu_l1(_f(A), _f(F), A.shape[0])


def verify_chol_l1(A, F):
    return lib().or_verify_chol_l1(_f(A), _f(F), A.shape[0])


def scaled_residual_lu(A, F):
    return lib().or_scaled_residual_lu(_f(A), _f(F), A.shape[0])


def scaled_residual_chol(A, F):
    return lib().or_scaled_residual_chol(_f(A), _f(F), A.shape[0])


def lu_solve(F, B):
    X = np.array(B, dtype=np.float64, order="F", copy=True)
    lib().or_lu_solve(F.shape[0], X.shape[1], _f(F), _f(X))
    return X


def chol_solve(F, B):
    X = np.array(B, dtype=np.float64, order="F", copy=True)
    lib().or_chol_solve(F.shape[0], X.shape[1], _f(F), _f(X))
    return X


def gen_cond(n: int, kappa: float, seed: int, nrefl: int = 4, _tau_scale: float = 2.0):
    """Restatement of mu_gen_cond: A = H_r .. H_1 diag(s) H_1 .. H_r with
    s_i = kappa^(-i/(n-1)) and H = I - 2 v v^T / (v.v), v uniform [-1, 1)
    from stream seed + r; each step A <- H A H is applied as
    A - tau (v w^T + w v^T) + tau^2 (v.w) v v^T with w = A v.  The eigenvalues
    of the result are exactly s (the closed form the tests check).
    `_tau_scale` exists only so a test can plant a wrong reflector."""
    s = kappa ** (-np.arange(n) / (n - 1)) if n > 1 else np.ones(1)
    A = np.diag(s)
    for r in range(nrefl):
        v = gen_uniform((n,), seed + r, -1.0, 1.0)
        w = A @ v
        tau = _tau_scale / (v @ v)
        c = v @ w
        A = A - tau * (np.outer(v, w) + np.outer(w, v)) + tau * tau * c * np.outer(v, v)
    return np.asfortranarray(A)


def backward_error(A, X, B):
    """Normwise backward error, max over columns j of
    ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)."""
    R = B - A @ X
    an = np.max(np.sum(np.abs(A), axis=1))
    num = np.max(np.abs(R), axis=0)
    den = an * np.max(np.abs(X), axis=0) + np.max(np.abs(B), axis=0)
    return float(np.max(num / den))


def solve_fp64(kind: str, A, B):
    """FP64 LAPACK-path solution: no-pivot getrf + getrs, or potrf + potrs."""
    F = np.array(A, dtype=np.float64, order="F", copy=True)
    if kind == "lu":
        assert getrf_nopiv(F) == 0
        return lu_solve(F, B)
    assert potrf_lower(F) == 0
    return chol_solve(F, B)


def chol_backward(L, Lbar):
    """Restated reverse mode of the Cholesky factorization (Murray,
    arXiv:1602.07527, eq. 10): G = (Abar + Abar^T)/2 with
    Abar = L^-T Phi(L^T Lbar) L^-1, Phi = lower triangle with the diagonal
    halved.  Dense FP64."""
    import scipy.linalg as sl
    L = np.tril(L)
    Lbar = np.tril(Lbar)
    X = L.T @ Lbar
    P = np.tril(X) - 0.5 * np.diag(np.diag(X))
    W = sl.solve_triangular(L.T, P.T, lower=False).T       # W L = P
    Abar = sl.solve_triangular(L.T, W, lower=False)        # L^T Abar = W
    return 0.5 * (Abar + Abar.T)


def c3_loss_grad(A, B):
    """Loss sum(log diag L) + 0.5 ||L^-1 B||_F^2 with A = L L^T, which equals
    0.5 log det A + 0.5 tr(B^T A^-1 B), its Lbar, and its closed-form symmetric
    gradient 0.5 A^-1 - 0.5 A^-1 B B^T A^-1."""
    import scipy.linalg as sl
    L = np.linalg.cholesky(A)
    Y = sl.solve_triangular(L, B, lower=True)
    loss = float(np.sum(np.log(np.diag(L))) + 0.5 * np.sum(Y * Y))
    Z = sl.solve_triangular(L.T, Y, lower=False)
    Lbar = np.diag(1.0 / np.diag(L)) - np.tril(Z @ Y.T)
    Ai = np.linalg.inv(A)
    G = 0.5 * Ai - 0.5 * Ai @ B @ B.T @ Ai
    return loss, L, Lbar, G


# --------------------------------------------------------------------------
# task DAG restatement
# --------------------------------------------------------------------------
def task_ops(kind: str, T: int):
    """The reference's task sequence: list of (op, write_tile, read_tiles).

    LU follows lu_mustard.cu:343-535, Cholesky cholesky_mustard.cu:221-400.
    Tiles are (row block, column block).
    """
    ops = []
    for k in range(T):
        if kind == "lu":
            ops.append(("GETRF", (k, k), [(k, k)]))
            for i in range(k + 1, T):
                ops.append(("TRSM_L", (k, i), [(k, k), (k, i)]))
            for i in range(k + 1, T):
                ops.append(("TRSM_R", (i, k), [(k, k), (i, k)]))
                for j in range(k + 1, T):
                    ops.append(("GEMM", (i, j), [(i, k), (k, j), (i, j)]))
        else:
            ops.append(("POTRF", (k, k), [(k, k)]))
            for i in range(k + 1, T):
                ops.append(("TRSM", (i, k), [(k, k), (i, k)]))
            for i in range(k + 1, T):
                ops.append(("SYRK", (i, i), [(i, i), (i, k)]))
                for j in range(i + 1, T):
                    ops.append(("GEMM", (j, i), [(j, i), (j, k), (i, k)]))
    return ops


def task_dag(kind: str, T: int):
    """Dependencies per task: sorted unique indices of the last writers of the
    read tiles and of the written tile (mustard/mustard.h:148-181, 279-296)."""
    last = {}
    deps = []
    ops = task_ops(kind, T)
    for idx, (_, w, r) in enumerate(ops):
        d = sorted({last[t] for t in list(r) + [w] if t in last})
        deps.append(d)
        last[w] = idx
    return ops, deps


def critical_path(deps):
    """Length (in tasks) of the longest chain of the DAG."""
    depth = []
    for d in deps:
        depth.append(1 + max((depth[p] for p in d), default=0))
    return max(depth) if depth else 0
