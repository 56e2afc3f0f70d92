"""Writes the golden fixtures of tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

The reference ships no tests, fixtures or golden vectors, and its arithmetic
lives in cuSOLVER/cuBLAS (closed, CUDA >= 12.3, not runnable here).  The
fixtures are therefore built from independent sources:

* generator.json -- include/gen.h:23-38 restated in pure Python on top of the
  C library's own rand()/srand() (through ctypes): the first 16 values of the
  N=600 matrix (row-major view), and SHA-256 digests of the whole N=64 and
  N=600 matrices;
* lapack_n64.npz -- LAPACK dgetrf (SciPy) and dpotrf factors of the N=64
  reference matrix (dgetrf picks identity pivots on it: the matrix is strictly
  diagonally dominant), plus the matrix;
* lapack_n600.json -- norms and sampled entries of the N=600 LAPACK factors;
* dag.json -- the task sequence and dependency lists of tiled LU and
  Cholesky for T = 2, 3, 4, derived by hand from lu_mustard.cu:343-535,
  cholesky_mustard.cu:221-400 and the last-writer rule of mustard.h:279-296,
  re-derived here with a second, independent implementation (a write-set scan
  rather than a last-writer map).
"""
import ctypes
import ctypes.util
import hashlib
import json
import os

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))


def gen_ref_python(n):
    """include/gen.h:23-38 in Python over libc rand()."""
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.rand.restype = ctypes.c_int
    RAND_MAX = 2147483647  # glibc
    libc.srand(420)
    A = np.zeros(n * n)
    for i in range(n):
        for j in range(i, n):
            A[i * n + j] = float(libc.rand()) / float(RAND_MAX)
    for i in range(n):
        for j in range(i, -1, -1):
            A[i * n + j] = A[j * n + i]
    for i in range(n):
        A[i * n + i] = A[i * n + i] + n
    return A.reshape(n, n)  # symmetric: row- and column-major views agree


def task_seq(kind, T):
    seq = []
    for k in range(T):
        if kind == "lu":
            seq.append(("GETRF", (k, k), [(k, k)]))
            seq += [("TRSM_L", (k, i), [(k, k), (k, i)]) for i in range(k + 1, T)]
            for i in range(k + 1, T):
                seq.append(("TRSM_R", (i, k), [(k, k), (i, k)]))
                seq += [("GEMM", (i, j), [(i, k), (k, j), (i, j)]) for j in range(k + 1, T)]
        else:
            seq.append(("POTRF", (k, k), [(k, k)]))
            seq += [("TRSM", (i, k), [(k, k), (i, k)]) for i in range(k + 1, T)]
            for i in range(k + 1, T):
                seq.append(("SYRK", (i, i), [(i, i), (i, k)]))
                seq += [("GEMM", (j, i), [(j, i), (j, k), (i, k)]) for j in range(i + 1, T)]
    return seq


def deps_by_scan(seq):
    """For each task: the latest earlier writer of each tile it touches."""
    out = []
    for idx, (_, w, r) in enumerate(seq):
        d = set()
        for t in r + [w]:
            for prev in range(idx - 1, -1, -1):
                if seq[prev][1] == t:
                    d.add(prev)
                    break
        out.append(sorted(d))
    return out


def main():
    g = {}
    A600 = gen_ref_python(600)
    g["n600_first16_rowmajor"] = A600.reshape(-1)[:16].tolist()
    g["n600_sha256"] = hashlib.sha256(np.asfortranarray(A600).tobytes(order="F")).hexdigest()
    A64 = gen_ref_python(64)
    g["n64_sha256"] = hashlib.sha256(np.asfortranarray(A64).tobytes(order="F")).hexdigest()
    with open(os.path.join(HERE, "generator.json"), "w") as f:
        json.dump(g, f, indent=1)

    lu, piv = sl.lu_factor(A64)
    assert np.array_equal(piv, np.arange(64)), "identity pivots expected"
    L = sl.cholesky(A64, lower=True)
    np.savez(os.path.join(HERE, "lapack_n64.npz"), A=A64, lu=lu, chol=L)

    lu6, piv6 = sl.lu_factor(A600)
    assert np.array_equal(piv6, np.arange(600))
    L6 = sl.cholesky(A600, lower=True)
    rng = np.random.default_rng(7)
    idx = rng.integers(0, 600, size=(64, 2)).tolist()
    j = {
        "lu_fro": float(np.linalg.norm(lu6)),
        "chol_fro": float(np.linalg.norm(L6)),
        "lu_diag_prod_log": float(np.sum(np.log(np.abs(np.diag(lu6))))),
        "chol_diag_sum_log": float(np.sum(np.log(np.diag(L6)))),
        "samples": [[a, b, float(lu6[a, b]), float(L6[a, b]) if a >= b else None] for a, b in idx],
    }
    with open(os.path.join(HERE, "lapack_n600.json"), "w") as f:
        json.dump(j, f, indent=1)

    dag = {}
    for kind in ("lu", "chol"):
        for T in (2, 3, 4):
            seq = task_seq(kind, T)
            dag[f"{kind}_{T}"] = {
                "ops": [[o, list(w)] for o, w, _ in seq],
                "deps": deps_by_scan(seq),
            }
    with open(os.path.join(HERE, "dag.json"), "w") as f:
        json.dump(dag, f)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
