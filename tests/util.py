"""Helpers shared by the tests (matrix builders and tolerances)."""
import numpy as np

EPS = np.finfo(np.float64).eps


def rand(rng, m, n, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, size=(m, n)))


def diag_dominant(rng, n):
    A = rand(rng, n, n, 0.0, 1.0)
    A[np.arange(n), np.arange(n)] += n
    return A


def spd(rng, n):
    M = rand(rng, n, n)
    A = M @ M.T / max(n, 1) + n * np.eye(n)
    return np.asfortranarray((A + A.T) / 2)


def close_rel(X, Y, rtol=1e-10, floor_frac=1e-3):
    """|x - y| <= rtol * max(|y|, floor_frac * max|Y|) elementwise; returns the
    worst ratio (<= 1 passes)."""
    X = np.asarray(X)
    Y = np.asarray(Y)
    if Y.size == 0:
        return 0.0
    scale = np.maximum(np.abs(Y), floor_frac * np.max(np.abs(Y)) + 1e-300)
    return float(np.max(np.abs(X - Y) / (rtol * scale)))
