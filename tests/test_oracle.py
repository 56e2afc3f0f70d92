"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something it was not derived from:
the golden fixtures (built from libc rand() in Python and from LAPACK), SciPy's
LAPACK, closed forms, brute force, and the reference's own task-count
formulas.  Tests that plant a mistake plant it in the oracle's inputs, on the
CPU, and show the checks catch it.
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

from tests.util import EPS, close_rel, diag_dominant, rand, spd

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _json(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ---- generator (include/gen.h:23-38) --------------------------------------
def test_gen_ref_matches_fixture(oracle):
    g = _json("generator.json")
    A = oracle.gen_ref(600)
    assert np.array_equal(A.T.reshape(-1)[:16], np.array(g["n600_first16_rowmajor"]))
    assert hashlib.sha256(A.tobytes(order="F")).hexdigest() == g["n600_sha256"]
    assert hashlib.sha256(oracle.gen_ref(64).tobytes(order="F")).hexdigest() == g["n64_sha256"]


def test_gen_ref_properties(oracle):
    n = 300
    A = oracle.gen_ref(n)
    assert np.array_equal(A, A.T)
    off = A - np.diag(np.diag(A))
    assert off.min() >= 0.0 and off.max() <= 1.0
    # strictly diagonally dominant: row sums of |off| < n - 1 < diagonal
    assert np.all(np.diag(A) > np.sum(np.abs(off), axis=1))


def test_product_generator_matches_oracle(oracle):
    import mustard_amd as mu
    assert np.array_equal(mu.gen_ref_host(257), oracle.gen_ref(257))


def test_counter_rng(oracle):
    M = oracle.gen_uniform((50, 40), 420, -1.0, 1.0)
    assert M.min() >= -1.0 and M.max() < 1.0
    assert abs(M.mean()) < 0.05
    # element k of the stream depends on (seed, k) only
    assert oracle.u01(420, 7) == (M.reshape(-1, order="F")[7] + 1.0) / 2.0
    assert oracle.u01(421, 7) != oracle.u01(420, 7)


def test_gen_spd(oracle):
    n = 200
    A = oracle.gen_spd(n, 420)
    M = oracle.gen_uniform((n, n), 420, -1.0, 1.0)
    ref = M @ M.T / n + n * np.eye(n)
    assert np.array_equal(A, A.T)
    assert np.max(np.abs(A - ref)) <= 1e-13 * n
    assert np.all(np.linalg.eigvalsh(A) > 0)


# ---- tile operations against LAPACK ---------------------------------------
@pytest.mark.parametrize("n", [1, 5, 64, 65, 200])
@pytest.mark.parametrize("unblocked", [False, True])
def test_getrf_nopiv_vs_lapack(oracle, n, unblocked):
    A = diag_dominant(np.random.default_rng(n), n)
    F = A.copy(order="F")
    assert oracle.getrf_nopiv(F, unblocked=unblocked) == 0
    lu, piv = sl.lu_factor(A)
    assert np.array_equal(piv, np.arange(n))
    assert close_rel(F, lu, rtol=1e-12) <= 1.0


def test_getrf_zero_pivot(oracle):
    A = np.asfortranarray(np.eye(6))
    A[3, 3] = 0.0
    assert oracle.getrf_nopiv(A.copy(order="F")) == 4
    assert oracle.getrf_nopiv(A.copy(order="F"), unblocked=True) == 4


@pytest.mark.parametrize("n", [1, 7, 64, 150])
def test_potrf_vs_numpy(oracle, n):
    A = spd(np.random.default_rng(n), n)
    for unblocked in (False, True):
        F = A.copy(order="F")
        assert oracle.potrf_lower(F, unblocked=unblocked) == 0
        assert close_rel(np.tril(F), np.linalg.cholesky(A), rtol=1e-12) <= 1.0
        assert np.array_equal(np.triu(F, 1), np.triu(A, 1))  # upper untouched


def test_potrf_not_spd(oracle):
    A = np.asfortranarray(np.diag([4.0, 1.0, -2.0, 3.0]))
    assert oracle.potrf_lower(A.copy(order="F")) == 3
    assert oracle.potrf_lower(A.copy(order="F"), unblocked=True) == 3


def test_trsm_variants(oracle):
    rng = np.random.default_rng(3)
    m, n = 37, 23
    Lm = np.tril(rand(rng, m, m)) + 3 * np.eye(m)
    B = rand(rng, m, n)
    X = B.copy(order="F")
    oracle.trsm_llnu(np.asfortranarray(Lm), X)
    Lu = np.tril(Lm, -1) + np.eye(m)
    assert close_rel(X, sl.solve_triangular(Lu, B, lower=True), 1e-12) <= 1
    U = np.triu(rand(rng, n, n)) + 3 * np.eye(n)
    C = rand(rng, m, n)
    X = C.copy(order="F")
    oracle.trsm_runn(np.asfortranarray(U), X)
    assert close_rel(X @ U, C, 1e-12) <= 1
    Ln = np.tril(rand(rng, n, n)) + 3 * np.eye(n)
    X = C.copy(order="F")
    oracle.trsm_rltn(np.asfortranarray(Ln), X)
    assert close_rel(X @ Ln.T, C, 1e-12) <= 1


def test_gemm_syrk_subs(oracle):
    rng = np.random.default_rng(4)
    A, B, C = rand(rng, 20, 9), rand(rng, 9, 13), rand(rng, 20, 13)
    X = C.copy(order="F")
    oracle.gemm_nn_sub(A, B, X)
    assert np.allclose(X, C - A @ B, atol=1e-14)
    Bt = rand(rng, 13, 9)
    X = C.copy(order="F")
    oracle.gemm_nt_sub(A, Bt, X)
    assert np.allclose(X, C - A @ Bt.T, atol=1e-14)
    S = rand(rng, 20, 20)
    X = S.copy(order="F")
    oracle.syrk_ln_sub(A, X)
    lo = np.tril(np.ones((20, 20), dtype=bool))
    assert np.allclose(X[lo], (S - A @ A.T)[lo], atol=1e-14)
    assert np.array_equal(X[~lo], S[~lo])


# ---- tiled factorizations (reference task order) --------------------------
@pytest.mark.parametrize("n,t", [(64, 1), (64, 4), (600, 2), (600, 6), (300, 10)])
def test_tiled_lu_equals_lapack(oracle, n, t):
    A = oracle.gen_ref(n)
    F = A.copy(order="F")
    assert oracle.tiled_lu(F, t) == 0
    lu, piv = sl.lu_factor(A)
    assert np.array_equal(piv, np.arange(n))
    assert close_rel(F, lu) <= 1.0
    assert oracle.scaled_residual_lu(A, F) <= 10.0


@pytest.mark.parametrize("n,t", [(64, 1), (64, 4), (600, 2), (300, 10)])
def test_tiled_cholesky_equals_lapack(oracle, n, t):
    A = oracle.gen_ref(n)
    F = A.copy(order="F")
    assert oracle.tiled_cholesky(F, t) == 0
    assert close_rel(np.tril(F), sl.cholesky(A, lower=True)) <= 1.0
    assert oracle.scaled_residual_chol(A, F) <= 10.0


def test_lapack_fixture_n64(oracle):
    z = np.load(os.path.join(GOLD, "lapack_n64.npz"))
    A = oracle.gen_ref(64)
    assert np.array_equal(A, z["A"])
    F = A.copy(order="F")
    oracle.tiled_lu(F, 4)
    assert close_rel(F, z["lu"]) <= 1.0
    C = A.copy(order="F")
    oracle.tiled_cholesky(C, 4)
    assert close_rel(np.tril(C), z["chol"]) <= 1.0


def test_lapack_fixture_n600(oracle):
    j = _json("lapack_n600.json")
    A = oracle.gen_ref(600)
    F = A.copy(order="F")
    oracle.tiled_lu(F, 2)
    C = A.copy(order="F")
    oracle.tiled_cholesky(C, 2)
    L = np.tril(C)
    assert abs(np.linalg.norm(F) - j["lu_fro"]) <= 1e-12 * j["lu_fro"]
    assert abs(np.linalg.norm(L) - j["chol_fro"]) <= 1e-12 * j["chol_fro"]
    for a, b, lu_v, ch_v in j["samples"]:
        assert abs(F[a, b] - lu_v) <= 1e-10 * max(abs(lu_v), 1e-3)
        if ch_v is not None:
            assert abs(L[a, b] - ch_v) <= 1e-10 * max(abs(ch_v), 1e-3)


@pytest.mark.parametrize("kind", ["lu", "chol"])
def test_openmp_dag_equals_sequential(oracle, kind):
    """The OpenMP task-DAG baseline computes bit-identical factors."""
    A = oracle.gen_ref(400)
    F1, F2 = A.copy(order="F"), A.copy(order="F")
    f = oracle.tiled_lu if kind == "lu" else oracle.tiled_cholesky
    f(F1, 8)
    f(F2, 8, threads=4)
    assert np.array_equal(F1, F2)


# ---- verification formulas (include/verify.h) -----------------------------
def test_verify_l1_formulas(oracle):
    A = oracle.gen_ref(40)
    F = A.copy(order="F")
    oracle.tiled_lu(F, 4)
    Lu = np.tril(F, -1) + np.eye(40)
    U = np.triu(F)
    assert abs(oracle.verify_lu_l1(A, F) - np.sum(np.abs(A - Lu @ U))) <= 1e-9
    C = A.copy(order="F")
    oracle.tiled_cholesky(C, 4)
    L = np.tril(C)
    assert abs(oracle.verify_chol_l1(A, C) - np.sum(np.abs(A - L @ L.T))) <= 1e-9
    assert oracle.verify_lu_l1(A, F) <= 1e-6 and oracle.verify_chol_l1(A, C) <= 1e-6


def test_planted_mistakes_are_caught(oracle):
    """A wrong factor fails every check: one perturbed entry, a transposed
    tile, and a factorization of the wrong matrix."""
    n = 120
    A = oracle.gen_ref(n)
    F = A.copy(order="F")
    oracle.tiled_lu(F, 4)
    bad = F.copy(order="F")
    bad[70, 20] += 1e-6
    assert oracle.verify_lu_l1(A, bad) > 1e-6
    assert oracle.scaled_residual_lu(A, bad) > 10
    bad = F.copy(order="F")
    bad[30:60, 60:90] = bad[30:60, 60:90].T.copy()
    assert oracle.scaled_residual_lu(A, bad) > 10
    A2 = A.copy(order="F")
    A2[5, 9] += 1.0
    F2 = A2.copy(order="F")
    oracle.tiled_lu(F2, 4)
    assert oracle.scaled_residual_lu(A, F2) > 10


def test_solves(oracle):
    rng = np.random.default_rng(9)
    A = diag_dominant(rng, 50)
    F = A.copy(order="F")
    oracle.getrf_nopiv(F)
    B = rand(rng, 50, 3)
    X = oracle.lu_solve(F, B)
    assert np.max(np.abs(A @ X - B)) <= 1e-12
    S = spd(rng, 50)
    C = S.copy(order="F")
    oracle.potrf_lower(C)
    X = oracle.chol_solve(C, B)
    assert np.max(np.abs(S @ X - B)) <= 1e-12


# ---- task DAG (mustard.h:148-181, 279-296) ---------------------------------
def _ref_total_nodes(kind, T):
    # lu_mustard.cu:322-326 and cholesky_mustard.cu:200-204
    tot = T
    for k in range(T):
        for i in range(k + 1, T):
            tot += 2 + (T - (k + 1) if kind == "lu" else T - (i + 1))
    return tot


@pytest.mark.parametrize("kind", ["lu", "chol"])
def test_task_counts_match_reference_formula(oracle, kind):
    for T in range(1, 31):
        ops, deps = oracle.task_dag(kind, T)
        assert len(ops) == _ref_total_nodes(kind, T)
        for i, d in enumerate(deps):
            assert all(p < i for p in d)


def test_task_dag_fixture_and_counts(oracle):
    gold = _json("dag.json")
    for key, v in gold.items():
        kind, T = key.split("_")
        ops, deps = oracle.task_dag(kind, int(T))
        assert [[o, list(w)] for o, w, _ in ops] == v["ops"]
        assert deps == v["deps"]
    ops, deps = oracle.task_dag("lu", 2)
    assert [o for o, _, _ in ops] == ["GETRF", "TRSM_L", "TRSM_R", "GEMM", "GETRF"]
    assert deps == [[], [0], [0], [1, 2], [3]]
    for kind, edges in (("lu", 13524), ("chol", 6900)):
        ops, deps = oracle.task_dag(kind, 24)
        assert len(ops) == (4900 if kind == "lu" else 2600)
        assert sum(map(len, deps)) == edges
        assert oracle.critical_path(deps) == 70


@pytest.mark.parametrize("kappa", [1.0, 1e3, 1e6])
def test_gen_cond_spectrum(oracle, kappa):
    """The conditioned generator's closed form: eigenvalues are exactly
    kappa^(-i/(n-1)), the matrix is symmetric, and a wrong reflector scale
    (planted) breaks the spectrum."""
    n = 120
    A = oracle.gen_cond(n, kappa, 11, nrefl=4)
    s = np.sort(kappa ** (-np.arange(n) / (n - 1)))
    assert np.allclose(A, A.T, rtol=0, atol=1e-15)
    ev = np.sort(np.linalg.eigvalsh(A))
    assert np.max(np.abs(ev - s)) <= 1e-13
    assert abs(np.linalg.cond(A) / kappa - 1) < 1e-6
    # not diagonal: the reflectors really mix it (kappa = 1 gives H I H = I)
    if kappa > 1:
        assert np.max(np.abs(A - np.diag(np.diag(A)))) > 1e-3 / kappa
    bad = oracle.gen_cond(n, kappa, 11, nrefl=4, _tau_scale=1.0)
    assert np.max(np.abs(np.sort(np.linalg.eigvalsh(bad)) - s)) > 1e-6 or kappa == 1.0


def test_backward_error_and_fp64_solve(oracle):
    n = 80
    A = oracle.gen_cond(n, 1e3, 3)
    B = oracle.gen_uniform((n, 4), 5, -1.0, 1.0)
    for kind in ("lu", "chol"):
        X = oracle.solve_fp64(kind, A, B)
        assert np.allclose(X, np.linalg.solve(A, B), rtol=1e-9, atol=0)
        be = oracle.backward_error(A, X, B)
        assert be <= 10 * n * EPS
        # a perturbed solution has a much larger backward error
        assert oracle.backward_error(A, X * (1 + 1e-6), B) > 100 * be


def test_chol_backward_oracle_pinned(oracle):
    """The restated Cholesky adjoint equals the closed-form gradient of the
    config loss, agrees with central differences along random symmetric
    directions, and a planted mistake (Phi without halving the diagonal)
    is caught."""
    rng = np.random.default_rng(3)
    n, m = 64, 32
    A = oracle.gen_spd(n, 5)
    B = rng.uniform(-1, 1, (n, m))
    loss, L, Lbar, Gc = oracle.c3_loss_grad(A, B)
    G = oracle.chol_backward(L, Lbar)
    assert np.max(np.abs(G - Gc)) <= 1e-12 * np.max(np.abs(Gc))
    h = 1e-6
    for _ in range(20):
        E = rng.uniform(-1, 1, (n, n))
        E = E + E.T
        fd = (oracle.c3_loss_grad(A + h * E, B)[0] - oracle.c3_loss_grad(A - h * E, B)[0]) / (2 * h)
        assert abs(fd - np.sum(G * E)) <= 1e-6 * max(1.0, abs(fd))
    X = L.T @ np.tril(Lbar)
    bad_P = np.tril(X)  # diagonal not halved
    import scipy.linalg as sl
    W = sl.solve_triangular(L.T, bad_P.T, lower=False).T
    Ab = sl.solve_triangular(L.T, W, lower=False)
    assert np.max(np.abs(0.5 * (Ab + Ab.T) - Gc)) > 1e-6
