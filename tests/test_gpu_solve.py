"""GPU tests of the solve phase and of the mixed-precision + FP64 iterative
refinement solver (mu_record_tiled_solve, mu_solve_ir), and of the
conditioned test-matrix generator (mu_gen_cond), against the oracle and the
FP64 LAPACK-path solution."""
import numpy as np
import pytest

from tests.util import EPS

pytestmark = pytest.mark.gpu


def _mu():
    import mustard_amd as mu
    return mu


def test_gen_cond_matches_oracle(torch_gpu, oracle):
    mu = _mu()
    n, kappa = 256, 1e6
    got = mu.to_host(mu.gen_cond(n, kappa, 21))
    ref = oracle.gen_cond(n, kappa, 21)
    assert np.max(np.abs(got - ref)) <= 1e-13
    assert np.array_equal(got, got.T)  # the update keeps it exactly symmetric


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("n,t,nrhs", [(600, 2, 16), (1200, 4, 16), (1000, 4, 3)])
def test_tiled_solve(torch_gpu, oracle, kind, n, t, nrhs):
    """Factor (FP64) then forward + back substitution on the same DAG
    machinery; X equals the oracle's FP64 solution."""
    mu = _mu()
    A = oracle.gen_cond(n, 1e3, 4)
    B = oracle.gen_uniform((n, nrhs), 9, -1.0, 1.0)
    dF, dX = mu.to_device(A), mu.to_device(B)
    info, _ = mu.factorize(kind, dF, n, t)
    assert info == 0
    dag = mu.Dag().record_solve(mu.ptr(dF), n, t, kind, mu.ptr(dX), nrhs)
    ex = mu.Exec(dag, "tiled")
    ex.launch()
    torch_gpu.cuda.synchronize()
    X = mu.to_host(dX)
    ref = oracle.solve_fp64(kind, A, B)
    assert np.max(np.abs(X - ref)) <= 1e3 * n * EPS * np.max(np.abs(ref))
    assert oracle.backward_error(A, X, B) <= 10 * n * EPS


# Variants expected to converge by refinement alone (kappa times the unit
# roundoff of the update precision well below one for FP32; for BF16 the
# updates accumulate in FP32 and refinement converges at kappa = 1e3).
CONVERGES = [("fp64", 1e3, 1), ("fp32", 1e3, 10), ("fp32", 1e6, 15), ("bf16", 1e3, 30)]


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("prec,kappa,maxit", CONVERGES)
@pytest.mark.parametrize("n,t", [(2000, 4), (3000, 3)])
def test_solve_ir(torch_gpu, oracle, kind, prec, kappa, maxit, n, t):
    """Reduced-precision trailing updates + FP64 refinement reach the FP64
    backward error (<= 10 n eps) and the LAPACK solution to kappa-level
    accuracy by refinement alone: run without the FP64 fallback, so a
    refinement that stopped converging fails here.  (3000, 3) uses the
    tiles of 1000, the benchmark's tile order."""
    mu = _mu()
    nrhs = 16
    A = oracle.gen_cond(n, kappa, 7)
    B = oracle.gen_uniform((n, nrhs), 8, -1.0, 1.0)
    dA, dB = mu.to_device(A), mu.to_device(B)
    dX, res = mu.solve_ir(kind, dA, dB, n, nrhs, t, prec=prec, fallback=False)
    torch_gpu.cuda.synchronize()
    X = mu.to_host(dX)
    print(kind, prec, kappa, n, t, res)
    assert res["info"] == 0
    assert res["converged"] == 1 and res["fell_back"] == 0 and res["iters"] <= maxit
    be = oracle.backward_error(A, X, B)
    assert be <= 10 * n * EPS and res["berr"] <= 10 * n * EPS
    ref = oracle.solve_fp64(kind, A, B)
    assert np.max(np.abs(X - ref)) <= 10 * kappa * n * EPS * np.max(np.abs(ref))
    if prec != "fp64":
        assert res["berr0"] > 1e3 * EPS   # the factors really are reduced precision
    # A and B are inputs only
    assert np.array_equal(mu.to_host(dA), A) and np.array_equal(mu.to_host(dB), B)


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_solve_ir_more_than_16_rhs(torch_gpu, oracle, kind, prec):
    """20 right-hand sides: the reduced-precision solve runs them as groups of
    16 and 4 columns of the row-stored right-hand sides (leading dimension
    20); the refinement still reaches the FP64 backward error."""
    mu = _mu()
    n, t, nrhs = 2000, 4, 20
    A = oracle.gen_cond(n, 1e3, 11)
    B = oracle.gen_uniform((n, nrhs), 12, -1.0, 1.0)
    dA, dB = mu.to_device(A), mu.to_device(B)
    dX, res = mu.solve_ir(kind, dA, dB, n, nrhs, t, prec=prec, fallback=False)
    torch_gpu.cuda.synchronize()
    X = mu.to_host(dX)
    print(kind, prec, res)
    assert res["info"] == 0 and res["converged"] == 1 and res["fell_back"] == 0
    assert oracle.backward_error(A, X, B) <= 10 * n * EPS
    ref = oracle.solve_fp64(kind, A, B)
    assert np.max(np.abs(X - ref)) <= 10 * 1e3 * n * EPS * np.max(np.abs(ref))


@pytest.mark.parametrize("kind", ["lu", "chol"])
def test_solve_ir_bf16_kappa_1e6(torch_gpu, oracle, kind):
    """BF16 updates at kappa = 1e6 (kappa u_bf16 ~ 4e3): plain refinement is
    not expected to converge.  With the fallback the run must end at the
    FP64 backward error either way, and the flags must say which way:
    fell_back == 0 only with a refinement that met the stopping test."""
    mu = _mu()
    n, t, nrhs = 2000, 4, 16
    A = oracle.gen_cond(n, 1e6, 7)
    B = oracle.gen_uniform((n, nrhs), 8, -1.0, 1.0)
    dX, res = mu.solve_ir(kind, mu.to_device(A), mu.to_device(B), n, nrhs, t, prec="bf16",
                          fallback=True)
    torch_gpu.cuda.synchronize()
    X = mu.to_host(dX)
    print(kind, "bf16 kappa 1e6 outcome:", "refined" if not res["fell_back"] else "FP64 fallback",
          res)
    assert res["info"] == 0 and res["converged"] == 1
    assert oracle.backward_error(A, X, B) <= 10 * n * EPS
    if res["fell_back"] == 0:
        assert res["iters"] <= 30 and res["berr"] <= 10 * n * EPS


@pytest.mark.parametrize("fallback", [False, True])
def test_solve_ir_nan_is_never_converged(torch_gpu, oracle, fallback):
    """A NaN in B spreads to X; the column maxima carry it (no fmax drop),
    so the stopping test fails and the run reports converged == 0 (after the
    FP64 fallback too) instead of a zero backward error."""
    mu = _mu()
    n, t, nrhs = 600, 2, 4
    A = oracle.gen_cond(n, 1e3, 7)
    B = oracle.gen_uniform((n, nrhs), 8, -1.0, 1.0)
    B[17, 2] = np.nan
    dX, res = mu.solve_ir("lu", mu.to_device(A), mu.to_device(B), n, nrhs, t, prec="fp32",
                          itermax=3, fallback=fallback)
    torch_gpu.cuda.synchronize()
    assert res["converged"] == 0
    assert res["fell_back"] == (1 if fallback else 0)
    assert not np.isfinite(res["berr"])


def test_solve_ir_out_of_fp32_range_goes_to_fp64(torch_gpu, oracle):
    """||A||_inf beyond FLT_MAX: like LAPACK dsgesv, no reduced-precision
    factorization is tried; the FP64 path solves it (fell_back, no steps),
    and without the fallback the call is refused."""
    mu = _mu()
    n, t, nrhs = 600, 2, 4
    A = oracle.gen_cond(n, 1e3, 7) * 1e300
    B = oracle.gen_uniform((n, nrhs), 8, -1.0, 1.0)
    dA, dB = mu.to_device(A), mu.to_device(B)
    dX, res = mu.solve_ir("chol", dA, dB, n, nrhs, t, prec="fp32", fallback=True)
    torch_gpu.cuda.synchronize()
    assert res["fell_back"] == 1 and res["iters"] == 0 and res["converged"] == 1
    assert oracle.backward_error(A, mu.to_host(dX), B) <= 10 * n * EPS
    with pytest.raises(mu.MustardError):
        mu.solve_ir("chol", dA, dB, n, nrhs, t, prec="fp32", fallback=False)


def test_executable_outlives_its_dag(torch_gpu, oracle):
    """A reduced-precision factorization's panel copies belong to the DAG and
    to every executable built from it: destroying the DAG first leaves the
    executable working, with the same bits as before."""
    mu = _mu()
    n, t = 2000, 4
    A = oracle.gen_ref(n)
    dA = mu.to_device(A)
    dag = mu.Dag().record_lu(mu.ptr(dA), n, t, prec=mu.PREC_FP32)
    ex = mu.Exec(dag, "tiled")
    ex.launch()
    torch_gpu.cuda.synchronize()
    first = dA.clone()
    ex._dag = None
    dag.close()
    del dag
    for _ in range(2):
        dA.copy_(mu.to_device(A))
        ex.launch()
        torch_gpu.cuda.synchronize()
        assert bool((dA == first).all())
