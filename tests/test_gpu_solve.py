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


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("prec,kappa", [("fp64", 1e3), ("fp32", 1e3), ("fp32", 1e6),
                                        ("bf16", 1e3)])
def test_solve_ir(torch_gpu, oracle, kind, prec, kappa):
    """Reduced-precision trailing updates + FP64 refinement reach the FP64
    backward error (<= 10 n eps) and the LAPACK solution to kappa-level
    accuracy; the unrefined solution is visibly worse for reduced precision."""
    mu = _mu()
    n, t, nrhs = 2000, 4, 16
    A = oracle.gen_cond(n, kappa, 7)
    B = oracle.gen_uniform((n, nrhs), 8, -1.0, 1.0)
    dA, dB = mu.to_device(A), mu.to_device(B)
    dX, res = mu.solve_ir(kind, dA, dB, n, nrhs, t, prec=prec, fallback=True)
    torch_gpu.cuda.synchronize()
    X = mu.to_host(dX)
    assert res["info"] == 0
    be = oracle.backward_error(A, X, B)
    print(kind, prec, kappa, res)
    assert res["converged"] == 1
    assert be <= 10 * n * EPS and res["berr"] <= 10 * n * EPS
    ref = oracle.solve_fp64(kind, A, B)
    assert np.max(np.abs(X - ref)) <= 10 * kappa * n * EPS * np.max(np.abs(ref))
    if prec == "fp64":
        assert res["iters"] <= 1
    if prec == "fp32" and kappa <= 1e3:
        assert res["berr0"] > 1e3 * EPS   # the factors really are reduced precision
        assert res["fell_back"] == 0 and res["iters"] <= 10
    # A and B are inputs only
    assert np.array_equal(mu.to_host(dA), A) and np.array_equal(mu.to_host(dB), B)
