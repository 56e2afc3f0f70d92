"""GPU tests of the tiled Cholesky reverse mode (mu_chol_backward) and its
torch.autograd binding, against the oracle's restated adjoint, the closed-form
gradient of the loss sum(log diag L) + 0.5 ||L^-1 B||_F^2, and central finite
differences."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mu():
    import mustard_amd as mu
    return mu


@pytest.mark.parametrize("n,t", [(256, 1), (512, 4), (1000, 4), (4096, 16)])
def test_chol_backward_vs_oracle(torch_gpu, oracle, n, t):
    mu = _mu()
    rng = np.random.default_rng(n + t)
    A = oracle.gen_spd(n, 3)
    L = np.linalg.cholesky(A)
    Lbar = np.tril(rng.uniform(-1, 1, (n, n)))
    # the strict upper part of the factor buffer holds A (as the tiled
    # Cholesky leaves it) and must be ignored, as must Lbar's
    F = np.where(np.tril(np.ones((n, n), dtype=bool)), L, A)
    Lb_in = Lbar + np.triu(rng.uniform(-1, 1, (n, n)), 1)
    dF, dLb = mu.to_device(F), mu.to_device(Lb_in)
    dG = torch_gpu.empty_like(dF)
    mu.call("mu_chol_backward", mu.ptr(dF), n, mu.ptr(dLb), n, mu.ptr(dG), n, n, t, 0, None)
    torch_gpu.cuda.synchronize()
    G = mu.to_host(dG)
    ref = oracle.chol_backward(L, Lbar)
    assert np.max(np.abs(G - ref)) <= 1e-10 * np.max(np.abs(ref))
    assert np.array_equal(G, G.T)


def _c3(torch, mu, A, B, tiles, prec):
    Ad = A.clone().requires_grad_(True)
    L = mu.autograd.cholesky(Ad, tiles, prec)
    Y = torch.linalg.solve_triangular(L, B, upper=False)
    loss = torch.log(torch.diagonal(L)).sum() + 0.5 * (Y * Y).sum()
    loss.backward()
    return loss.item(), Ad.grad


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_autograd_c3_loss_n4096(torch_gpu, oracle, prec, tol):
    """N = 4096 in tiles of 256, B 4096 x 32: gradient vs the dense FP64
    closed form 0.5 A^-1 - 0.5 A^-1 B B^T A^-1."""
    import mustard_amd.autograd  # noqa: F401
    mu = _mu()
    torch = torch_gpu
    n, m = 4096, 32
    A = oracle.gen_spd(n, 420)
    B = oracle.gen_uniform((n, m), 7, -1.0, 1.0)
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    loss, g = _c3(torch, mu, dA, dB, 16, prec)
    ref_loss, _, _, Gc = oracle.c3_loss_grad(A, B)
    G = g.cpu().numpy()
    assert abs(loss - ref_loss) <= tol * abs(ref_loss)
    assert np.max(np.abs(G - Gc)) <= tol * np.max(np.abs(Gc))
    if prec == "fp64":
        assert np.max(np.abs(G - Gc)) <= 1e-10 * np.max(np.abs(Gc))


def test_autograd_finite_differences_n64(torch_gpu, oracle):
    """200 random symmetric directions at N = 64 (tiles of 32): <G, E> equals
    the central difference of the loss with step 1e-6."""
    import mustard_amd.autograd  # noqa: F401
    mu = _mu()
    torch = torch_gpu
    n, m = 64, 32
    rng = np.random.default_rng(11)
    A = oracle.gen_spd(n, 9)
    B = rng.uniform(-1, 1, (n, m))
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    _, g = _c3(torch, mu, dA, dB, 2, "fp64")
    G = g.cpu().numpy()
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for _ in range(200):
            E = rng.uniform(-1, 1, (n, n))
            E = E + E.T
            dE = torch.from_numpy(E).cuda()
            lp = _loss_nograd(torch, mu, dA + h * dE, dB)
            lm = _loss_nograd(torch, mu, dA - h * dE, dB)
            fd = (lp - lm) / (2 * h)
            worst = max(worst, abs(fd - float(np.sum(G * E))) / max(1.0, abs(fd)))
    assert worst <= 1e-6


def _loss_nograd(torch, mu, A, B):
    L = mu.autograd.cholesky(A, 2, "fp64")
    Y = torch.linalg.solve_triangular(L, B, upper=False)
    return (torch.log(torch.diagonal(L)).sum() + 0.5 * (Y * Y).sum()).item()
