"""Reduced-precision trailing updates: the level-batched graph runs the
updates that read the converted panels on k_batch_gemm_low (both operands
k-contiguous, long k-slices, 128 x 128 blocks), while the stream mode runs
the same units on the general kernel (gemm_block_low).  In FP32 both
accumulate each element over the same k order (4-wide MFMA steps,
ascending), so the factors must agree bit for bit; in BF16 the graph's
kernel sums 32 products per MFMA (v_mfma_f32_16x16x32_bf16) against the
general kernel's 16, so the two agree to the FP32 accumulation's rounding.
Tile orders 1000 and 512 qualify for the k-contiguous kernel; 500 does not
for BF16 (500 is not a multiple of 8) and takes the general kernel in both
modes."""
import numpy as np
import pytest

from tests.util import spd

pytestmark = pytest.mark.gpu


def _factor(mu, torch, kind, A, t, prec, mode):
    n = A.shape[0]
    dA = mu.to_device(A)
    info = mu.info_cell(dA.device)
    d = mu.Dag()
    (d.record_lu if kind == "lu" else d.record_cholesky)(mu.ptr(dA), n, t, None, mu.ptr(info),
                                                         prec)
    ex = mu.Exec(d, mode)
    ex.launch()
    torch.cuda.synchronize()
    ex.check()
    kernels = {ex.level_info(lv)[0] for lv in range(ex.stats()["levels"])} if mode == "tiled" \
        else set()
    assert mu.info_value(info) == 0
    return mu.to_host(dA).copy(), kernels


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("n,t,prec", [(3000, 3, 2), (3000, 3, 1), (2048, 4, 2), (2000, 4, 2)])
def test_low_precision_kernels_agree(torch_gpu, oracle, kind, n, t, prec):
    import mustard_amd as mu
    rng = np.random.default_rng(n + 7 * t + prec)
    A = oracle.gen_ref(n) if kind == "lu" else spd(rng, n)
    print("low-precision", kind, n, t, prec, flush=True)
    Fg, kern = _factor(mu, torch_gpu, kind, A, t, prec, "tiled")
    Fs, _ = _factor(mu, torch_gpu, kind, A, t, prec, "stream")
    b = n // t
    ve = 8 if prec == 2 else 4
    assert (("k_batch_gemm_low" in kern) == (b % ve == 0)), kern
    if prec == 1 or "k_batch_gemm_low" not in kern:
        assert np.array_equal(Fg, Fs)
    else:
        # the two BF16 kernels round the FP32 sums at different points: a
        # few FP32 ulps of the updates, grown through the later steps
        lo = np.tril(np.ones(A.shape, dtype=bool)) if kind == "chol" else np.ones(A.shape, bool)
        assert np.max(np.abs(Fg - Fs)[lo]) <= 1e-4 * np.max(np.abs(Fs[lo]))
    # and the factors are those of A to the precision of the updates
    ref = A.copy(order="F")
    (oracle.tiled_lu if kind == "lu" else oracle.tiled_cholesky)(ref, t)
    if kind == "chol":
        lo = np.tril(np.ones(A.shape, dtype=bool))
        Fg, ref = Fg[lo], ref[lo]
    tol = 2e-2 if prec == 2 else 1e-5
    assert np.max(np.abs(Fg - ref)) <= tol * np.max(np.abs(ref))
