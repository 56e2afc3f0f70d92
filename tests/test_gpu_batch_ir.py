"""Batched solves with reduced-precision factorizations and FP64 refinement
(mustard_amd.batch.BatchSolve): every problem's backward error, computed on
the device by mu_batch_residual, reaches the FP64 level; the device norms
agree with numpy's; and the persistent-scheduler launch of the batch gives
the same factors as the graph launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -53


def _problems(count, seed, lo=32, hi=300):
    rng = np.random.default_rng(seed)
    ns = [int(x) for x in rng.integers(lo, hi + 1, count)]
    kinds = ["chol" if p % 3 else "lu" for p in range(count)]
    return kinds, ns


@pytest.mark.parametrize("prec,mode", [("fp32", "tiled"), ("bf16", "tiled"),
                                       ("fp32", "subgraph")])
def test_batched_ir_reaches_fp64_backward_error(torch_gpu, oracle, prec, mode):
    from mustard_amd import autotune
    from mustard_amd.batch import BatchSolve, make_problems, make_rhs, unpack
    kinds, ns = _problems(96, 7)
    try:
        table = autotune.load()
    except OSError:
        table = {}
    tiles = [autotune.choose(k, n, table) for k, n in zip(kinds, ns)]
    A = make_problems(kinds, ns, 11)
    b = make_rhs(ns, 12)
    bs = BatchSolve(kinds, ns, tiles, prec=prec, mode=mode)
    bs.load(A, b)
    out = bs.solve(itermax=30)
    print(prec, mode, "steps", out["steps"], "max berr", np.max(out["berr"]),
          "max unrefined", np.max(out["berr0"]), "ms", out["ms"])
    assert not out["infos"].any()
    assert out["converged"] and out["steps"] <= 30
    X = bs.X.cpu().numpy()
    voff = np.concatenate([[0], np.cumsum(ns)])
    for p in range(len(ns)):
        Ap = unpack(A, ns, p)
        xp, bp = X[voff[p]:voff[p + 1]], b[voff[p]:voff[p + 1]]
        # FP64-level backward error, by numpy and by the device residual (the
        # two differ only in the rounding of the residual's sums)
        ref = oracle.backward_error(Ap, xp[:, None], bp[:, None])
        assert ref <= 10 * ns[p] * EPS, (p, ref)
        assert np.isfinite(out["berr"][p]) and out["berr"][p] <= 10 * ns[p] * EPS, p
    bs.close()


def test_batch_residual_norms_match_numpy(torch_gpu):
    import torch
    import mustard_amd as mu
    from mustard_amd.batch import make_problems, unpack
    kinds, ns = _problems(40, 3, 1, 700)
    A = make_problems(kinds, ns, 5)
    rng = np.random.default_rng(9)
    nv = sum(ns)
    x, b = rng.standard_normal(nv), rng.standard_normal(nv)
    x[17] = np.nan  # the problem holding entry 17 must report NaN norms
    aoff = np.concatenate([[0], np.cumsum([n * n for n in ns])])[:-1]
    voff = np.concatenate([[0], np.cumsum(ns)])
    d = lambda v, dt=torch.float64: torch.tensor(v, dtype=dt, device="cuda")
    dA, dx, db = d(A), d(x), d(b)
    dR, dN = torch.empty(nv, dtype=torch.float64, device="cuda"), \
        torch.empty(4 * len(ns), dtype=torch.float64, device="cuda")
    dn, da, dv = d(ns, torch.int32), d(aoff, torch.int64), d(voff[:-1], torch.int64)
    mu.call("mu_batch_residual", len(ns), mu.ptr(dn), mu.ptr(da), mu.ptr(dv), mu.ptr(dA),
            mu.ptr(dx), mu.ptr(db), mu.ptr(dR), mu.ptr(dN), None)
    torch.cuda.synchronize()
    R, N = dR.cpu().numpy(), dN.cpu().numpy().reshape(-1, 4)
    for p, n in enumerate(ns):
        Ap = unpack(A, ns, p)
        xp, bp = x[voff[p]:voff[p + 1]], b[voff[p]:voff[p + 1]]
        rp = bp - Ap @ xp
        if np.isnan(xp).any():
            assert np.isnan(N[p, 0]) and np.isnan(N[p, 1])
            continue
        assert np.allclose(R[voff[p]:voff[p + 1]], rp, rtol=1e-12, atol=1e-12 * np.abs(Ap).sum())
        assert np.isclose(N[p, 0], np.abs(rp).max(), rtol=1e-12, atol=1e-12 * np.abs(Ap).sum())
        assert N[p, 1] == np.abs(xp).max() and N[p, 3] == np.abs(bp).max()
        assert np.isclose(N[p, 2], np.abs(Ap).sum(axis=1).max(), rtol=1e-12)


def test_batch_persistent_launch_matches_graph(torch_gpu):
    from mustard_amd import autotune
    from mustard_amd.batch import Batch, make_problems
    kinds, ns = _problems(64, 21, 40, 400)
    tiles = [autotune.choose(k, n, {}) for k, n in zip(kinds, ns)]
    A = make_problems(kinds, ns, 22)
    out = {}
    for mode in ("tiled", "subgraph"):
        bt = Batch(kinds, ns, tiles, mode=mode)
        bt.load(A)
        bt.launch()
        torch_gpu.cuda.synchronize()
        bt.exec.check()
        assert not bt.infos().any()
        out[mode] = bt.buf.cpu().numpy()
        bt.close()
    # same item bodies and the same per-element order of the sums
    print("bit-identical:", np.array_equal(out["tiled"], out["subgraph"]))
    assert np.allclose(out["tiled"], out["subgraph"], rtol=1e-12, atol=0)
