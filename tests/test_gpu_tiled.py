"""Whole tiled factorizations on the GPU against the oracle.

Each case records the reference's task sequence (lu_mustard.cu:343-535,
cholesky_mustard.cu:221-400) through the C ABI and runs it in every execution
mode.  Checks: elementwise agreement with the host LAPACK / oracle factors
(1e-10 relative, floored at 1e-3 of the largest entry), the scaled residual
||A - LU||_F / (||A||_F N eps) <= 10 computed on the GPU, and for N = 600 the
reference's own verify line (sum |A - LU| <= 1e-6, include/verify.h:70).
"""
import numpy as np
import pytest
import scipy.linalg as sl

from tests.util import close_rel, spd

pytestmark = pytest.mark.gpu

MODES = ["tiled", "stream", "subgraph"]


def _mu():
    import mustard_amd as mu
    return mu


def _run(torch, kind, A, t, mode, **kw):
    mu = _mu()
    n = A.shape[0]
    dA = mu.to_device(A)
    info, stats = mu.factorize(kind, dA, n, t, mode=mode, **kw)
    F = mu.to_host(dA)
    res = mu.scaled_residual(mu.to_device(A), dA, n, chol=(kind == "chol"))
    torch.cuda.synchronize()
    return F, info, res, stats


@pytest.mark.parametrize("mode", MODES)
def test_config0_lu_600_t2(torch_gpu, oracle, mode):
    """N=600 split in 2x2 tiles of 300, reference generator, --verify."""
    mu = _mu()
    A = mu.gen_ref_host(600)
    assert np.array_equal(A, oracle.gen_ref(600))
    F, info, res, _ = _run(torch_gpu, "lu", A, 2, mode)
    assert info == 0
    lu, piv = sl.lu_factor(A)
    assert np.array_equal(piv, np.arange(600))
    assert close_rel(F, lu) <= 1.0
    assert res <= 10.0
    assert mu.lib().mu_verify_lu_l1_host(A.ctypes.data, F.copy(order="F").ctypes.data,
                                        600) <= 1e-6


@pytest.mark.parametrize("mode", MODES)
def test_cholesky_600_t2(torch_gpu, oracle, mode):
    mu = _mu()
    A = mu.gen_ref_host(600)
    F, info, res, _ = _run(torch_gpu, "chol", A, 2, mode)
    assert info == 0
    L = sl.cholesky(A, lower=True)
    lo = np.tril(np.ones(A.shape, dtype=bool))
    assert close_rel(F[lo], L[lo]) <= 1.0
    assert np.array_equal(F[~lo], A[~lo])  # upper triangle untouched
    assert res <= 10.0
    Fc = np.asfortranarray(F)
    assert mu.lib().mu_verify_chol_l1_host(A.ctypes.data, Fc.ctypes.data, 600) <= 1e-6


CASES = [
    # n, t -- B = n / t; covers B = 1000 (nested GETRF/POTRF), ragged sub-tiles,
    # B > 1024 (nested TRSM), T = 1 and B = 1, and a whole-matrix task nested in
    # two levels (4500 -> 5 blocks of 900 -> 8 sub-tiles each)
    (3000, 3), (3000, 6), (3000, 12), (1000, 8), (130, 1), (8, 8), (4000, 2), (2100, 1),
    (4500, 1),
]


@pytest.mark.parametrize("n,t", CASES)
@pytest.mark.parametrize("kind", ["lu", "chol"])
def test_tiled_vs_oracle(torch_gpu, oracle, kind, n, t):
    rng = np.random.default_rng(n * 10 + t)
    A = oracle.gen_ref(n) if kind == "lu" else spd(rng, n)
    ref = A.copy(order="F")
    if kind == "lu":
        assert oracle.tiled_lu(ref, t) == 0
    else:
        assert oracle.tiled_cholesky(ref, t) == 0
    for mode in (["tiled", "stream", "subgraph"] if n <= 3000 else ["tiled", "subgraph"]):
        F, info, res, _ = _run(torch_gpu, kind, A, t, mode)
        assert info == 0
        if kind == "lu":
            assert close_rel(F, ref) <= 1.0, mode
        else:
            lo = np.tril(np.ones(A.shape, dtype=bool))
            assert close_rel(F[lo], ref[lo]) <= 1.0, mode
        assert res <= 10.0, mode


@pytest.mark.parametrize("grid", [1, 3, 64])
def test_persistent_small_grids(torch_gpu, oracle, grid):
    """The scheduler is correct with any number of workgroups, down to one."""
    A = oracle.gen_ref(1000)
    ref = A.copy(order="F")
    oracle.tiled_lu(ref, 4)
    F, info, res, st = _run(torch_gpu, "lu", A, 4, "subgraph", grid=grid)
    assert st["grid"] == grid
    assert info == 0 and res <= 10.0
    assert close_rel(F, ref) <= 1.0


def test_persistent_trace(torch_gpu, oracle):
    mu = _mu()
    n, t = 1200, 4
    A = oracle.gen_ref(n)
    dA = mu.to_device(A)
    info = mu.info_cell()
    dag = mu.Dag().record_lu(mu.ptr(dA), n, t, None, mu.ptr(info))
    ex = mu.Exec(dag, "subgraph", trace=True)
    ex.launch()
    torch_gpu.cuda.synchronize()
    ex.check()
    tr = ex.trace()
    assert len(tr) == ex.stats()["items"]
    assert all(e >= s > 0 for s, e, *_ in tr)
    # an item starts only after every item of every predecessor task ended
    first = {}
    last = {}
    for s, e, _, task, _ in tr:
        first[task] = min(first.get(task, s), s)
        last[task] = max(last.get(task, e), e)
    for task in range(dag.num_tasks()):
        for p in dag.deps(task):
            assert last[p] <= first[task]


def test_persistent_timeout_aborts_cleanly(torch_gpu, oracle):
    """A zero time-out makes idle workgroups abort; every workgroup exits and
    the run reports MU_ERUN instead of hanging."""
    mu = _mu()
    A = oracle.gen_ref(600)
    dA = mu.to_device(A)
    dag = mu.Dag().record_lu(mu.ptr(dA), 600, 6)
    ex = mu.Exec(dag, "subgraph", grid=8, timeout_s=1e-9)
    ex.launch()
    torch_gpu.cuda.synchronize()
    with pytest.raises(mu.MustardError) as e:
        ex.check()
    assert e.value.code == -4


@pytest.mark.parametrize("sm", [100, 20, 5])
def test_persistent_cu_budget(torch_gpu, oracle, sm):
    """--sm: each recorded task runs on at most sm% of the workgroups, and a
    task starts only while the running tasks' budgets sum to less than 100
    (at sm=100 the tasks run one at a time); the factors are the oracle's."""
    mu = _mu()
    n, t = 2000, 4
    A = oracle.gen_ref(n)
    ref = A.copy(order="F")
    oracle.tiled_lu(ref, t)
    dA = mu.to_device(A)
    info = mu.info_cell()
    dag = mu.Dag().record_lu(mu.ptr(dA), n, t, None, mu.ptr(info))
    ex = mu.Exec(dag, "subgraph", trace=True, sm=sm)
    ex.launch()
    torch_gpu.cuda.synchronize()
    ex.check()
    assert mu.info_value(info) == 0
    assert close_rel(mu.to_host(dA), ref) <= 1.0
    cap = -(-sm * ex.stats()["grid"] // 100)
    ev = []
    for s_, e_, _, task, _ in ex.trace():
        ev += [(s_, 1, task), (e_, -1, task)]
    running, tasks = {}, 0
    for _, d, task in sorted(ev, key=lambda x: (x[0], x[1])):
        running[task] = running.get(task, 0) + d
        assert running[task] <= cap
        tasks = sum(1 for v in running.values() if v > 0)
        if sm == 100:
            assert tasks <= 1


def test_persistent_host_watchdog(torch_gpu, oracle):
    """A run whose device time-out is longer than the host deadline is ended
    by the host: the abort word in pinned memory is set after 3 x timeout_s
    + 1 s, every polling loop reads it, and the run reports MU_ERUN.  The
    condition is forced from the host: a tiny grid on a long run with a
    device time-out the run never reaches."""
    mu = _mu()
    n, t = 8000, 8
    dA = mu.to_device(oracle.gen_ref(n))
    dag = mu.Dag().record_lu(mu.ptr(dA), n, t)
    # one workgroup: about 4.5 s of work (one CU runs some 77 GFLOP/s), four
    # times the host deadline of 3 * 1e-3 + 1 s
    ex = mu.Exec(dag, "subgraph", grid=1, timeout_s=1e-3)
    ex.launch()
    torch_gpu.cuda.synchronize()
    with pytest.raises(mu.MustardError) as e:
        ex.check()
    assert e.value.code == -4


def test_repeated_launches_identical(torch_gpu, oracle):
    """Re-running the same executable gives bit-identical factors."""
    mu = _mu()
    n, t = 2000, 4
    A = oracle.gen_ref(n)
    dA = mu.to_device(A)
    orig = dA.clone()
    dag = mu.Dag().record_lu(mu.ptr(dA), n, t)
    ex = mu.Exec(dag, "tiled")
    outs = []
    for _ in range(3):
        dA.copy_(orig)
        ex.launch()
        torch_gpu.cuda.synchronize()
        ex.check()
        outs.append(dA.clone())
    assert all(bool((o == outs[0]).all()) for o in outs)


def test_level_batching_stats(torch_gpu, oracle):
    """Tiled mode runs one kernel per DAG level; LU with T tiles of size
    <= 128 has 3 levels per step except the last (GETRF, TRSMs, GEMMs)."""
    mu = _mu()
    n, t = 1024, 8
    dA = mu.to_device(oracle.gen_ref(n))
    dag = mu.Dag().record_lu(mu.ptr(dA), n, t)
    ex = mu.Exec(dag, "tiled", lookahead=False)
    st = ex.stats()
    assert st["levels"] == st["depth"] == 3 * (t - 1) + 1
    # with lookahead (the default) the bulk updates of steps 0 .. t-3 (those
    # not feeding the next panel) get nodes of their own on the second chain
    la = mu.Exec(dag, "tiled").stats()
    assert la["depth"] == st["depth"] and la["levels"] == st["depth"] + (t - 2)


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("lookahead", [True, False])
def test_lookahead_graph_matches_oracle(torch_gpu, oracle, kind, lookahead):
    """The two-chain (lookahead) graph and the single chain give the oracle's
    factors, and the same bits as each other."""
    mu = _mu()
    n, t = 2000, 8
    A = oracle.gen_ref(n)
    ref = A.copy(order="F")
    assert (oracle.tiled_lu if kind == "lu" else oracle.tiled_cholesky)(ref, t) == 0
    dA = mu.to_device(A)
    info, st = mu.factorize(kind, dA, n, t, lookahead=lookahead)
    F = mu.to_host(dA)
    m = np.tril(np.ones((n, n), dtype=bool)) if kind == "chol" else np.ones((n, n), dtype=bool)
    err = np.max(np.abs(F[m] - ref[m]) / np.maximum(np.abs(ref[m]), 1e-3 * np.abs(ref).max()))
    assert info == 0 and err <= 1e-10


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("trsm_max,kernel", [(0, "k_batch_trsm_256"), (512, "k_batch_trsm_512"),
                                             (1024, "k_batch_trsm")])
def test_trsm_nesting_threshold(torch_gpu, oracle, kind, trsm_max, kernel):
    """Panel solves of 1000-tiles nest into blocks of at most trsm_max (256 by
    default) and run on the kernel sized for that block; every threshold gives
    the oracle's factors."""
    mu = _mu()
    n, t = 3000, 3
    A = oracle.gen_ref(n)
    ref = A.copy(order="F")
    assert (oracle.tiled_lu if kind == "lu" else oracle.tiled_cholesky)(ref, t) == 0
    dA = mu.to_device(A)
    dag = mu.Dag()
    (dag.record_lu if kind == "lu" else dag.record_cholesky)(mu.ptr(dA), n, t)
    ex = mu.Exec(dag, "tiled", trsm_max=trsm_max)
    kernels = {ex.level_info(lv)[0] for lv in range(ex.stats()["levels"])}
    assert kernel in kernels
    ex.launch()
    torch_gpu.cuda.synchronize()
    F = mu.to_host(dA)
    m = np.tril(np.ones((n, n), dtype=bool)) if kind == "chol" else np.ones((n, n), dtype=bool)
    err = np.max(np.abs(F[m] - ref[m]) / np.maximum(np.abs(ref[m]), 1e-3 * np.abs(ref).max()))
    assert err <= 1e-10


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("n,t,prec", [(3000, 6, 0), (2400, 4, 0), (3600, 6, 1), (1000, 1, 0)])
def test_merged_update_units_are_bit_identical(torch_gpu, oracle, kind, n, t, prec):
    """The tiled graph merges the GEMM/SYRK units of a node that tile one
    rectangle (LU) or lower triangle (Cholesky) of C into one unit over the
    whole region.  Each element is accumulated in the same order either way,
    so the factors must be bit-identical with and without merging; the
    merged graph runs fewer GEMM items (no ragged blocks at tile edges)."""
    mu = _mu()
    rng = np.random.default_rng(n + t)
    A = oracle.gen_ref(n) if kind == "lu" else spd(rng, n)
    out = {}
    for fuse in (False, True):
        dA = mu.to_device(A)
        info = mu.info_cell(dA.device)
        dag = mu.Dag()
        (dag.record_lu if kind == "lu" else dag.record_cholesky)(mu.ptr(dA), n, t, None,
                                                                 mu.ptr(info), prec)
        ex = mu.Exec(dag, "tiled", fuse=fuse)
        ex.launch()
        torch_gpu.cuda.synchronize()
        assert mu.info_value(info) == 0
        gemm_items = sum(ex.level_items(lv) for lv in range(ex.stats()["levels"])
                         if ex.level_info(lv)[0] == "k_batch_gemm")
        out[fuse] = (mu.to_host(dA).copy(), gemm_items)
    assert np.array_equal(out[False][0], out[True][0])
    if n // t == 600:
        # 600-wide tiles are 10 blocks of 64 each, a merged 2 x 2 region 19
        assert out[True][1] < out[False][1]
