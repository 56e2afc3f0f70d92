"""CPU tests of libmustard_hip.so: exports, DAG recording, the scheduler
protocol (host emulation) and error paths.  No GPU is touched."""
import ctypes
import os
import re

import pytest

import mustard_amd as mu
from mustard_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    with open(os.path.join(ROOT, "include", "mustard_hip.h")) as f:
        decl = set(re.findall(r"\b(mu_[a-z0-9_]+)\s*\(", f.read()))
    L = mu.lib()
    missing = [s for s in decl if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding covers every declared entry point
    assert decl <= set(_lib.SIGNATURES), decl - set(_lib.SIGNATURES)


def _fake(addr=1 << 20):
    return ctypes.c_void_p(addr)


@pytest.mark.parametrize("kind", ["lu", "chol"])
@pytest.mark.parametrize("T", [1, 2, 3, 4, 5, 6, 9])
def test_recorded_deps_equal_oracle(oracle, kind, T):
    d = mu.Dag()
    if kind == "lu":
        d.record_lu(_fake(), 60 * T, T)
    else:
        d.record_cholesky(_fake(), 60 * T, T)
    ops, deps = oracle.task_dag(kind, T)
    assert d.num_tasks() == len(ops)
    assert [d.deps(t) for t in range(len(ops))] == deps
    names = {1: "GEMM", 2: "SYRK", 3: "TRSM", 4: "GETRF", 5: "POTRF"}
    for t, (op, w, _) in enumerate(ops):
        r, c, o, n = ctypes.c_int(), ctypes.c_int(), (ctypes.c_int * 4)(), ctypes.c_int()
        mu.call("mu_dag_task_info", d.h, t, ctypes.byref(r), ctypes.byref(c), o, 4,
                ctypes.byref(n))
        assert (r.value, c.value) == w and n.value == 1
        assert names[o[0]] == op.split("_")[0]


def test_recorded_counts_at_24(oracle):
    for kind, tasks, edges in (("lu", 4900, 13524), ("chol", 2600, 6900)):
        d = mu.Dag()
        (d.record_lu if kind == "lu" else d.record_cholesky)(_fake(), 24000, 24)
        assert d.num_tasks() == tasks
        assert sum(len(d.deps(t)) for t in range(tasks)) == edges


def _emulate(d, threads, timeout=20.0, bump=-1):
    r = (ctypes.c_longlong * 6)()
    mu.call("mu_dag_emulate", d.h, threads, timeout, bump, r)
    return list(r)


@pytest.mark.parametrize("n,t,kind,threads", [
    (600, 2, "lu", 1), (600, 2, "lu", 7), (600, 6, "chol", 48), (3000, 3, "lu", 7),
    (3000, 3, "lu", 48), (4000, 2, "lu", 7), (6000, 6, "chol", 7), (6000, 6, "lu", 48)])
def test_scheduler_protocol_emulated(n, t, kind, threads):
    """Every item runs exactly once, after all items of its predecessor units,
    and the run terminates: the kernel's own worker loop (mu_sched.hpp), run
    with four host threads per emulated workgroup and a barrier each."""
    d = mu.Dag()
    (d.record_lu if kind == "lu" else d.record_cholesky)(_fake(), n, t)
    executed, dup, viol, err, done, total = _emulate(d, threads)
    assert (executed, dup, viol, err, done) == (total, 0, 0, 0, total)


def _emulate_ex(d, workgroups, timeout=20.0, deadline=0.0, bump=-1, delay_us=0, sm=0):
    o = _lib.EmuOpts(workgroups, timeout, deadline, bump, delay_us, sm)
    r = (ctypes.c_longlong * 9)()
    mu.call("mu_dag_emulate_ex", d.h, ctypes.byref(o), r)
    return list(r)


def test_scheduler_survives_stalled_producers():
    """A producer that stalls between reserving ring positions and storing
    the item ids: the consumers that claimed those positions wait for them
    (they never give a claimed position up), and every item runs once."""
    d = mu.Dag().record_lu(_fake(), 3000, 3)
    r = _emulate_ex(d, 12, delay_us=300)
    executed, dup, viol, err, done, total = r[:6]
    assert (executed, dup, viol, err, done) == (total, 0, 0, 0, total)


@pytest.mark.parametrize("sm", [5, 20, 50, 100])
def test_scheduler_cu_budget(sm):
    """--sm (FLAGS_OCCUP): at most ceil(sm% of the workgroups) run one task's
    items at once, and tasks start only while their budgets sum to less than
    100, so sm=100 runs one task at a time."""
    d = mu.Dag().record_cholesky(_fake(), 2000, 4)
    r = _emulate_ex(d, 16, sm=sm)
    executed, dup, viol, err, done, total, most, cap, tasks = r
    assert (executed, dup, viol, err, done) == (total, 0, 0, 0, total)
    assert cap == -(-sm * 16 // 100) and most <= cap
    assert tasks <= (100 + sm - 1) // sm
    if sm == 100:
        assert tasks == 1


def test_scheduler_host_watchdog_ends_a_stalled_run():
    """A dependency that never resolves and a device time-out longer than the
    host deadline: the host sets the abort word and every worker leaves."""
    d = mu.Dag().record_lu(_fake(), 600, 2)
    r = _emulate_ex(d, 4, timeout=600.0, deadline=0.3, bump=7)
    executed, dup, viol, err, done, total = r[:6]
    assert err == 4 and done < total and dup == 0 and viol == 0


def test_scheduler_watchdog_detects_deadlock():
    d = mu.Dag().record_lu(_fake(), 600, 2)
    executed, dup, viol, err, done, total = _emulate(d, 4, timeout=0.2, bump=7)
    assert err == 1 and done < total and dup == 0 and viol == 0


def test_dag_error_paths():
    d = mu.Dag()
    with pytest.raises(mu.MustardError) as e:
        d.end_task()
    assert e.value.code == _lib.MU_ESTATE
    d.begin_task((0, 0), [(0, 0)])
    with pytest.raises(mu.MustardError) as e:
        d.begin_task((0, 1), [])
    assert e.value.code == _lib.MU_ESTATE
    d.end_task()
    d.begin_task((1, 1), [(0, 0)])
    d.end_task()
    assert d.deps(1) == [0]
    with pytest.raises(mu.MustardError):
        d.add_edge(1, 0)  # edges run from earlier to later tasks
    with pytest.raises(mu.MustardError):
        mu.call("mu_record_tiled_lu", d.h, _fake(), 600, 7, 600, None, 0)


def test_tile_op_argument_errors():
    s = ctypes.c_void_p(0)
    L = mu.lib()
    assert L.mu_dgemm(s, b"X", b"N", 2, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2) == _lib.MU_EINVAL
    assert L.mu_dgemm(s, b"N", b"N", 4, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 4) == _lib.MU_EINVAL
    assert L.mu_dtrsm(s, b"Q", b"L", b"N", b"U", 2, 2, 1.0, None, 2, None, 2) == _lib.MU_EINVAL
    assert L.mu_dsyrk(s, b"U", b"N", 2, 2, 1.0, None, 2, 0.0, None, 2) == _lib.MU_ENOTSUP
    assert L.mu_dpotrf(s, b"U", 2, None, 2, None) == _lib.MU_ENOTSUP
    assert L.mu_dgetrf_nopiv(s, 4, None, 2, None) == _lib.MU_EINVAL
    assert b"leading" in L.mu_last_error() or len(L.mu_last_error()) > 0


def test_capture_records_instead_of_running():
    """Ops issued on a stream that a DAG is recording are added to its current
    task (nothing runs, so fake pointers are fine here)."""
    d = mu.Dag()
    key = ctypes.c_void_p(12345)
    d.capture(key)
    d.begin_task((0, 0), [(0, 0)])
    mu.call("mu_dgetrf_nopiv", key, 100, _fake(), 100, None)
    mu.call("mu_dgemm", key, b"N", b"T", 8, 8, 8, -1.0, _fake(), 8, _fake(), 8, 1.0, _fake(), 8)
    d.end_task()
    d.end_capture(key)
    n = ctypes.c_int()
    ops = (ctypes.c_int * 4)()
    mu.call("mu_dag_task_info", d.h, 0, None, None, ops, 4, ctypes.byref(n))
    assert n.value == 2 and list(ops[:2]) == [4, 1]
    with pytest.raises(mu.MustardError):
        d.end_capture(key)


def _record_lu_tasks(d, key, A, n, t, nested=None):
    """The CLI's tiled LU task loop (lu_mustard.cu:343-535) recorded through
    the tile operations on `key`; for the diagonal tiles k in `nested` the
    GETRF task records no operation and the caller inserts a sub-DAG."""
    b = n // t
    esz = 8

    def blk(i, j):
        return ctypes.c_void_p(A.value + (i * b + j * b * n) * esz)

    d.capture(key)
    for k in range(t):
        d.begin_task((k, k), [(k, k)])
        if not (nested and k in nested):
            mu.call("mu_dgetrf_nopiv", key, b, blk(k, k), n, None)
        d.end_task()
        if nested and k in nested:
            yield k
        for i in range(k + 1, t):
            d.begin_task((k, i), [(k, k), (k, i)])
            mu.call("mu_dtrsm", key, b"L", b"L", b"N", b"U", b, b, 1.0, blk(k, k), n, blk(k, i), n)
            d.end_task()
        for i in range(k + 1, t):
            d.begin_task((i, k), [(k, k), (i, k)])
            mu.call("mu_dtrsm", key, b"R", b"U", b"N", b"N", b, b, 1.0, blk(k, k), n, blk(i, k), n)
            d.end_task()
            for j in range(k + 1, t):
                d.begin_task((i, j), [(i, k), (k, j), (i, j)])
                mu.call("mu_dgemm", key, b"N", b"N", b, b, b, -1.0, blk(i, k), n, blk(k, j), n, 1.0,
                        blk(i, j), n)
                d.end_task()
    d.end_capture(key)


def test_insert_subdag_lowers_the_child_inside_the_task():
    """insertSubgraph (mustard.h:214-248): the child's tasks run as a nested
    sub-DAG of the last ended task.  The lowered DAG holds exactly the
    parent's units plus the child's (item counts add up), and the scheduler
    protocol runs it with every item once and in dependency order."""
    n, t, subt = 9000, 2, 2
    A = _fake()
    key = ctypes.c_void_p(777)
    parent = mu.Dag()
    children = []
    for k in _record_lu_tasks(parent, key, A, n, t, nested={0, 1}):
        child = mu.Dag()
        b = n // t
        child.record_lu(ctypes.c_void_p(A.value + (k * b + k * b * n) * 8), b, subt, lda=n)
        parent.insert_subdag(child)
        children.append(child)
    bare = mu.Dag()
    for _ in _record_lu_tasks(bare, ctypes.c_void_p(778), A, n, t, nested={0, 1}):
        pass
    r_par = _emulate_ex(parent, 16)
    r_bare = _emulate_ex(bare, 16)
    r_child = [_emulate_ex(c, 16) for c in children]
    executed, dup, viol, err, done, total = r_par[:6]
    assert (executed, dup, viol, err, done) == (total, 0, 0, 0, total)
    assert total == r_bare[5] + sum(r[5] for r in r_child)
    # the inserted child may be destroyed afterwards (the parent holds a copy)
    for c in children:
        c.close()
    assert _emulate_ex(parent, 4)[5] == total


def test_insert_subdag_errors_and_nested_recording():
    parent, child = mu.Dag(), mu.Dag()
    with pytest.raises(mu.MustardError) as e:
        parent.insert_subdag(child)  # no task ended yet
    assert e.value.code == _lib.MU_ESTATE
    key = ctypes.c_void_p(4321)
    parent.capture(key)
    parent.begin_task((0, 0), [(0, 0)])
    parent.end_task()
    # a sub-graph recorded on the same stream takes it, then hands it back
    child.capture(key)
    child.begin_task((0, 0), [(0, 0)])
    mu.call("mu_dgetrf_nopiv", key, 100, _fake(), 100, None)
    child.end_task()
    with pytest.raises(mu.MustardError):
        parent.end_capture(key)  # the inner recording ends first
    child.end_capture(key)
    parent.begin_task((0, 1), [(0, 0)])
    mu.call("mu_dgemm", key, b"N", b"N", 8, 8, 8, -1.0, _fake(), 8, _fake(), 8, 1.0, _fake(), 8)
    parent.end_task()
    parent.end_capture(key)
    n = ctypes.c_int()
    ops = (ctypes.c_int * 4)()
    mu.call("mu_dag_task_info", parent.h, 1, None, None, ops, 4, ctypes.byref(n))
    assert n.value == 1 and ops[0] == 1
    mu.call("mu_dag_task_info", child.h, 0, None, None, ops, 4, ctypes.byref(n))
    assert n.value == 1 and ops[0] == 4
    child.begin_task((1, 1), [])
    with pytest.raises(mu.MustardError):
        parent.insert_subdag(child)  # the child is still recording a task
    child.end_task()
    parent.insert_subdag(child)
    with pytest.raises(mu.MustardError):
        parent.insert_subdag(parent)


@pytest.mark.parametrize("kind", ["lu", "chol"])
def test_whole_matrix_task_nests_like_the_tiled_dag(kind):
    """A factorization of order above 2 x 1000 nests in two levels: blocks of
    about 1000 whose diagonal operations nest again.  One whole-matrix task of
    N = 4500 therefore lowers to exactly the units of the T = 5 tiled DAG
    (blocks of 900), not to (4500 / 128)^2 sub-tiles, and the scheduler
    protocol runs it."""
    whole, tiled = mu.Dag(), mu.Dag()
    rec = "record_lu" if kind == "lu" else "record_cholesky"
    getattr(whole, rec)(_fake(), 4500, 1)
    getattr(tiled, rec)(_fake(), 4500, 5)
    r_w, r_t = _emulate_ex(whole, 16), _emulate_ex(tiled, 16)
    assert r_w[5] == r_t[5]
    executed, dup, viol, err, done, total = r_w[:6]
    assert (executed, dup, viol, err, done) == (total, 0, 0, 0, total)
