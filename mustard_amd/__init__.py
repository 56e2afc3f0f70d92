"""mustard_amd -- MI355X-native tiled LU / Cholesky (host-side Python API).

Python mirror of the reference's single-GPU path: a task DAG recorded with
the TiledGraphCreator rule (mustard/mustard.h:132-384) in the order of
tiledLU / tiledCholesky (mustard/lu_mustard.cu:343-535,
mustard/cholesky_mustard.cu:221-400), executed as one hipGraph of
level-batched kernels (``--tiled``) or as one launch of the persistent device
scheduler (``--subgraph``).

Device memory and streams come from PyTorch-ROCm; all arithmetic runs in
``libmustard_hip.so`` (hand-written gfx950 kernels).  Column-major matrices are
held in torch tensors of the transposed shape: an m x n column-major matrix is
a contiguous tensor of shape (n, m).
"""
from __future__ import annotations

import ctypes

from ._lib import (MODE_STREAM, MODE_SUBGRAPH, MODE_TILED, PREC_BF16, PREC_FP32, PREC_FP64,
                   ExecOpts, ExecStats, MustardError, call, lib)

__all__ = [
    "lib", "call", "MustardError", "Dag", "Exec", "factorize", "MODE_TILED", "MODE_SUBGRAPH",
    "MODE_STREAM", "PREC_FP64", "PREC_FP32", "PREC_BF16", "to_device", "to_host", "ptr",
    "stream_handle", "dgemm", "dsyrk", "dtrsm", "dgetrf_nopiv", "dpotrf", "scaled_residual",
    "gen_spd", "gen_ref_host", "MODES",
]

MODES = {"tiled": MODE_TILED, "subgraph": MODE_SUBGRAPH, "stream": MODE_STREAM}


def ptr(t, offset_elems: int = 0) -> ctypes.c_void_p:
    """Device address of a tensor (plus an element offset)."""
    return ctypes.c_void_p(t.data_ptr() + offset_elems * t.element_size())


def stream_handle(stream=None) -> ctypes.c_void_p:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def to_device(A, device="cuda"):
    """Fortran-ordered (or any) float64 numpy matrix -> column-major device tensor."""
    import numpy as np
    import torch
    A = np.asarray(A, dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(A.T)).to(device)


def to_host(T):
    """Column-major device tensor -> Fortran-ordered numpy matrix."""
    return T.detach().cpu().numpy().T


def gen_ref_host(n: int):
    """include/gen.h:23-38 through the library (glibc rand, seed 420)."""
    import numpy as np
    A = np.empty((n, n), dtype=np.float64, order="F")
    lib().mu_gen_ref_host(A.ctypes.data_as(ctypes.c_void_p), n)
    return A


# --------------------------------------------------------------------------
# tile operations (eager on a stream)
# --------------------------------------------------------------------------
def _b(c):
    return c.encode() if isinstance(c, str) else c


def dgemm(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
    call("mu_dgemm", stream_handle(stream), _b(transa), _b(transb), m, n, k, alpha, A, lda, B,
         ldb, beta, C, ldc)


PREC = {"fp64": 0, "fp32": 1, "bf16": 2}


def set_update_precision(prec, stream=None):
    """Precision of the GEMM/SYRK calls issued next on `stream` ("fp64",
    "fp32": FP32 operands and accumulation, "bf16": BF16 operands, FP32
    accumulation); C stays FP64."""
    call("mu_set_update_precision", stream_handle(stream), PREC[prec] if isinstance(prec, str) else prec)


def dsyrk(uplo, trans, n, k, alpha, A, lda, beta, C, ldc, stream=None):
    call("mu_dsyrk", stream_handle(stream), _b(uplo), _b(trans), n, k, alpha, A, lda, beta, C,
         ldc)


def dtrsm(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, stream=None):
    call("mu_dtrsm", stream_handle(stream), _b(side), _b(uplo), _b(transa), _b(diag), m, n,
         alpha, A, lda, B, ldb)


def dgetrf_nopiv(n, A, lda, info=None, stream=None):
    call("mu_dgetrf_nopiv", stream_handle(stream), n, A, lda, info)


def dpotrf(uplo, n, A, lda, info=None, stream=None):
    call("mu_dpotrf", stream_handle(stream), _b(uplo), n, A, lda, info)


# --------------------------------------------------------------------------
# DAGs and executables
# --------------------------------------------------------------------------
class Dag:
    """A recorded task DAG (mustard::TiledGraphCreator)."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        call("mu_dag_create", ctypes.byref(self.h))

    def close(self):
        if self.h:
            lib().mu_dag_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def record_lu(self, A, n, t, lda=None, info=None, prec=PREC_FP64):
        call("mu_record_tiled_lu", self.h, A, n, t, lda or n, info, prec)
        return self

    def record_cholesky(self, A, n, t, lda=None, info=None, prec=PREC_FP64):
        call("mu_record_tiled_cholesky", self.h, A, n, t, lda or n, info, prec)
        return self

    def record_solve(self, F, n, t, kind, X, nrhs, ldf=None, ldx=None):
        """Forward + back substitution with the factors F of a tiled LU
        (kind "lu") or Cholesky ("chol") on the right-hand sides X, in place."""
        call("mu_record_tiled_solve", self.h, F, n, t, ldf or n,
             _lib.SOLVE_LU if kind == "lu" else _lib.SOLVE_CHOL, X, ldx or n, nrhs)
        return self

    def set_tag(self, tag):
        call("mu_dag_set_tag", self.h, tag)

    def begin_task(self, w, reads):
        flat = (ctypes.c_int * (2 * len(reads)))(*[x for rc in reads for x in rc])
        t = ctypes.c_int()
        call("mu_dag_begin_task", self.h, w[0], w[1], flat, len(reads), ctypes.byref(t))
        return t.value

    def end_task(self):
        call("mu_dag_end_task", self.h)

    def capture(self, stream):
        call("mu_dag_capture_begin", self.h, stream)

    def end_capture(self, stream):
        call("mu_dag_capture_end", self.h, stream)

    def num_tasks(self):
        n = ctypes.c_int()
        call("mu_dag_num_tasks", self.h, ctypes.byref(n))
        return n.value

    def deps(self, task):
        n = ctypes.c_int()
        call("mu_dag_task_deps", self.h, task, None, 0, ctypes.byref(n))
        buf = (ctypes.c_int * max(n.value, 1))()
        call("mu_dag_task_deps", self.h, task, buf, n.value, ctypes.byref(n))
        return list(buf[:n.value])

    def add_edge(self, src, dst):
        call("mu_dag_add_edge", self.h, src, dst)

    def insert_subdag(self, child):
        """insertSubgraph (mustard.h:214-248): child's tasks run as a nested
        sub-DAG of the last task ended here, after its own operations."""
        call("mu_dag_insert_subdag", self.h, child.h)


class Exec:
    """A lowered, uploaded DAG ready to launch."""

    def __init__(self, dag: Dag, mode=MODE_TILED, grid=0, timeout_s=10.0, trace=False,
                 small_max=0, trsm_max=0, lookahead=True, sm=0, fuse=True):
        if isinstance(mode, str):
            mode = MODES[mode]
        self.mode = mode
        self.h = ctypes.c_void_p()
        # the executable points into buffers the DAG recorded; keep it alive
        self._dag = dag
        # lookahead: True / 1 (two lanes), 2 (three lanes: only the chain between
        # diagonal factorizations on the first), False (one chain)
        la = (1 if lookahead else -1) if isinstance(lookahead, bool) else int(lookahead)
        opts = ExecOpts(grid, timeout_s, int(trace), small_max, trsm_max, la, sm, 0 if fuse else 1)
        call("mu_exec_create", dag.h, mode, ctypes.byref(opts), ctypes.byref(self.h))

    def launch(self, stream=None):
        call("mu_exec_launch", self.h, stream_handle(stream))

    def check(self):
        call("mu_exec_check", self.h)

    def launch_level(self, level, stream=None):
        call("mu_exec_launch_level", self.h, level, stream_handle(stream))

    def level_ops(self, level):
        """Items of one level per operation name."""
        c = (ctypes.c_longlong * 8)()
        call("mu_exec_level_ops", self.h, level, c)
        names = ["NONE", "GEMM", "SYRK", "TRSM", "GETRF", "POTRF", "TRINV", "CVT"]
        return {names[i]: c[i] for i in range(8) if c[i]}

    def level_info(self, level):
        """(kernel name, algorithmic flops) of one level of a tiled executable."""
        k, f = ctypes.c_int(), ctypes.c_double()
        call("mu_exec_level_info", self.h, level, ctypes.byref(k), ctypes.byref(f))
        return {1: "k_batch_gemm", 2: "k_batch_small", 3: "k_batch_trsm",
                4: "k_batch_trsm_small", 5: "k_batch_trsm_256",
                6: "k_batch_trsm_512", 7: "k_batch_gemm_narrow",
                8: "k_batch_gemm_low", 9: "k_batch_cvt"}.get(k.value, "k_batch"), f.value

    def level_items(self, level):
        n = ctypes.c_longlong()
        call("mu_exec_level_items", self.h, level, ctypes.byref(n))
        return n.value

    def stats(self):
        s = ExecStats()
        call("mu_exec_stats", self.h, ctypes.byref(s))
        return {"mode": s.mode, "grid": s.grid, "units": s.units, "items": s.items,
                "edges": s.edges, "flops": s.flops, "critical_path_us": s.critical_path_us,
                "priority_units": list(s.priority_units), "wall_clock_khz": s.wall_clock_khz,
                "levels": s.levels, "depth": s.depth}

    def trace(self):
        """Per-item (start, end, unit, task, op) of the last persistent run."""
        st = self.stats()
        n = st["items"]
        buf = (ctypes.c_ulonglong * (2 * n))()
        call("mu_exec_trace", self.h, buf, 2 * n)
        iu = (ctypes.c_int * n)()
        tu = (ctypes.c_int * st["units"])()
        ou = (ctypes.c_int * st["units"])()
        call("mu_exec_item_units", self.h, iu, tu, ou)
        return [(buf[2 * i], buf[2 * i + 1], iu[i], tu[iu[i]], ou[iu[i]]) for i in range(n)]

    def close(self):
        if self.h:
            lib().mu_exec_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def info_cell(device="cuda"):
    """A device int initialised to INT_MAX (no failure)."""
    import torch
    return torch.full((1,), 2 ** 31 - 1, dtype=torch.int32, device=device)


def info_value(cell) -> int:
    v = int(cell.item())
    return 0 if v == 2 ** 31 - 1 else v


def factorize(kind, A, n, t, mode="tiled", prec=PREC_FP64, lda=None, **exec_kw):
    """Tiled LU ('lu') or Cholesky ('chol') of the column-major device matrix A
    in place.  Returns (info, stats)."""
    import torch
    info = info_cell(A.device)
    dag = Dag()
    if kind == "lu":
        dag.record_lu(ptr(A), n, t, lda, ptr(info), prec)
    else:
        dag.record_cholesky(ptr(A), n, t, lda, ptr(info), prec)
    ex = Exec(dag, mode, **exec_kw)
    ex.launch()
    torch.cuda.synchronize()
    ex.check()
    st = ex.stats()
    ex.close()
    dag.close()
    return info_value(info), st


def scaled_residual(A, F, n, chol=False, lda=None, stream=None):
    """||A - LU||_F / (||A||_F n eps) (or L L^T) computed on the device."""
    import torch
    work = torch.empty(3 * n * n, dtype=torch.float64, device=A.device)
    out = ctypes.c_double()
    call("mu_scaled_residual", ptr(A), ptr(F), n, lda or n, int(chol), ptr(work),
         ctypes.byref(out), stream_handle(stream))
    del work
    return out.value


def gen_cond(n, kappa, seed, nrefl=4, device="cuda", stream=None):
    """A = Q diag(kappa^(-i/(n-1))) Q^T on the device, Q a product of `nrefl`
    Householder reflectors: symmetric positive definite, cond_2(A) = kappa."""
    import torch
    A = torch.empty((n, n), dtype=torch.float64, device=device)
    work = torch.empty(3 * n + 2, dtype=torch.float64, device=device)
    call("mu_gen_cond", ptr(A), n, n, float(kappa), nrefl, seed, ptr(work), stream_handle(stream))
    torch.cuda.synchronize()
    del work
    return A


def solve_ir(kind, A, B, n, nrhs, tiles, prec="fp32", itermax=0, fallback=False, stream=None):
    """X solving A X = B (column-major device matrices as from to_device) by a
    tiled factorization with `prec` trailing updates and FP64 iterative
    refinement.  Returns (X, result dict)."""
    import torch
    X = torch.zeros((nrhs, n), dtype=torch.float64, device=A.device)
    o = _lib.IROpts(_lib.SOLVE_LU if kind == "lu" else _lib.SOLVE_CHOL, tiles,
                    PREC[prec] if isinstance(prec, str) else prec, itermax, 1 if fallback else 0)
    r = _lib.IRResult()
    call("mu_solve_ir", ptr(A), n, n, ptr(B), n, ptr(X), n, nrhs, ctypes.byref(o),
         ctypes.byref(r), stream_handle(stream))
    res = {f: getattr(r, f) for f, _ in _lib.IRResult._fields_ if f != "berr_hist"}
    res["berr_hist"] = list(r.berr_hist[:min(32, r.iters + 1)])
    return X, res


def gen_spd(n, seed, device="cuda", stream=None):
    """A = M M^T / n + n I on the device (M uniform [-1, 1) from `seed`)."""
    import torch
    A = torch.empty((n, n), dtype=torch.float64, device=device)
    work = torch.empty((n, n), dtype=torch.float64, device=device)
    call("mu_gen_spd", ptr(A), n, n, seed, ptr(work), stream_handle(stream))
    torch.cuda.synchronize()
    del work
    return A
