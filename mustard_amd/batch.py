"""Batched uneven factorizations: thousands of independent tiled LU / Cholesky
problems of different orders recorded into one DAG (mu_record_batch) and run
by one executable -- one hipGraph launch whose level kernels mix the work
items of every problem.  Tile counts per problem come from the autotuner
(mustard_amd.autotune) or are given explicitly.

Layout: the problems are packed one after another in one device buffer, each
column-major with leading dimension equal to its order; right-hand sides and
solutions of BatchSolve likewise, one vector per problem.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import Dag, Exec, PREC, call, ptr, to_host
from ._lib import SOLVE_CHOL, SOLVE_LU

KINDS = {"lu": SOLVE_LU, "chol": SOLVE_CHOL}


def flops(kind, n):
    """Algorithmic flops of one factorization (LU 2n^3/3, Cholesky n^3/3)."""
    return (2.0 if kind == "lu" else 1.0) * n ** 3 / 3.0


class Batch:
    """A batch of problems (kinds[p] in {"lu", "chol"}, orders ns[p]) on the
    device, with one recorded DAG and one executable: a hipGraph of
    level-batched kernels (mode "tiled") or one persistent-scheduler launch
    (mode "subgraph")."""

    def __init__(self, kinds, ns, tiles, prec="fp64", device="cuda", mode="tiled"):
        import torch
        self.kinds, self.ns = list(kinds), [int(n) for n in ns]
        self.tiles = [int(t) for t in tiles]
        assert len(self.kinds) == len(self.ns) == len(self.tiles)
        for n, t in zip(self.ns, self.tiles):
            assert t >= 1 and n % t == 0, (n, t)
        self.off = np.concatenate([[0], np.cumsum([n * n for n in self.ns])]).astype(np.int64)
        self.buf = torch.empty(int(self.off[-1]), dtype=torch.float64, device=device)
        self.info = torch.empty(len(self.ns), dtype=torch.int32, device=device)
        np_ = len(self.ns)
        k = (ctypes.c_int * np_)(*[KINDS[x] for x in self.kinds])
        nn = (ctypes.c_int * np_)(*self.ns)
        ld = (ctypes.c_int * np_)(*self.ns)
        tt = (ctypes.c_int * np_)(*self.tiles)
        base = self.buf.data_ptr()
        ps = (ctypes.c_void_p * np_)(*[base + 8 * int(o) for o in self.off[:-1]])
        self.dag = Dag()
        call("mu_record_batch", self.dag.h, np_, k, nn, ps, ld, tt, ptr(self.info),
             PREC[prec] if isinstance(prec, str) else prec)
        self.exec = Exec(self.dag, mode)

    @property
    def flops(self):
        return sum(flops(k, n) for k, n in zip(self.kinds, self.ns))

    def view(self, p):
        """Device view of problem p as a column-major tensor (n x n stored
        transposed, as mustard_amd.to_device makes them)."""
        n = self.ns[p]
        return self.buf[int(self.off[p]):int(self.off[p + 1])].view(n, n)

    def load(self, host_buf):
        """Copy a packed host buffer (same layout) to the device."""
        import torch
        self.buf.copy_(torch.from_numpy(host_buf))

    def launch(self, stream=None):
        import torch
        with torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream()):
            self.info.fill_(2 ** 31 - 1)
        self.exec.launch(stream)

    def infos(self):
        v = self.info.cpu().numpy().astype(np.int64)
        v[v == 2 ** 31 - 1] = 0
        return v

    def factor(self, p):
        return to_host(self.view(p))

    def close(self):
        self.exec.close()
        self.dag.close()


class BatchSolve:
    """A_p x_p = b_p for every problem p of a batch, by a batched
    factorization whose trailing updates run in `prec` (FP32 or BF16; FP64
    panels and diagonal factors), followed by FP64 iterative refinement of
    all problems together: x = x + solve(b - A x), every step one launch of
    the batched solve DAG (forward and back substitution tasks of every
    problem, mu_record_tiled_solve) and one launch of the batched residual
    kernel (mu_batch_residual), until every problem meets LAPACK dsgesv's
    stopping test ||r||_inf <= ||x||_inf ||A||_inf 2^-53 sqrt(n) or `itermax`
    steps have run.  The backward error of every problem,
    ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf), comes from the device.

    `mode` runs the factorization and the solves as one hipGraph ("tiled") or
    one persistent-scheduler launch ("subgraph") each."""

    def __init__(self, kinds, ns, tiles, prec="fp32", mode="tiled", device="cuda"):
        import torch
        self.bt = Batch(kinds, ns, tiles, prec=prec, device=device, mode=mode)
        self.prec = prec
        ns_ = self.bt.ns
        self.voff = np.concatenate([[0], np.cumsum(ns_)]).astype(np.int64)
        nv = int(self.voff[-1])
        f64 = dict(dtype=torch.float64, device=device)
        self.A0 = torch.empty_like(self.bt.buf)    # the matrices (the factors overwrite bt.buf)
        self.X = torch.empty(nv, **f64)            # solution
        self.Bv = torch.empty(nv, **f64)           # right-hand sides
        self.R = torch.empty(nv, **f64)            # residual
        self.S = torch.empty(nv, **f64)            # the solve DAG's vector (in place)
        self.norms = torch.empty(4 * len(ns_), **f64)
        self.ns_dev = torch.tensor(ns_, dtype=torch.int32, device=device)
        self.aoff_dev = torch.tensor(self.bt.off[:-1], dtype=torch.int64, device=device)
        self.voff_dev = torch.tensor(self.voff[:-1], dtype=torch.int64, device=device)
        self.sdag = Dag()
        fb, sb = self.bt.buf.data_ptr(), self.S.data_ptr()
        for p, (k, n, t) in enumerate(zip(self.bt.kinds, ns_, self.bt.tiles)):
            call("mu_dag_set_tag", self.sdag.h, p + 1)
            call("mu_record_tiled_solve", self.sdag.h, ctypes.c_void_p(fb + 8 * int(self.bt.off[p])),
                 n, t, n, KINDS[k], ctypes.c_void_p(sb + 8 * int(self.voff[p])), n, 1)
        call("mu_dag_set_tag", self.sdag.h, 0)
        self.sexec = Exec(self.sdag, mode)
        self.eps_n = np.ldexp(1.0, -53) * np.sqrt(np.asarray(ns_, dtype=np.float64))

    def load(self, host_A, host_b):
        """Packed host matrices (Batch layout) and right-hand sides (packed
        vectors, problem p at voff[p])."""
        import torch
        self.A0.copy_(torch.from_numpy(host_A))
        self.Bv.copy_(torch.from_numpy(host_b))

    def _residual(self, stream=None):
        from . import stream_handle
        call("mu_batch_residual", len(self.bt.ns), ptr(self.ns_dev), ptr(self.aoff_dev),
             ptr(self.voff_dev), ptr(self.A0), ptr(self.X), ptr(self.Bv), ptr(self.R),
             ptr(self.norms), stream_handle(stream))
        nm = self.norms.view(-1, 4).cpu().numpy()
        r, x, a, b = nm[:, 0], nm[:, 1], nm[:, 2], nm[:, 3]
        with np.errstate(invalid="ignore", divide="ignore"):
            berr = np.where(a * x + b > 0, r / (a * x + b), 0.0)
        berr[~np.isfinite(r) | ~np.isfinite(x)] = np.nan
        ok = np.isfinite(r) & np.isfinite(x) & (r <= x * a * self.eps_n)
        return berr, ok

    def solve(self, itermax=30, stream=None):
        """Factor, solve and refine; returns a dict with the per-problem
        backward errors (`berr`), steps taken, whether every problem met the
        stopping test, and the device time of the whole solve in ms."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(s):
            self.bt.buf.copy_(self.A0)
            e0.record(s)
            self.bt.launch(s)
            self.S.copy_(self.Bv)
            self.sexec.launch(s)
            self.X.copy_(self.S)
            it = 0
            while True:
                berr, ok = self._residual(s)
                if it == 0:
                    berr0 = berr.copy()
                if ok.all() or it >= itermax or not np.isfinite(berr).all():
                    break
                self.S.copy_(self.R)
                self.sexec.launch(s)
                self.X.add_(self.S)
                it += 1
            e1.record(s)
        s.synchronize()
        return {"berr": berr, "berr0": berr0, "steps": it, "converged": bool(ok.all()),
                "ms": e0.elapsed_time(e1), "infos": self.bt.infos()}

    def close(self):
        self.sexec.close()
        self.sdag.close()
        self.bt.close()


def make_rhs(ns, seed):
    """Packed right-hand sides, uniform [-1, 1), numpy PCG64 stream `seed`."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, int(sum(ns)))


def make_problems(kinds, ns, seed):
    """Packed host buffer of the test problems: Cholesky problems are
    M M^T / n + n I with M uniform [-1, 1), LU problems are uniform [0, 1) with
    n added to the diagonal (diagonally dominant, no pivoting needed), all
    from numpy's PCG64 stream `seed` (column-major, leading dimension n)."""
    rng = np.random.default_rng(seed)
    off = np.concatenate([[0], np.cumsum([n * n for n in ns])]).astype(np.int64)
    buf = np.empty(int(off[-1]), dtype=np.float64)
    for p, (k, n) in enumerate(zip(kinds, ns)):
        if k == "chol":
            M = rng.uniform(-1.0, 1.0, (n, n))
            A = M @ M.T / n + n * np.eye(n)
        else:
            A = rng.uniform(0.0, 1.0, (n, n)) + n * np.eye(n)
        buf[off[p]:off[p + 1]] = A.reshape(-1, order="F")
    return buf


def unpack(buf, ns, p):
    off = int(sum(n * n for n in ns[:p]))
    n = ns[p]
    return buf[off:off + n * n].reshape((n, n), order="F")
