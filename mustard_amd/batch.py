"""Batched uneven factorizations: thousands of independent tiled LU / Cholesky
problems of different orders recorded into one DAG (mu_record_batch) and run
by one executable -- one hipGraph launch whose level kernels mix the work
items of every problem.  Tile counts per problem come from the autotuner
(mustard_amd.autotune) or are given explicitly.

Layout: the problems are packed one after another in one device buffer, each
column-major with leading dimension equal to its order.
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
    device, with one recorded DAG and one tiled executable."""

    def __init__(self, kinds, ns, tiles, prec="fp64", device="cuda"):
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
        self.exec = Exec(self.dag, "tiled")

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
