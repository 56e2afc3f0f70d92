"""Benchmark runner and log reader in the reference's layout.

The reference drives its binaries with scripts/lu_all.sh / chol_all.sh and
reads the logs with scripts/calc.py.  This module does both for the
lu_mustard / cholesky_mustard CLIs built here, with the same file layout and
parsing rules, so existing analyses read these runs unchanged:

  * layout (lu_all.sh:71-126): <out>/<lu|chol>/<N>/ holds
      log0_<N>.log             single-kernel mode (no --tiled/--subgraph),
      log1_<N>_<T>.log         --tiled,
      log2_<N>_<T>_<g>GPU.log  --subgraph on g GPUs (g = 1 here);
    each run appends to its log, `-r` runs per invocation, `--sm` as given;
  * reading (calc.py:128-173): the method is the integer after "log" in the
    file name; methods 0 and 1 are one GPU; every line starting with
    "device" carries a run time as its last space-separated token; each
    group of g such lines is one data point, whose rate uses the reference's
    convention 2N^3/3 flops for LU and Cholesky alike (calc.py:53-59).

Usage:
  python -m mustard_amd.runs -m lu -s 24000 -t 24 -r 11 --out logs
  python -m mustard_amd.runs --read logs -m lu -s 24000
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "bin")
METHODS = {0: "single-kernel", 1: "tiled", 2: "mustard"}
PROGRAMS = {"lu": "lu_mustard", "chol": "cholesky_mustard"}


def log_name(method: int, n: int, t: int, gpus: int = 1) -> str:
    """The reference's log file name (lu_all.sh:71-126)."""
    if method == 0:
        return f"log0_{n}.log"
    if method == 1:
        return f"log1_{n}_{t}.log"
    return f"log2_{n}_{t}_{gpus}GPU.log"


def run(kind: str, n: int, t: int, runs: int, out: str, methods=(0, 1, 2), sm: int = 100,
        bin_dir: str = BIN, extra=()) -> list:
    """Run the CLI once per method, appending its output to the method's log
    under <out>/<kind>/<n>/; returns the log paths."""
    folder = os.path.join(out, kind, str(n))
    os.makedirs(folder, exist_ok=True)
    exe = os.path.join(bin_dir, PROGRAMS[kind])
    paths = []
    for m in methods:
        args = [exe, f"-N={n}", f"-run={runs}"]
        if m >= 1:
            args += [f"-T={t}", "-tiled" if m == 1 else "-subgraph", f"-sm={sm}"]
        args += list(extra)
        path = os.path.join(folder, log_name(m, n, t))
        with open(path, "a") as f:
            p = subprocess.run(args, stdout=f, stderr=subprocess.STDOUT)
        if p.returncode != 0:
            raise RuntimeError(f"{' '.join(args)} exited with {p.returncode} (see {path})")
        paths.append(path)
    return paths


def flops_ref_convention(n: int) -> float:
    """calc.py:53-59: 2N^3/3 for LU and Cholesky alike."""
    return 2.0 * float(n * n * n) / 3.0


def read_logs(folder: str, size: int) -> list:
    """Data points (method name, gpu count, seconds, TFLOP/s) of the method
    0-2 logs in `folder`, read by calc.py's rules (readData / readMustard)."""
    points = []
    for fname in sorted(os.listdir(folder)):
        if not fname.endswith(".log"):
            continue
        parts = fname[:-4].split("_")
        method = int(parts[0][3:])
        if method >= 3:
            continue  # other libraries' logs (multi-GPU cuSOLVER, StarPU, SLATE)
        if method == 0:
            parts.append("1")
        if method <= 1:
            parts.append("1GPU")
        gpus = int(parts[3][:-3])
        seen = 0
        with open(os.path.join(folder, fname)) as f:
            for line in f:
                if not line.startswith("device"):
                    continue
                seen += 1
                if seen == gpus:
                    secs = float(line.split(" ")[-1])
                    points.append((METHODS.get(method, str(method)), gpus, secs,
                                   flops_ref_convention(size) / secs * 1e-12))
                    seen = 0
    return points


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("-m", "--method", choices=["lu", "chol"], default="lu")
    ap.add_argument("-s", "--size", type=int, default=24000)
    ap.add_argument("-t", "--tiles", type=int, default=24)
    ap.add_argument("-r", "--runs", type=int, default=11)
    ap.add_argument("--sm", type=int, default=100)
    ap.add_argument("--out", default="logs")
    ap.add_argument("--methods", default="0,1,2")
    ap.add_argument("--read", metavar="DIR", help="only read the logs under DIR")
    a = ap.parse_args(argv)
    root = a.read or a.out
    if not a.read:
        run(a.method, a.size, a.tiles, a.runs, a.out,
            tuple(int(x) for x in a.methods.split(",")), a.sm)
    pts = read_logs(os.path.join(root, a.method, str(a.size)), a.size)
    for name, g, secs, tf in pts:
        print(f"{name:14s} {g} GPU  {secs:.4f} s  {tf:.2f} TFLOP/s (2N^3/3)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
