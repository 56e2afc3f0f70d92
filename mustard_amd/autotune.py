"""Tile-count autotuner for the tiled factorizations.

For a problem of order n the candidates are the tile counts t with n % t == 0
and tile size b = n / t in [MIN_TILE, MAX_TILE] (t = 1 is always allowed: the
lowering nests a large single tile by itself).  `tune` measures, on the
device, each candidate for a set of representative orders -- a small batch of
identical problems factored by one graph launch, so the figure is the
per-problem cost inside a batch -- and keeps the fastest tile size per
(kind, order).  The table is plain JSON; `choose` is a pure function of the
table: for order n it takes the tabulated order of the same kind nearest to n
(ties: the smaller one), and among n's candidates the tile size nearest to
the tabulated best tile size (ties: the larger tile).  Correctness does not
depend on the choice (every candidate factors to the same tolerance; the GPU
tests sweep them against the oracle).
"""
from __future__ import annotations

import json
import os
import time

MIN_TILE, MAX_TILE = 32, 1024
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning",
                             "mi355x.json")


def candidates(n):
    """Tile counts t for order n (ascending)."""
    out = [1]
    for t in range(2, n + 1):
        if n % t == 0 and MIN_TILE <= n // t <= MAX_TILE:
            out.append(t)
    return out


def choose(kind, n, table):
    """Tile count for (kind, n) from a tuning table {kind: {str(order): t}}."""
    entries = table.get(kind) or {}
    if not entries:
        return 1
    orders = sorted(int(k) for k in entries)
    n0 = min(orders, key=lambda o: (abs(o - n), o))
    b0 = n0 / entries[str(n0)]
    return min(candidates(n), key=lambda t: (abs(n / t - b0), t))


def load(path=DEFAULT_TABLE):
    with open(path) as f:
        return json.load(f)


def save(table, path=DEFAULT_TABLE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def measure(kind, n, t, reps=3, copies=32, seed=0):
    """Milliseconds per problem for `copies` identical order-n problems with
    t tiles each, factored by one graph launch (best of `reps`)."""
    import torch

    from .batch import Batch, make_problems
    kinds, ns = [kind] * copies, [n] * copies
    host = make_problems(kinds, ns, seed)
    b = Batch(kinds, ns, [t] * copies)
    best = float("inf")
    s = torch.cuda.Stream()
    for _ in range(reps):
        b.load(host)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(s):
            e0.record(s)
            b.launch(s)
            e1.record(s)
        s.synchronize()
        best = min(best, e0.elapsed_time(e1) / copies)
    ok = not b.infos().any()
    b.close()
    if not ok:
        raise RuntimeError(f"autotune: {kind} n={n} t={t} failed to factor")
    return best


def tune(kinds=("chol", "lu"), orders=(32, 64, 128, 192, 256, 384, 512, 768, 1024),
         path=None, log=print):
    """Measure every candidate of every (kind, order); returns (and saves, if
    `path`) the table of best tile counts plus the raw measurements."""
    table = {"_meta": {"device": _device_name(), "time": time.strftime("%Y-%m-%d %H:%M:%S"),
                       "method": "ms per problem, batch of 32 identical problems, one launch"}}
    raw = {}
    for kind in kinds:
        table[kind] = {}
        for n in orders:
            ms = {t: measure(kind, n, t) for t in candidates(n)}
            best = min(ms, key=lambda t: (ms[t], t))
            table[kind][str(n)] = best
            raw[f"{kind}/{n}"] = {str(t): round(v, 4) for t, v in ms.items()}
            if log:
                log(f"[autotune] {kind} n={n}: " +
                    ", ".join(f"t={t} {v:.3f}" for t, v in ms.items()) + f" -> t={best}")
    table["_raw"] = raw
    if path:
        save(table, path)
    return table


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # noqa: BLE001 - informational only
        return "unknown"


if __name__ == "__main__":
    # python -m mustard_amd.autotune [path]: measure and write the table
    import sys
    tune(path=sys.argv[1] if len(sys.argv) > 1 else DEFAULT_TABLE)
