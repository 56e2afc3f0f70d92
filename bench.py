"""Benchmark of tiled LU and tiled Cholesky at N = 24000 on one MI355X.

Workload: tiled LU and tiled Cholesky, FP64, N = 24000 cut in 24 x 24 tiles of
1000, each run both as one hipGraph launch whose nodes are level-batched
kernels (every independent task of one DAG level in one launch; the
reference's --tiled mode) and as one persistent-scheduler kernel launch (the
reference's --subgraph mode); the headline `value` is the faster mode, and
`modes` holds both.  LU uses the reference generator (include/gen.h, glibc
rand, seed 420); Cholesky uses A = M M^T / N + N I with M uniform [-1, 1)
(seed 420), built on the device.

One step = one LU plus one Cholesky factorization of the resident matrices
(each restored from a pristine device copy before it runs; the restore is
outside the timed window, as in the reference, lu_mustard.cu:605-611).
value = (2N^3/3 + N^3/3) / (LU time + Cholesky time), in GFLOP/s.

Extra objects on the JSON line: `roofline` for the dominant kernel
(k_batch_gemm, the level-batched GEMM/SYRK kernel of the trailing updates:
every launch of both factorizations timed with HIP events on the stream it
runs on, in a separate pass), `cpu_baseline` (the oracle's OpenMP task DAG
over host BLAS, timed on a bounded sample on this host), `mixed_ir` (A X = B
by FP64 / FP32+IR / BF16+IR at condition numbers 1e3 and 1e6), `batched`
(4608 uneven problems in one launch) and `single_kernel_mode` (T = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LU/Cholesky GFLOP/s at N=24000 FP64 vs mixed+IR at equal backward error"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 MFMA (public spec)
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n=24000, t=24):
    """Oracle OpenMP tiled LU + Cholesky on the host cores: the same N and tile
    size as the GPU workload, both on the reference generator's matrix (it is
    symmetric positive definite), roughly 10-20 s of CPU work."""
    from oracle import oracle as O
    O.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1, 16))
    A = O.gen_ref(n)
    F = A.copy(order="F")
    t0 = time.perf_counter()
    O.tiled_lu(F, t, threads=threads)
    t_lu = time.perf_counter() - t0
    F = A.copy(order="F")
    t0 = time.perf_counter()
    O.tiled_cholesky(F, t, threads=threads)
    t_ch = time.perf_counter() - t0
    flops = 2.0 * n ** 3 / 3 + n ** 3 / 3.0
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(flops / (t_lu + t_ch) * 1e-9, 1), "unit": "GFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"tiled LU + Cholesky, N={n}, {t}x{t} tiles of {n // t}, OpenMP task DAG "
                      f"with single-threaded OpenBLAS per tile ({os.path.basename(O.BLAS_PATH or 'loops')}), "
                      f"LU {t_lu:.2f} s, Cholesky {t_ch:.2f} s, CPU {cpu}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=24000)
    ap.add_argument("--t", type=int, default=24)
    ap.add_argument("--mode", default="both", choices=["both", "tiled", "subgraph", "stream"],
                    help="both: time the level-batched hipGraph (--tiled) and the persistent "
                         "scheduler (--subgraph); the headline is the faster")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-residual", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision + IR solves")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven factorizations")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-kernel (T = 1) mode factorizations")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build runs the tiled factorizations on a "
              "single GPU (multi-GPU distribution is out of scope); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import mustard_amd as mu

    torch.cuda.set_device(0)
    n, t = args.n, args.t
    b = n // t
    stream = torch.cuda.Stream()
    log(f"[bench] generating N={n} matrices")
    A_lu0 = mu.to_device(mu.gen_ref_host(n))
    A_ch0 = mu.gen_spd(n, 420)
    A_lu = torch.empty_like(A_lu0)
    A_ch = torch.empty_like(A_ch0)
    info = mu.info_cell()

    modes = ["tiled", "subgraph"] if args.mode == "both" else [args.mode]
    execs = {}
    with torch.cuda.stream(stream):
        for kind, A in (("lu", A_lu), ("chol", A_ch)):
            dag = mu.Dag()
            if kind == "lu":
                dag.record_lu(mu.ptr(A), n, t, None, mu.ptr(info))
            else:
                dag.record_cholesky(mu.ptr(A), n, t, None, mu.ptr(info))
            for mode in set(modes) | {"tiled"}:
                execs[mode, kind] = (dag, mu.Exec(dag, mode))
    for mode in modes:
        st = {k: execs[mode, k][1].stats() for k in ("lu", "chol")}
        log(f"[bench] {mode}: LU {st['lu']['units']} units / {st['lu']['items']} items / "
            f"{st['lu']['levels']} graph nodes / grid {st['lu']['grid']}, Cholesky "
            f"{st['chol']['units']} units / {st['chol']['items']} items")

    def run(mode, kind):
        src, dst = (A_lu0, A_lu) if kind == "lu" else (A_ch0, A_ch)
        with torch.cuda.stream(stream):
            dst.copy_(src)
            info.fill_(2 ** 31 - 1)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            execs[mode, kind][1].launch(stream)
            e1.record(stream)
        stream.synchronize()
        execs[mode, kind][1].check()
        assert mu.info_value(info) == 0, f"{kind}: info {mu.info_value(info)}"
        return e0.elapsed_time(e1)

    f_lu, f_ch = 2.0 * n ** 3 / 3.0, n ** 3 / 3.0
    timed = {}
    for mode in modes:
        for w in range(args.warmup):
            tl, tc = run(mode, "lu"), run(mode, "chol")
            log(f"[bench] {mode} warmup {w}: LU {tl:.1f} ms, Cholesky {tc:.1f} ms")
        torch.cuda.synchronize()
        lu_ms, ch_ms = [], []
        wall0 = time.perf_counter()
        for s in range(args.steps):
            lu_ms.append(run(mode, "lu"))
            ch_ms.append(run(mode, "chol"))
            log(f"[bench] {mode} step {s}: LU {lu_ms[-1]:.1f} ms, Cholesky {ch_ms[-1]:.1f} ms")
        torch.cuda.synchronize()
        wall = (time.perf_counter() - wall0) / max(args.steps, 1) * 1e3
        m_lu, m_ch = sum(lu_ms) / len(lu_ms), sum(ch_ms) / len(ch_ms)
        timed[mode] = {"lu_ms": lu_ms, "ch_ms": ch_ms, "m_lu": m_lu, "m_ch": m_ch, "wall": wall,
                       "value": (f_lu + f_ch) / ((m_lu + m_ch) * 1e-3) * 1e-9}
    head = max(timed, key=lambda m: timed[m]["value"])
    lu_ms, ch_ms, m_lu, m_ch, wall, value = (timed[head][k] for k in
                                             ("lu_ms", "ch_ms", "m_lu", "m_ch", "wall", "value"))
    how = {"tiled": "one hipGraph of level-batched kernels",
           "subgraph": "one persistent-scheduler kernel launch",
           "stream": "a stream of unit kernels"}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(m_lu + m_ch, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"tiled LU + tiled Cholesky, FP64, N={n}, {t}x{t} tiles of {b}, "
                               f"{how[head]} each",
                   "n": n, "tiles": t, "tile_size": b, "mode": head,
                   "lu_matrix": "reference generator (gen.h, glibc rand, seed 420)",
                   "cholesky_matrix": "M M^T/N + N I, M uniform[-1,1) seed 420"},
        "lu": {"ms": round(m_lu, 3), "ms_min": round(min(lu_ms), 3),
               "gflops": round(f_lu / (m_lu * 1e-3) * 1e-9, 1),
               "frac_fp64_peak": round(f_lu / (m_lu * 1e-3) * 1e-12 / PEAK_FP64_TFLOPS, 4)},
        "cholesky": {"ms": round(m_ch, 3), "ms_min": round(min(ch_ms), 3),
                     "gflops": round(f_ch / (m_ch * 1e-3) * 1e-9, 1),
                     "gflops_ref_convention_2n3_3": round(2 * f_ch / (m_ch * 1e-3) * 1e-9, 1),
                     "frac_fp64_peak": round(f_ch / (m_ch * 1e-3) * 1e-12 / PEAK_FP64_TFLOPS, 4)},
        "wall_ms_per_step_incl_restore": round(wall, 3),
        "modes": {m: {"lu_ms": round(v["m_lu"], 3), "cholesky_ms": round(v["m_ch"], 3),
                      "gflops": round(v["value"], 1),
                      "frac_fp64_peak": round(v["value"] * 1e-3 / PEAK_FP64_TFLOPS, 4)}
                  for m, v in timed.items()},
    }
    # Dominant kernel: k_batch_gemm, the GEMM/SYRK batch kernel that runs the
    # trailing updates.  Every level of each factorization is launched alone
    # between two HIP events on the bench stream; achieved = the algorithmic
    # flops of the GEMM/SYRK levels / their summed duration (= average flops
    # per launch / average launch duration).  The generic kernel (k_batch:
    # TRSM panels, diagonal factors) is reported beside it.
    if "tiled" in modes:
        per = {}
        for kind in ("lu", "chol"):
            src, dst = (A_lu0, A_lu) if kind == "lu" else (A_ch0, A_ch)
            ex = execs["tiled", kind][1]
            nlev = ex.stats()["levels"]
            info_l = [ex.level_info(l) for l in range(nlev)]
            with torch.cuda.stream(stream):
                dst.copy_(src)
                info.fill_(2 ** 31 - 1)
                evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * nlev)]
                for l in range(nlev):
                    evs[2 * l].record(stream)
                    ex.launch_level(l, stream)
                    evs[2 * l + 1].record(stream)
            stream.synchronize()
            assert mu.info_value(info) == 0
            for l in range(nlev):
                name, fl = info_l[l]
                p = per.setdefault(name, {"launches": 0, "ms": 0.0, "flops": 0.0})
                p["launches"] += 1
                p["ms"] += evs[2 * l].elapsed_time(evs[2 * l + 1])
                p["flops"] += fl
        g = per["k_batch_gemm"]
        achieved = g["flops"] / (g["ms"] * 1e-3) * 1e-12
        # HBM bytes per k_batch_gemm launch: FETCH_SIZE + WRITE_SIZE from the
        # rocprofv3 --pmc passes summarised in profiles/ (scaled by the same
        # counters on a copy of known size); not measurable inside this run
        traffic, per_bytes, per_util = None, {}, {}
        tp = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tp):
            with open(tp) as f:
                tj = json.load(f)
            if tj.get("kernel") == "k_batch_gemm" and n == 24000 and t == 24:
                traffic = tj["bytes_per_launch"]
                per_bytes = {k: v["fetch_bytes_per_launch"] + v["write_bytes_per_launch"]
                             for k, v in tj.get("per_kernel", {}).items()}
                per_util = {k: v["mfma_util"] for k, v in tj.get("per_kernel", {}).items()
                            if "mfma_util" in v}
        out["roofline"] = {"bound": "mfma", "achieved": round(achieved, 3),
                           "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                           "kernel": "k_batch_gemm", "launches": g["launches"],
                           "avg_launch_ms": round(g["ms"] / g["launches"], 4),
                           "avg_launch_gflop": round(g["flops"] / g["launches"] * 1e-9, 3)}
        if "k_batch_gemm" in per_util:  # from the rocprofv3 --pmc pass, not this run
            out["roofline"]["mfma_util_pmc"] = per_util["k_batch_gemm"]
        # the panel solves and diagonal factorizations beside it; hbm_gbs =
        # their HBM bytes per launch from the same counter passes as
        # `traffic` / this run's live average launch time
        pmc_name = {"k_batch_trsm_256": "k_batch_trsm_mid"}
        others = {}
        for name, o in per.items():
            if name == "k_batch_gemm":
                continue
            e = {"launches": o["launches"], "ms": round(o["ms"], 3),
                 "tflops": round(o["flops"] / (o["ms"] * 1e-3) * 1e-12, 3)}
            nb = per_bytes.get(pmc_name.get(name, name))
            if nb:
                e["hbm_bytes_per_launch"] = nb
                e["hbm_gbs"] = round(nb * o["launches"] / (o["ms"] * 1e-3) * 1e-9, 1)
                e["hbm_frac"] = round(e["hbm_gbs"] / PEAK_HBM_GBS, 4)
            if pmc_name.get(name, name) in per_util:
                e["mfma_util_pmc"] = per_util[pmc_name.get(name, name)]
            others[name] = e
        out["roofline"]["other_kernels"] = others
        for name, p in per.items():
            log(f"[bench] {name}: {p['launches']} launches, {p['ms']:.1f} ms, "
                f"{p['flops'] / (p['ms'] * 1e-3) * 1e-12:.2f} TFLOP/s")
    if not args.no_residual:
        for kind, A0, A in (("lu", A_lu0, A_lu), ("cholesky", A_ch0, A_ch)):
            r = mu.scaled_residual(A0, A, n, chol=(kind == "cholesky"))
            out[kind]["scaled_residual"] = round(r, 4)
            log(f"[bench] {kind} scaled residual {r:.4f}")
    # Mixed precision + FP64 refinement (the metric's second half): A X = B
    # with 16 right-hand sides on an SPD matrix of condition number 1e3
    # (A = Q diag(s) Q^T, Householder Q), time-to-solution of the plain FP64
    # tiled solve vs FP32 (and BF16) trailing updates + refinement, each to
    # the same stopping test and reported with its backward error.
    if not args.no_mixed:
        nrhs = 16

        def mixed_leg(kappa):
            Ac = mu.gen_cond(n, kappa, 420)
            Bm = torch.empty((nrhs, n), dtype=torch.float64, device="cuda")
            mu.call("mu_gen_uniform", mu.ptr(Bm), n * nrhs, 421, -1.0, 1.0, None)
            torch.cuda.synchronize()
            leg = {"rhs": nrhs, "kappa": kappa, "bound_10_n_eps": 10 * n * 2.0 ** -52}
            for kind in ("lu", "chol"):
                r = {}
                for prec in ("fp64", "fp32", "bf16"):
                    best = None
                    for _ in range(2):  # the first run warms the code paths
                        _, res = mu.solve_ir(kind, Ac, Bm, n, nrhs, t, prec=prec, fallback=True)
                        if best is None or res["total_ms"] < best["total_ms"]:
                            best = res
                    r[prec] = {"ms": round(best["total_ms"], 3),
                               "fact_ms": round(best["fact_ms"], 3), "iters": best["iters"],
                               "converged": best["converged"], "fell_back": best["fell_back"],
                               "berr": best["berr"], "berr_unrefined": best["berr0"]}
                    log(f"[bench] kappa {kappa:g} {kind} {prec}+IR: {best['total_ms']:.1f} ms, "
                        f"{best['iters']} steps, fell back {best['fell_back']}, "
                        f"berr {best['berr']:.2e}")
                r["speedup_fp32_ir"] = round(r["fp64"]["ms"] / r["fp32"]["ms"], 3)
                r["speedup_bf16_ir"] = round(r["fp64"]["ms"] / r["bf16"]["ms"], 3)
                leg[kind] = r
            del Ac, Bm
            return leg

        mixed = mixed_leg(1e3)
        # the ill-conditioned cell: kappa u_bf16 ~ 4e3, so plain BF16
        # refinement is not expected to converge; whatever happens (refined,
        # or the FP64 fallback) is reported with its backward error
        mixed["kappa_1e6"] = mixed_leg(1e6)
        out["mixed_ir"] = mixed
    # The reference's single-kernel mode (no --tiled/--subgraph, T = 1): one
    # task per factorization on the whole matrix, nested by the lowering in
    # two levels (blocks of 1000, then sub-tiles of 125), one graph launch.
    if not args.no_single:
        single = {}
        for kind, A0, A in (("lu", A_lu0, A_lu), ("cholesky", A_ch0, A_ch)):
            dag = mu.Dag()
            with torch.cuda.stream(stream):
                if kind == "lu":
                    dag.record_lu(mu.ptr(A), n, 1, None, mu.ptr(info))
                else:
                    dag.record_cholesky(mu.ptr(A), n, 1, None, mu.ptr(info))
            ex = mu.Exec(dag, "tiled")
            ms = []
            for _ in range(2):  # warm-up, then timed
                with torch.cuda.stream(stream):
                    A.copy_(A0)
                    info.fill_(2 ** 31 - 1)
                    e0, e1 = (torch.cuda.Event(enable_timing=True),
                              torch.cuda.Event(enable_timing=True))
                    e0.record(stream)
                    ex.launch(stream)
                    e1.record(stream)
                stream.synchronize()
                assert mu.info_value(info) == 0
                ms.append(e0.elapsed_time(e1))
            fl = f_lu if kind == "lu" else f_ch
            single[kind] = {"ms": round(ms[-1], 3),
                            "gflops": round(fl / (ms[-1] * 1e-3) * 1e-9, 1),
                            "units": ex.stats()["units"]}
            log(f"[bench] single-kernel mode {kind}: {ms[-1]:.1f} ms")
            del ex, dag
        out["single_kernel_mode"] = single
    # Batched uneven factorizations in one graph launch: 4096 SPD Cholesky
    # problems with n uniform in [32, 512] and 512 diagonally dominant LU
    # problems with n uniform in [64, 1024] (seed 420), tile counts from the
    # autotuner's table; aggregate GFLOP/s, and per-problem scaled residuals
    # and oracle comparison on a sample.
    if not args.no_batch:
        import numpy as np
        from mustard_amd import autotune
        from mustard_amd.batch import Batch, make_problems, unpack
        rng = np.random.default_rng(420)
        kinds = ["chol"] * 4096 + ["lu"] * 512
        ns = [int(x) for x in rng.integers(32, 513, 4096)] + \
             [int(x) for x in rng.integers(64, 1025, 512)]
        try:
            table = autotune.load()
            tsrc = os.path.relpath(autotune.DEFAULT_TABLE, ROOT)
        except OSError:
            table, tsrc = {}, "none (one tile per problem; the lowering nests)"
        tiles = [autotune.choose(k, nn, table) for k, nn in zip(kinds, ns)]
        host = make_problems(kinds, ns, 420)
        t0 = time.perf_counter()
        bt = Batch(kinds, ns, tiles)
        build_s = time.perf_counter() - t0
        ms = []
        for r in range(3):
            bt.load(host)
            torch.cuda.synchronize()
            with torch.cuda.stream(stream):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                bt.launch(stream)
                e1.record(stream)
            stream.synchronize()
            ms.append(e0.elapsed_time(e1))
        assert not bt.infos().any()
        worst = 0.0
        sample = list(range(0, len(ns), 36))
        for p in sample:
            A = unpack(host, ns, p)
            F = bt.factor(p)
            nn = ns[p]
            if kinds[p] == "lu":
                Lf = np.tril(F, -1) + np.eye(nn)
                R = A - Lf @ np.triu(F)
            else:
                Lf = np.tril(F)
                R = A - Lf @ Lf.T
            worst = max(worst, np.linalg.norm(R) / (np.linalg.norm(A) * nn * 2.0 ** -52))
        best = min(ms)
        out["batched"] = {"problems": len(ns), "cholesky": 4096, "lu": 512,
                          "gflop": round(bt.flops * 1e-9, 3), "ms": round(best, 3),
                          "gflops": round(bt.flops / (best * 1e-3) * 1e-9, 1),
                          "levels": bt.exec.stats()["levels"], "graph_launches": 1,
                          "tile_table": tsrc, "record_and_build_s": round(build_s, 2),
                          "max_scaled_residual_sample": round(worst, 4),
                          "sample": f"{len(sample)} problems (every 36th)"}
        log(f"[bench] batched: {len(ns)} problems, {bt.flops * 1e-9:.1f} GFLOP in {best:.2f} ms "
            f"({bt.flops / (best * 1e-3) * 1e-9:.0f} GFLOP/s), {out['batched']['levels']} levels, "
            f"max scaled residual {worst:.3f}")
        bt.close()
        # the same factorizations as one persistent-scheduler launch
        bp = Batch(kinds, ns, tiles, mode="subgraph")
        pms = []
        for r in range(2):
            bp.load(host)
            torch.cuda.synchronize()
            with torch.cuda.stream(stream):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                bp.launch(stream)
                e1.record(stream)
            stream.synchronize()
            bp.exec.check()
            pms.append(e0.elapsed_time(e1))
        assert not bp.infos().any()
        out["batched"]["persistent_ms"] = round(min(pms), 3)
        out["batched"]["persistent_gflops"] = round(bt.flops / (min(pms) * 1e-3) * 1e-9, 1)
        log(f"[bench] batched, one persistent launch: {min(pms):.2f} ms")
        bp.close()
        # A x = b for every problem: FP64 factorization + solve, and FP32 /
        # BF16 trailing updates + FP64 refinement of all problems together;
        # the backward error of every problem from the device residual kernel
        from mustard_amd.batch import BatchSolve, make_rhs
        rhs = make_rhs(ns, 421)
        solves = {}
        for prec in ("fp64", "fp32", "bf16"):
            bs = BatchSolve(kinds, ns, tiles, prec=prec)
            bs.load(host, rhs)
            best_r = None
            for r in range(2):  # the first run warms the code paths
                res = bs.solve(itermax=30, stream=stream)
                if best_r is None or res["ms"] < best_r["ms"]:
                    best_r = res
            assert not best_r["infos"].any()
            solves[prec] = {"ms": round(best_r["ms"], 3), "steps": best_r["steps"],
                            "converged": best_r["converged"],
                            "max_berr": float(np.max(best_r["berr"])),
                            "max_berr_unrefined": float(np.max(best_r["berr0"]))}
            log(f"[bench] batched solve {prec}: {best_r['ms']:.2f} ms, {best_r['steps']} steps, "
                f"max backward error {np.max(best_r['berr']):.2e} over {len(ns)} problems")
            bs.close()
        solves["speedup_fp32_ir"] = round(solves["fp64"]["ms"] / solves["fp32"]["ms"], 3)
        solves["speedup_bf16_ir"] = round(solves["fp64"]["ms"] / solves["bf16"]["ms"], 3)
        solves["berr_bound_10_nmax_eps"] = 10 * max(ns) * 2.0 ** -53
        out["batched"]["solve"] = solves
        del host
    if not args.no_cpu:
        log("[bench] CPU baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
