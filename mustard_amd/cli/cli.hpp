// cli.hpp -- command line of lu_mustard / cholesky_mustard (include/cli.h of
// the reference) and the run / verify / report steps they share.
//
// Parsing follows the argh conventions the reference relies on
// (include/argh.h:227-356, default mode PREFER_FLAG_FOR_UNREG_OPTION):
//   * leading dashes are dropped, so -tiled and --tiled are the same flag;
//   * "name=value" is a parameter; any other option is a flag, and a value
//     written after it as a separate word is a positional argument, so
//     "-n 600" leaves N at its default (only "-n=600" sets it).
// Validation and messages follow include/cli.h:67-97.  --sm is the CU budget
// of a task in the persistent scheduler, in percent of the GPU (1-100): the
// reference counts A100 SMs (1-108) against an admission limit of 100, and a
// percentage keeps that limit meaning "the whole GPU" on any CU count.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "mustard.h"

namespace cli {

struct Args {
    std::map<std::string, std::string> params;
    std::set<std::string> flags;
    std::vector<std::string> positional;

    bool flag(std::initializer_list<const char *> names) const
    {
        for (const char *n : names)
            if (flags.count(n))
                return true;
        return false;
    }
    const std::string *param(std::initializer_list<const char *> names) const
    {
        for (const char *n : names) {
            auto it = params.find(n);
            if (it != params.end())
                return &it->second;
        }
        return nullptr;
    }
    // argh's `cmdl(names, default) >> value`: true when the parameter is
    // absent (value keeps its default) or parses completely.
    template <class T>
    bool get(std::initializer_list<const char *> names, T &value) const
    {
        const std::string *s = param(names);
        if (!s)
            return true;
        std::istringstream is(*s);
        T v{};
        if (!(is >> v))
            return false;
        value = v;
        return true;
    }
    std::string str(std::initializer_list<const char *> names) const
    {
        const std::string *s = param(names);
        return s ? *s : std::string();
    }
};

inline Args parse(int argc, char **argv)
{
    Args a;
    for (int i = 1; i < argc; i++) {
        std::string w = argv[i];
        if (w.size() > 1 && w[0] == '-' && !std::isdigit((unsigned char)w[1])) {
            size_t k = w.find_first_not_of('-');
            std::string body = k == std::string::npos ? std::string() : w.substr(k);
            size_t eq = body.find('=');
            if (eq != std::string::npos)
                a.params[body.substr(0, eq)] = body.substr(eq + 1);
            else
                a.flags.insert(body);
        } else {
            a.positional.push_back(w);
        }
    }
    return a;
}

// include/cli.h:9-18
struct MustardConfig {
    size_t N = 15;
    size_t B = 3;
    size_t T = 5;
    int myPE = 0;
    int verbose = 0;
    int workspace = 256;  // kB, accepted and echoed (the tile kernels need none)
    int smLimit = 20;     // CU budget per task, percent of the GPU (--subgraph)
    int runs = 1;
};

inline void printCommonUsage()
{
    std::cerr << "\n  Common options:\n"
              << "    -n, -N=<int>         Matrix dimension N                       [default: 15]\n"
              << "    -t, -T=<int>         Number of tiles (N must be divisible)    [default: 5]\n"
              << "    --sm, --smLimit=<int> CU budget per task, % of the GPU (1-100) [default: 20]\n"
              << "    --ws, --workspace=<int> workspace in kB (1-1048576)          [default: 256]\n"
              << "    -r, --runs=<int>     Number of timing runs                    [default: 1]\n"
              << "    -v, --verbose        Enable verbose output\n"
              << "    --verify             Verify result correctness\n"
              << "    --dot                Dump the task graph in DOT format (./graph.dot)\n"
              << "\n  MI355X options:\n"
              << "    --gen=ref|spd        Matrix: reference generator (gen.h) or M M^T/N + N I\n"
              << "    --seed=<int>         Seed of the spd generator               [default: 420]\n"
              << "    --stream             Launch units one by one (debugging mode)\n"
              << "    --grid=<int>         Persistent scheduler workgroups (0 = all CUs)\n"
              << "    --trace=<file>       Chrome-trace JSON of the last --subgraph run\n"
              << "    --peak=<TFLOP/s>     FP64 MFMA peak for the fraction line [default: 78.6]\n"
              << "    --strict             Exit 1 when the factorization reports d_info != 0\n";
}

inline void printSingleNodeUsage(const char *progName, const char *decomposition)
{
    std::cerr << "Usage: " << progName << " [options]\n"
              << "\n  " << decomposition
              << " decomposition on one MI355X (tile kernels on gfx950, device-side scheduler).\n"
              << "\n  Mode (pick one; default is single-kernel if none given):\n"
              << "    --tiled              Tiled execution (one hipGraph, a kernel node per unit)\n"
              << "    --subgraph           Sub-graph (mustard) execution: persistent scheduler\n";
    printCommonUsage();
    std::cerr << "\n  Examples:\n"
              << "    " << progName << " -n=600 -t=2 --tiled --verify\n"
              << "    " << progName << " -n=24000 -t=24 --subgraph -r=5\n"
              << std::endl;
}

// include/cli.h:67-97
inline bool parseCommonArgs(const Args &cmdl, MustardConfig &cfg)
{
    if (cmdl.flag({"h", "help"}))
        return false;
    if (!cmdl.get({"N", "n"}, cfg.N)) {
        std::cerr << "Error: Must provide a valid N value! Got '" << cmdl.str({"N", "n"}) << "'"
                  << std::endl;
        return false;
    }
    if (!cmdl.get({"t", "T"}, cfg.T)) {
        std::cerr << "Error: Must provide a valid T value! Got '" << cmdl.str({"T", "t"}) << "'"
                  << std::endl;
        return false;
    }
    if (cfg.T == 0 || cfg.N % cfg.T > 0) {
        std::cerr << "Error: N must be divisible by T! Got 'N=" << cfg.N << " & T=" << cfg.T
                  << "'" << std::endl;
        return false;
    }
    if (!cmdl.get({"sm", "SM", "smLimit"}, cfg.smLimit) || cfg.smLimit > 100 || cfg.smLimit < 1) {
        std::cerr << "Error: Must provide a valid SM Limit value! Got '"
                  << cmdl.str({"sm", "SM", "smLimit"}) << "'" << std::endl;
        return false;
    }
    if (!cmdl.get({"workspace", "ws", "w", "W"}, cfg.workspace) || cfg.workspace > 1024 * 1024 ||
        cfg.workspace < 1) {
        std::cerr << "Error: Must provide a valid workspace (in kBytes) value! Got '"
                  << cmdl.str({"workspace", "ws", "w"}) << "'" << std::endl;
        return false;
    }
    if (!cmdl.get({"run", "runs", "r", "R"}, cfg.runs) || cfg.runs < 1) {
        std::cerr << "Error: Must provide a valid number of runs! Got '"
                  << cmdl.str({"run", "r", "R"}) << "'" << std::endl;
        return false;
    }
    return true;
}

// Replaces initNvshmemDevice (include/cli.h:101-116): one process, one GPU.
inline void initDevice(const Args &cmdl, MustardConfig &cfg)
{
    cfg.myPE = 0;
    checkMustard(mu_set_device(0));
    cfg.verbose = cmdl.flag({"v", "verbose"});
    if (cfg.verbose) {
        int n = 0;
        checkMustard(mu_get_device_count(&n));
        printf("%d GPUs detected, asked to use 1 GPU\n", n);
    }
}

inline void banner(const Args &cmdl, const MustardConfig &cfg)
{
    // lu_mustard.cu:698-711
    if (cmdl.flag({"tiled"}))
        std::cout << "TILED";
    else if (cmdl.flag({"subgraph"}))
        std::cout << "SUBGRAPH";
    else
        std::cout << "Single-kernel";
    std::cout << " with N=" << cfg.N << " (" << cfg.T << " of " << cfg.B << "x" << cfg.B
              << " tiles)" << std::endl;
    if (cmdl.flag({"subgraph", "tiled"})) {
        std::cout << "SM Limit per kernel = " << cfg.smLimit << std::endl;
        std::cout << "cuBLAS workspace = " << cfg.workspace << " kB" << std::endl;
    }
}

// Host matrix of the run: the reference generator (include/gen.h:23-38) or
// the SPD generator built on the device.
inline void makeMatrix(const Args &cmdl, size_t N, std::vector<double> &h, double *d_tmp,
                       double *d_work, hipStream_t s)
{
    h.assign(N * N, 0.0);
    const std::string gen = cmdl.param({"gen"}) ? *cmdl.param({"gen"}) : "ref";
    if (gen == "spd") {
        uint64_t seed = 420;
        cmdl.get({"seed"}, seed);
        checkMustard(mu_gen_spd(d_tmp, (int)N, (int)N, seed, d_work, s));
        checkMustard(mu_stream_synchronize(s));
        checkMustard(mu_memcpy(h.data(), d_tmp, N * N * sizeof(double)));
    } else {
        mu_gen_ref_host(h.data(), N);
    }
}

inline std::string opName(int op)
{
    static const char *names[] = {"NONE", "GEMM", "SYRK", "TRSM", "GETRF", "POTRF"};
    return (op >= 0 && op <= 5) ? names[op] : "?";
}

// Chrome-trace JSON of a traced persistent run.
inline void writeTrace(const std::string &path, mustard::Executable &ex)
{
    mu_exec_stats_t st = ex.stats();
    std::vector<unsigned long long> stamps(2 * st.items);
    checkMustard(mu_exec_trace(ex.handle(), stamps.data(), stamps.size()));
    std::vector<int> iu(st.items), tu(st.units), ou(st.units);
    checkMustard(mu_exec_item_units(ex.handle(), iu.data(), tu.data(), ou.data()));
    unsigned long long t0 = ~0ull;
    for (long long i = 0; i < st.items; i++)
        t0 = std::min(t0, stamps[2 * i]);
    const double us_per_tick = 1000.0 / st.wall_clock_khz;
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f)
        return;
    std::fprintf(f, "{\"traceEvents\":[\n");
    for (long long i = 0; i < st.items; i++) {
        const int u = iu[i];
        std::fprintf(f,
                     "%s{\"name\":\"%s\",\"cat\":\"task %d\",\"ph\":\"X\",\"pid\":0,\"tid\":%lld,"
                     "\"ts\":%.3f,\"dur\":%.3f,\"args\":{\"unit\":%d,\"task\":%d}}\n",
                     i ? "," : "", opName(ou[u]).c_str(), tu[u], i % 256,
                     (stamps[2 * i] - t0) * us_per_tick,
                     (stamps[2 * i + 1] - stamps[2 * i]) * us_per_tick, u, tu[u]);
    }
    std::fprintf(f, "]}\n");
    std::fclose(f);
}

// Runs the recorded DAG `runs` times as the reference does
// (lu_mustard.cu:605-648): upload A (untimed), time launch -> stream sync
// with events, print the `device` line.  Returns the per-run seconds.
inline std::vector<double> timedRuns(const Args &cmdl, const MustardConfig &cfg,
                                     mustard::Executable &ex, double *d_matrix,
                                     const std::vector<double> &h, int *d_info, hipStream_t s)
{
    void *e0 = nullptr, *e1 = nullptr;
    checkMustard(mu_event_create(&e0));
    checkMustard(mu_event_create(&e1));
    std::vector<double> times;
    for (int i = 0; i < cfg.runs; i++) {
        checkMustard(mu_memcpy(d_matrix, h.data(), h.size() * sizeof(double)));
        checkMustard(mu_info_reset(d_info, s));
        checkMustard(mu_stream_synchronize(s));
        checkMustard(mu_event_record(e0, s));
        ex.launch(s);
        checkMustard(mu_event_record(e1, s));
        checkMustard(mu_stream_synchronize(s));
        ex.check();
        float ms = 0;
        checkMustard(mu_event_elapsed_ms(&ms, e0, e1));
        printf("device %d | %d run | time (s): %4.4f\n", cfg.myPE, i, ms * 1e-3);
        times.push_back(ms * 1e-3);
    }
    checkMustard(mu_event_destroy(e0));
    checkMustard(mu_event_destroy(e1));
    (void)cmdl;
    return times;
}

// Reports, verification and the exit status shared by both programs.
inline int finish(const Args &cmdl, const MustardConfig &cfg, bool chol,
                  const std::vector<double> &times, const std::vector<double> &h,
                  double *d_matrix, int *d_info, hipStream_t s, double flops)
{
    const size_t N = cfg.N;
    int info = 0;
    checkMustard(mu_memcpy(&info, d_info, sizeof(int)));
    if (info == 0x7fffffff)
        info = 0;
    if (info != 0)
        std::cout << "Unsuccessful " << (chol ? "potrf" : "getrf") << " execution\n\n"
                  << "d_info = " << info << "\n\n";

    double total = 0, best = 1e300;
    for (double t : times) {
        total += t;
        best = std::min(best, t);
    }
    if (cmdl.flag({"verify"})) {
        // include/verify.h; the triple loop on the host for small N, the
        // same sum |A - LU| on the device above that.
        std::vector<double> F(N * N);
        checkMustard(mu_memcpy(F.data(), d_matrix, N * N * sizeof(double)));
        double err = 0, fr = 0, fa = 0;
        double *dA = nullptr, *work = nullptr;
        checkMustard(mu_malloc((void **)&dA, N * N * sizeof(double)));
        checkMustard(mu_malloc((void **)&work, 3 * N * N * sizeof(double)));
        checkMustard(mu_memcpy(dA, h.data(), N * N * sizeof(double)));
        checkMustard(mu_residual_norms(dA, d_matrix, (int)N, (int)N, chol ? 1 : 0, work, &fr, &fa,
                                       &err, s));
        if (N <= 2000)
            err = chol ? mu_verify_chol_l1_host(h.data(), F.data(), (int)N)
                       : mu_verify_lu_l1_host(h.data(), F.data(), (int)N);
        checkMustard(mu_free(dA));
        checkMustard(mu_free(work));
        printf("error = %.6f\n", err);
        printf("Result passes verification: %d\n", err <= 1e-6 ? 1 : 0);
        printf("scaled residual ||A-%s||_F/(||A||_F*N*eps) = %.4f (%s)\n", chol ? "LL^T" : "LU",
               fr / (fa * (double)N * 2.220446049250313e-16),
               fr / (fa * (double)N * 2.220446049250313e-16) <= 10.0 ? "pass" : "FAIL");
    }
    printf("Total time used (s): %4.4f\n", total);
    double peak = 78.6;
    cmdl.get({"peak"}, peak);
    const double n3 = (double)N * N * N;
    const double ref_conv = 2.0 * n3 / 3.0;  // scripts/calc.py:53-54 uses 2N^3/3 for both
    if (!times.empty() && best > 0) {
        printf("rate: best %.1f GFLOP/s (%.1f%% of %.1f TFLOP/s FP64 MFMA peak), %s flops %.4g "
               "(recorded tasks %.4g)\n",
               (chol ? n3 / 3.0 : 2.0 * n3 / 3.0) / best * 1e-9,
               (chol ? n3 / 3.0 : 2.0 * n3 / 3.0) / best * 1e-12 / peak * 100.0, peak,
               chol ? "N^3/3" : "2N^3/3", chol ? n3 / 3.0 : 2.0 * n3 / 3.0, flops);
        if (chol)
            printf("rate (reference convention, 2N^3/3): best %.1f GFLOP/s\n",
                   ref_conv / best * 1e-9);
    }
    // like the reference (lu_mustard.cu:716), a failed factorization is
    // reported above and the program still exits 0; --strict makes it exit 1
    return (info != 0 && cmdl.flag({"strict"})) ? 1 : 0;
}

}  // namespace cli
