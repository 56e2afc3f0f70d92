// mustard.h -- the graph-building API of the reference (mustard/mustard.h),
// re-implemented over libmustard_hip.so for MI355X.
//
// Same names and argument meaning as the reference:
//   * MatrixTile, FLAGS_SUBG_COUNT, FLAGS_OCCUP, MAX_TILE, extern int myPE
//     (mustard/mustard.h:19-21, 51-53);
//   * mustard::TiledGraphCreator with subgraphs, graph, subgraphDependencies,
//     beginCaptureOperation, endCaptureOperation, printDeps,
//     insertDependencyKernel, insertSubgraph (mustard/mustard.h:132-248);
//   * BrokerWorkDistributor(size), free_mem() and the device enqueue /
//     dequeue / size (mustard/broker_queue.h:12-150), in mustard_queue.h.
//     The executables' persistent scheduler keeps its own per-priority rings
//     in HBM; the class is for kernels written against the reference's queue.
//
// What differs, and why:
//   * Between beginCaptureOperation and endCaptureOperation the caller issues
//     the library's tile operations (mu_dgemm, mu_dsyrk, mu_dtrsm,
//     mu_dgetrf_nopiv, mu_dpotrf from mustard_hip.h) on the creator's stream
//     instead of cuBLAS / cuSOLVER calls; they are recorded, not captured.
//   * HIP has no device-side graph launch, so the reference's per-task
//     sub-graphs, kernel_scheduler and kernel_dep_update nodes do not exist:
//     instantiate() lowers the recorded DAG either to one hipGraph (the
//     reference's --tiled mode) or to one persistent scheduler kernel with
//     its queue and dependency counters in HBM (--subgraph mode).
//   * Large diagonal factorizations are nested into sub-tiles automatically;
//     insertSubgraph still nests a hand-recorded sub-graph the way the
//     reference does for B > MAX_TILE (lu_mustard.cu:381-388): a creator
//     built on a graph handle that ends without being instantiated leaves its
//     recorded DAG under that handle, and insertSubgraph(handle) makes it a
//     nested sub-DAG of the last ended task.
//   * Errors print a message and exit(EXIT_FAILURE), as checkCudaErrors does
//     (include/utils.h:8-18).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "mustard_hip.h"
#include "mustard_queue.h"

#define FLAGS_SUBG_COUNT 0
#define FLAGS_OCCUP 4
#define MAX_TILE 4000

extern int myPE;

typedef std::pair<int, int> MatrixTile;

// checkCudaErrors for the C ABI: print and exit on failure.
#define checkMustard(val) ::mustard::check_((val), #val, __FILE__, __LINE__)
#define checkHip(val) ::mustard::check_hip_((val), #val, __FILE__, __LINE__)

namespace mustard {

inline void check_(int rc, const char *expr, const char *file, int line)
{
    if (rc != MU_OK) {
        std::fprintf(stderr, "MUSTARD error at %s:%d code=%d \"%s\": %s\n", file, line, rc, expr,
                     mu_last_error());
        std::exit(EXIT_FAILURE);
    }
}

inline void check_hip_(hipError_t e, const char *expr, const char *file, int line)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "HIP error at %s:%d code=%d \"%s\": %s\n", file, line, (int)e, expr,
                     hipGetErrorString(e));
        std::exit(EXIT_FAILURE);
    }
}

// Execution modes of an instantiated DAG.
enum class Mode : int {
    Tiled = MU_MODE_TILED,       // one hipGraph, one batched kernel node per DAG level
    Subgraph = MU_MODE_SUBGRAPH, // one persistent scheduler kernel
    Stream = MU_MODE_STREAM,     // units launched one by one (debugging)
};

// A lowered, uploaded DAG (the instantiated graph / scheduler of the
// reference, lu_mustard.cu:578-600, 636).
class Executable {
  public:
    Executable() = default;
    explicit Executable(mu_exec_t h) : h_(h) {}
    Executable(const Executable &) = delete;
    Executable &operator=(const Executable &) = delete;
    Executable(Executable &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    Executable &operator=(Executable &&o) noexcept
    {
        std::swap(h_, o.h_);
        return *this;
    }
    ~Executable()
    {
        if (h_)
            mu_exec_destroy(h_);
    }
    void launch(hipStream_t s) { checkMustard(mu_exec_launch(h_, s)); }
    // After synchronizing the stream: exits if the scheduler timed out.
    void check() { checkMustard(mu_exec_check(h_)); }
    mu_exec_stats_t stats() const
    {
        mu_exec_stats_t st;
        checkMustard(mu_exec_stats(h_, &st));
        return st;
    }
    mu_exec_t handle() const { return h_; }

  private:
    mu_exec_t h_ = nullptr;
};

// DAGs recorded by creators built on a graph handle that ended without being
// instantiated (the reference's recordSubgraph leaves its tasks in the
// cudaGraph_t it returns, lu_mustard.cu:133-254); insertSubgraph reads them.
inline std::map<hipGraph_t, mu_dag_t> &recordedGraphs()
{
    static std::map<hipGraph_t, mu_dag_t> m;
    return m;
}

// Drop the DAG recorded under `graph` (after its last insertSubgraph).
inline void releaseGraph(hipGraph_t graph)
{
    auto &m = recordedGraphs();
    auto it = m.find(graph);
    if (it != m.end()) {
        mu_dag_destroy(it->second);
        m.erase(it);
    }
}

class TiledGraphCreator {
  public:
    // Per-task graphs do not exist on HIP (no device-side graph launch); the
    // member is kept for source compatibility and stays empty.
    hipGraph_t *subgraphs = nullptr;
    hipGraph_t graph;
    std::vector<std::vector<int>> subgraphDependencies;

    TiledGraphCreator(hipStream_t stream, hipGraph_t graph, bool subgraph = false,
                      int totalNodes = 1)
        : graph(graph), stream_(stream), subgraph_(subgraph)
    {
        subgraphDependencies.resize(totalNodes > 0 ? totalNodes : 1);
        checkMustard(mu_dag_create(&dag_));
        checkMustard(mu_dag_capture_begin(dag_, stream_));
        capturing_ = true;
    }
    ~TiledGraphCreator()
    {
        if (capturing_)
            mu_dag_capture_end(dag_, stream_);
        capturing_ = false;
        if (!instantiated_ && graph && numTasks() > 0) {
            releaseGraph(graph);
            recordedGraphs()[graph] = dag_;  // kept for insertSubgraph(graph)
        } else {
            mu_dag_destroy(dag_);
        }
    }
    TiledGraphCreator(const TiledGraphCreator &) = delete;
    TiledGraphCreator &operator=(const TiledGraphCreator &) = delete;

    // mustard/mustard.h:148-167
    void beginCaptureOperation(MatrixTile tileToWrite, std::initializer_list<MatrixTile> tilesToRead)
    {
        std::vector<int> r;
        for (const MatrixTile &t : tilesToRead) {
            r.push_back(t.first);
            r.push_back(t.second);
        }
        checkMustard(mu_dag_begin_task(dag_, tileToWrite.first, tileToWrite.second, r.data(),
                                       (int)tilesToRead.size(), &current_));
        lastModifiedTile_ = tileToWrite;
    }

    // mustard/mustard.h:169-181
    void endCaptureOperation()
    {
        if (lastModifiedTile_.first == -1 || lastModifiedTile_.second == -1) {
            std::fprintf(stderr, "endCaptureOperation without beginCaptureOperation\n");
            std::exit(EXIT_FAILURE);
        }
        checkMustard(mu_dag_end_task(dag_));
        if (subgraph_) {
            if ((int)subgraphDependencies.size() <= current_)
                subgraphDependencies.resize(current_ + 1);
            subgraphDependencies[current_] = taskDeps(current_);
        }
        lastModifiedTile_ = std::make_pair(-1, -1);
    }

    // mustard/mustard.h:183-195 (same output format)
    void printDeps()
    {
        for (size_t i = 0; i < subgraphDependencies.size(); i++) {
            std::cout << i << ":";
            for (int d : subgraphDependencies[i])
                std::cout << " " << d;
            std::cout << std::endl;
        }
    }

    // mustard/mustard.h:197-211: an explicit edge src -> dst.  The recorded
    // dependencies already contain every edge the reference inserts; extra
    // ones are added.
    void insertDependencyKernel(int src, int dst, BrokerWorkDistributor, int *)
    {
        checkMustard(mu_dag_add_edge(dag_, src, dst));
    }

    // mustard/mustard.h:214-248: the sub-graph recorded on `getrfSubgraph`
    // (by a creator built on that handle and since destroyed) runs as a
    // nested sub-DAG of the last ended task, after that task's own
    // operations.  The library also nests large diagonal factorizations by
    // itself, so recording mu_dgetrf_nopiv on the whole tile is equivalent.
    void insertSubgraph(hipGraph_t getrfSubgraph)
    {
        auto it = recordedGraphs().find(getrfSubgraph);
        if (it == recordedGraphs().end()) {
            std::fprintf(stderr, "insertSubgraph: no sub-graph was recorded on this graph "
                                 "handle\n");
            std::exit(EXIT_FAILURE);
        }
        checkMustard(mu_dag_insert_subdag(dag_, it->second));
    }

    // The same with a creator still in scope (its recording ends here).
    void insertSubgraph(TiledGraphCreator &child)
    {
        child.endRecording();
        checkMustard(mu_dag_insert_subdag(dag_, child.dag_));
    }

    // ---- additions: lowering and execution --------------------------------
    std::vector<int> taskDeps(int task) const
    {
        int n = 0;
        checkMustard(mu_dag_task_deps(dag_, task, nullptr, 0, &n));
        std::vector<int> d(n);
        if (n)
            checkMustard(mu_dag_task_deps(dag_, task, d.data(), n, &n));
        return d;
    }

    int numTasks() const
    {
        int n = 0;
        checkMustard(mu_dag_num_tasks(dag_, &n));
        return n;
    }

    // Stop recording the stream's tile operations into this DAG (the stream
    // returns to an enclosing recording, if any, or to running them).
    void endRecording()
    {
        if (capturing_) {
            checkMustard(mu_dag_capture_end(dag_, stream_));
            capturing_ = false;
        }
    }

    // End recording and lower the DAG for `mode`.
    Executable instantiate(Mode mode, const mu_exec_opts_t *opts = nullptr)
    {
        endRecording();
        instantiated_ = true;
        mu_exec_t h = nullptr;
        checkMustard(mu_exec_create(dag_, (int)mode, opts, &h));
        return Executable(h);
    }

    // DOT file of the task DAG (the reference dumps its CUDA graphs with
    // cudaGraphDebugDotPrint, lu_mustard.cu:583-585, 634-635).
    void writeDot(const std::string &path) const
    {
        static const char *names[] = {"NONE", "GEMM", "SYRK", "TRSM", "GETRF", "POTRF"};
        std::ofstream f(path);
        f << "digraph mustard {\n";
        const int n = numTasks();
        for (int t = 0; t < n; t++) {
            int r = 0, c = 0, ops[4] = {0, 0, 0, 0}, nops = 0;
            checkMustard(mu_dag_task_info(dag_, t, &r, &c, ops, 4, &nops));
            f << "  t" << t << " [label=\"" << t << ": " << names[nops ? ops[0] % 6 : 0] << " ("
              << r << "," << c << ")\"];\n";
            for (int d : taskDeps(t))
                f << "  t" << d << " -> t" << t << ";\n";
        }
        f << "}\n";
    }

    mu_dag_t dag() const { return dag_; }
    hipStream_t stream() const { return stream_; }

  private:
    mu_dag_t dag_ = nullptr;
    hipStream_t stream_;
    bool subgraph_;
    bool capturing_ = false;
    bool instantiated_ = false;
    int current_ = -1;
    MatrixTile lastModifiedTile_ = std::make_pair(-1, -1);
};

}  // namespace mustard
