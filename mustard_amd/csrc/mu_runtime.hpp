// mu_runtime.hpp -- execution of lowered unit DAGs on one gfx950 GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "mu_dag.hpp"
#include "mu_kernels.hpp"

namespace mu {

enum Mode : int {
    MODE_GRAPH = 0,       // one hipGraph, one batched kernel node per DAG level ("--tiled")
    MODE_PERSISTENT = 1,  // one persistent scheduler kernel ("--subgraph")
    MODE_STREAM = 2,      // units launched in topological order on the stream
};

struct ExecOpts {
    int grid = 0;              // persistent grid; 0 = one workgroup per resident slot
    double timeout_s = 10.0;   // persistent: seconds without progress before abort (the host
                               // watchdog aborts a run after 3 x timeout_s + 1 s)
    int trace = 0;             // persistent: record per-item wall-clock stamps
    int lookahead = 1;         // tiled: bulk updates on a second, concurrent chain
    int sm = 0;                // persistent: CU budget per recorded task, percent of the
                               // GPU (the reference's --sm); 0 = none
    int fuse = 1;              // tiled: merge a node's GEMM/SYRK units that tile a region
};


struct Exec {
    int mode = MODE_PERSISTENT;
    int device = 0;
    Lowered low;
    // device memory
    char *dbuf = nullptr;
    size_t dbytes = 0;
    double *ws = nullptr;            // inverted diagonal blocks of the TRSM units
    ExecView ev{};
    uint32_t *slots_init = nullptr;  // initial queue contents, same layout as the rings
    uint32_t slots_total = 0;
    uint32_t init_cnt[NPRIO] = {0, 0, 0, 0};
    uint32_t ntasks = 0;
    uint64_t *trace = nullptr;
    int grid = 0;
    uint64_t wall_khz = 100000;
    // host watchdog of a persistent run
    uint32_t *host_abort = nullptr;  // pinned host word the kernel polls
    hipEvent_t done_ev = nullptr;    // recorded after the persistent kernel
    std::thread watchdog;
    std::atomic<bool> wd_stop{false};
    std::atomic<bool> host_aborted{false};
    double host_timeout_s = 31.0;
    // graph mode
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int levels = 0;                  // graph nodes (level kernels of both chains)
    int depth = 0;                   // DAG levels
    std::vector<UnitDesc> gunits;  // tiled: the graph's unit table (lowered + merged units)
    std::vector<uint32_t> level_items;
    std::vector<size_t> level_off;
    std::vector<char> level_gemm;    // LEVEL_*: the kernel the level runs on
    std::vector<double> level_flops; // algorithmic flops of the units of each level
    const void *level_args[3] = {nullptr, nullptr, nullptr};
};

int exec_create(const Dag &dag, int mode, const LowerOpts &lopts, const ExecOpts &eopts,
                Exec **out, std::string &err);
int exec_launch(Exec *e, hipStream_t s, std::string &err);
// Tiled mode: launch the kernel of one DAG level alone (for per-launch timing).
int exec_launch_level(Exec *e, int level, hipStream_t s, std::string &err);
// After the stream has finished: scheduler status (0 = fine).
int exec_check(Exec *e, std::string &err);
int exec_trace(Exec *e, std::vector<uint64_t> &stamps, std::string &err);
void exec_destroy(Exec *e);

// Eager execution of the units of a lowered DAG in topological order.
int run_units(const Lowered &low, hipStream_t s, std::string &err);

// Device helpers used by the C ABI.
int gen_uniform_device(double *M, size_t count, uint64_t seed, double lo, double hi,
                       hipStream_t s);
int add_diag_scale(double *A, int n, int lda, double div, double shift, hipStream_t s);
int symmetrize_lower(double *A, int n, int lda, hipStream_t s);
int fill_int(int32_t *p, size_t count, int32_t v, hipStream_t s);
int split_lu(const double *F, int n, int ldf, double *L, double *U, int ld, int chol,
             hipStream_t s);
int sub_frob2(const double *A, int lda, const double *B, int ldb, int m, int n, double *acc,
              hipStream_t s);
int sum_abs(const double *A, int lda, int m, int n, double *acc, hipStream_t s);
int copy_matrix(const double *src, int lds, double *dst, int ldd, int m, int n, hipStream_t s);
// out[j] = max_i |M(i, j)| for the n columns of M (m x n)
int colmax_abs(const double *M, int ld, int m, int n, double *out, hipStream_t s);
int batch_residual(int nprob, const int *ns, const long long *a_off, const long long *v_off,
                   const double *A, const double *X, const double *B, double *R, double *norms,
                   hipStream_t s);
// *out = ||A||_inf (max row sum of |A|), through a work buffer of
// inf_norm_work_doubles(m, n) doubles
size_t inf_norm_work_doubles(int m, int n);
int inf_norm(const double *A, int lda, int m, int n, double *out, double *work, hipStream_t s);
// inf_norm of A that also copies A to D (ldd) in the same pass
int copy_inf_norm(const double *A, int lda, double *D, int ldd, int m, int n, double *out,
                  double *work, hipStream_t s);
// m x n column-major (ld lds) -> stored by rows (D[i*n + j]) and back
int to_rows(const double *S, int lds, double *D, int m, int n, hipStream_t s);
int from_rows(const double *S, double *D, int ldd, int m, int n, hipStream_t s);
// X += D
int add_matrix(double *X, int ldx, const double *D, int ldd, int m, int n, hipStream_t s);
// R = B - A X (A m x n, X n x nrhs) through a K-split partial buffer P of
// residual_partial_doubles(m, n) doubles; deterministic.
size_t residual_partial_doubles(int m, int n);
int residual(const double *A, int lda, const double *X, int ldx, const double *B, int ldb,
             double *R, int ldr, int m, int n, int nrhs, double *P, hipStream_t s);
// conditioned test matrices: A = diag(kappa^(-i/(n-1))); acc += x.y; A <- H A H
int gen_cond_start(double *A, int n, int lda, double kappa, hipStream_t s);
// t consecutive b x b identity tiles (tile k at D + k b^2, leading dimension b)
int identity_tiles(double *D, int b, int t, hipStream_t s);
// dst = tril(src) with the diagonal times dscale (zeros above); dst = (src + src^T) / 2
int tril_scaled(const double *src, int lds, double *dst, int ldd, int n, double dscale,
                hipStream_t s);
int sym_avg(const double *src, int lds, double *dst, int ldd, int n, hipStream_t s);
int dot_acc(const double *x, const double *y, int n, double *acc, hipStream_t s);
int reflect2(double *A, int lda, int n, const double *v, const double *w, const double *sc,
             hipStream_t s);

}  // namespace mu
