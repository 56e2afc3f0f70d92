/*
 * mustard_hip.h -- C ABI of libmustard_hip.so, the MI355X (gfx950) tiled LU /
 * Cholesky engine.
 *
 * Plain C types only: device pointers are `double *`, HIP streams and events
 * are `void *` (a hipStream_t / hipEvent_t), handles are opaque.  Every call
 * returns 0 on success or a negative MU_E* code; mu_last_error() gives the
 * message of the last failure on the calling thread.  Nothing here calls
 * exit().
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the ParCoreLab/mustard tree).
 */
#ifndef MUSTARD_HIP_H
#define MUSTARD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MU_OK 0
#define MU_EINVAL (-1)  /* bad argument */
#define MU_EHIP (-2)    /* HIP runtime error */
#define MU_ESTATE (-3)  /* call out of order (e.g. end without begin) */
#define MU_ERUN (-4)    /* device run failed (scheduler time-out) */
#define MU_ENOTSUP (-5) /* combination not supported */

/* Execution modes of a recorded DAG. */
#define MU_MODE_TILED 0    /* one hipGraph of batched kernels, one per DAG level (reference --tiled) */
#define MU_MODE_SUBGRAPH 1 /* one persistent scheduler kernel pulling ready items from HBM
                              queues (reference --subgraph) */
#define MU_MODE_STREAM 2   /* units launched one by one on the stream (debugging) */

/* Precision of trailing GEMM/SYRK updates. */
#define MU_PREC_FP64 0
#define MU_PREC_FP32 1 /* operands rounded to FP32, FP32 MFMA, FP32 accumulation */
#define MU_PREC_BF16 2 /* operands rounded to BF16, FP32 accumulation */

typedef struct mu_dag_s *mu_dag_t;
typedef struct mu_exec_s *mu_exec_t;

typedef struct {
    int grid;           /* persistent grid size, 0 = one workgroup per CU slot */
    double timeout_s;   /* persistent: abort after this long without progress (0 = 10 s);
                           a host watchdog aborts a run after 3 * timeout_s + 1 s */
    int trace;          /* persistent: record per-item wall-clock stamps */
    int small_max;      /* diagonal factorizations above this order nest (0 = 128) */
    int trsm_max;       /* triangular solves above this order nest (0 = 256; at most 1024) */
    int lookahead;      /* tiled: 0 (default) or 1 = two concurrent chains of level
                           kernels, the trailing-update GEMM/SYRK tasks that feed only
                           other updates beside everything else (lookahead);
                           -1 = one chain */
    int sm;             /* persistent: CU budget of each recorded task in percent of the
                           GPU (the reference's --sm / FLAGS_OCCUP, mustard.h:82-85,
                           lu_mustard.cu:290, 352-377): at most sm% of the workgroups run
                           one task's items at once, and a task starts only while the
                           budgets of the running tasks sum to less than 100.
                           0 = no budget. */
    int no_fuse;        /* tiled: 1 keeps every GEMM/SYRK unit of a graph node separate;
                           0 (default) merges units of one node that tile a rectangle
                           (or, for a Cholesky update, a lower triangle) of C into one
                           GEMM (SYRK) over the whole region: its 64 x 64 blocks then
                           ignore the tile boundaries (no ragged blocks at the edge of
                           every 1000-wide tile); results are bit-identical */
} mu_exec_opts_t;

typedef struct {
    int mode;
    int grid;
    long long units;
    long long items;
    long long edges;
    double flops;             /* algorithmic flops of the recorded operations */
    double critical_path_us;  /* scheduler's estimate */
    int priority_units[4];    /* units per priority level */
    double wall_clock_khz;
    int levels;               /* tiled mode: kernel nodes of the graph (launches per run) */
    int depth;                /* tiled mode: DAG levels (longest chain of units) */
} mu_exec_stats_t;

/* ---- errors and devices ------------------------------------------------ */
/* utils.h:8-18 (checkCudaErrors) becomes a status code plus this message. */
const char *mu_last_error(void);
int mu_get_device_count(int *count);
/* cli.h:101-116 (initNvshmemDevice -> cudaSetDevice(myPE)); one GPU here. */
int mu_set_device(int device);
int mu_device_props(int device, int *cu_count, int *clock_khz, size_t *lds_per_block,
                    char *name, int name_len);

/* ---- memory, streams and events (utils.h:70-110, CudaEventClock) -------- */
int mu_malloc(void **ptr, size_t bytes);
int mu_free(void *ptr);
int mu_memcpy(void *dst, const void *src, size_t bytes); /* direction inferred */
int mu_memcpy_async(void *dst, const void *src, size_t bytes, void *stream);
int mu_memset_zero(void *ptr, size_t bytes);
int mu_stream_create(void **stream);
int mu_stream_destroy(void *stream);
int mu_stream_synchronize(void *stream);
int mu_device_synchronize(void);
int mu_event_create(void **event);
int mu_event_destroy(void *event);
int mu_event_record(void *event, void *stream);
int mu_event_elapsed_ms(float *ms, void *start, void *end);
/* utils.h:20-39 (showMemUsage) */
int mu_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- tile operations ------------------------------------------------------
 * The arithmetic the reference takes from cuBLAS / cuSOLVER.  Column-major,
 * device pointers, LAPACK argument meaning.  When `stream` is being recorded
 * by mu_dag_capture_begin the call is added to the current task of that DAG
 * (the analogue of stream capture in mustard.h:155-166); otherwise it runs on
 * `stream` at once.                                                         */
/* cublasGemmEx(..., CUDA_R_64F, COMPUTE_64F) -- lu_mustard.cu:511-522,
 * cholesky_mustard.cu:376-387 */
int mu_dgemm(void *stream, char transa, char transb, int m, int n, int k, double alpha,
             const double *A, int lda, const double *B, int ldb, double beta, double *C,
             int ldc);
/* cublasDsyrk -- cholesky_mustard.cu:327-333 (uplo 'L', trans 'N') */
int mu_dsyrk(void *stream, char uplo, char trans, int n, int k, double alpha, const double *A,
             int lda, double beta, double *C, int ldc);
/* cublasDtrsm -- lu_mustard.cu:413-422, 459-468; cholesky_mustard.cu:281-290.
 * All side/uplo/trans/diag combinations. */
int mu_dtrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
             double alpha, const double *A, int lda, double *B, int ldb);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  mu_dtrsm with the same arguments, meaning and result
 * (bit for bit), run at once on `stream`; the inverted diagonal blocks of a
 * triangle of order > 128 go to a fixed device array of
 * MU_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of a
 * per-call allocation, so triangles of order <= 8192 are accepted and larger
 * ones rejected with MU_EINVAL.  Calls that overlap on one GPU share that
 * array and race; a stream being recorded into a DAG is refused
 * (MU_ESTATE).  Allocates no device memory. */
#define MU_NOALLOC_SCRATCH_BYTES (1u << 21)
#define MU_NOALLOC_MAX_ORDER 8192
int mu_dtrsm_noalloc(void *stream, char side, char uplo, char transa, char diag, int m, int n,
                     double alpha, const double *A, int lda, double *B, int ldb);
/* No counterpart in the reference: the substitution step of mu_solve_ir's
 * reduced-precision refinement solves.  C = beta C + alpha A B with A an
 * m x k FP32 matrix (column-major, lda in elements, op(A) = A), B k x n and
 * C m x n FP64, 0 <= n <= 16: the products run in FP32 on B rounded to FP32
 * and the sums are added into C in FP64.  B and C are column-major, or with
 * `rows` != 0 stored by rows (B(k, j) at B[k*ldb + j], C(i, j) at
 * C[i*ldc + j]).  Recorded or run like mu_dgemm. */
int mu_gemm_a32(void *stream, int m, int n, int k, double alpha, const float *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc, int rows);
/* cusolverDnDgetrf with devIpiv = NULL -- lu_mustard.cu:360-369.  `info_dev`
 * (device int, may be NULL) receives the 1-based column of the first zero
 * pivot through an atomic minimum: initialise it to INT_MAX (mu_info_reset). */
int mu_dgetrf_nopiv(void *stream, int n, double *A, int lda, int *info_dev);
/* cusolverDnDpotrf(LOWER) -- cholesky_mustard.cu:238-246 */
int mu_dpotrf(void *stream, char uplo, int n, double *A, int lda, int *info_dev);
/* Set an info cell to INT_MAX (= no failure) on `stream`. */
int mu_info_reset(int *info_dev, void *stream);
/* Precision of the GEMM/SYRK calls recorded next on `stream` (MU_PREC_*). */
int mu_set_update_precision(void *stream, int prec);

/* ---- task DAG (mustard::TiledGraphCreator, mustard.h:132-384) ------------ */
int mu_dag_create(mu_dag_t *dag);
int mu_dag_destroy(mu_dag_t dag);
int mu_dag_capture_begin(mu_dag_t dag, void *stream);
int mu_dag_capture_end(mu_dag_t dag, void *stream);
/* beginCaptureOperation(W, R) -- mustard.h:148-167.  `read_tiles` holds
 * nread (row, col) pairs.  Returns the new task's index in *task. */
int mu_dag_begin_task(mu_dag_t dag, int w_row, int w_col, const int *read_tiles, int nread,
                      int *task);
/* endCaptureOperation -- mustard.h:169-181 */
int mu_dag_end_task(mu_dag_t dag);
int mu_dag_num_tasks(mu_dag_t dag, int *ntasks);
/* subgraphDependencies[task] -- mustard.h:137, 279-296 */
int mu_dag_task_deps(mu_dag_t dag, int task, int *deps, int max, int *ndeps);
/* insertDependencyKernel(src, dst, ...) -- mustard.h:197-211 */
int mu_dag_add_edge(mu_dag_t dag, int src, int dst);
/* insertSubgraph (mustard.h:214-248, used at lu_mustard.cu:381-388): the
 * tasks of `child` (recorded like any DAG, typically the tiled factorization
 * of one large diagonal tile, lu_mustard.cu:133-254) become a nested
 * sub-DAG of the last task ended in `dag`; it runs after that task's own
 * operations and before the task counts as done, with the child's own
 * dependencies inside it.  The child is copied: it may be destroyed (or
 * inserted again) afterwards; device buffers it owns stay alive with `dag`.
 * Recordings on one stream nest: capture_begin of the child on the parent's
 * stream records there until the child's capture_end. */
int mu_dag_insert_subdag(mu_dag_t dag, mu_dag_t child);
/* Problem tag for batched recording: tiles of different tags never alias. */
int mu_dag_set_tag(mu_dag_t dag, int tag);
/* Operation codes reported by mu_dag_task_info. */
#define MU_OP_GEMM 1
#define MU_OP_SYRK 2
#define MU_OP_TRSM 3
#define MU_OP_GETRF 4
#define MU_OP_POTRF 5
#define MU_OP_TRINV 6 /* internal: inverted 16x16 diagonal blocks of a large TRSM triangle */
#define MU_OP_CVT 7   /* internal: a panel tile rounded to the update precision */
/* Written tile and operation codes of a recorded task (for DOT output and
 * inspection; the reference prints its graphs with cudaGraphDebugDotPrint,
 * lu_mustard.cu:583-585, 634-635). */
int mu_dag_task_info(mu_dag_t dag, int task, int *w_row, int *w_col, int *ops, int max_ops,
                     int *nops);

/* ---- executables ---------------------------------------------------------
 * mu_exec_create lowers the DAG and uploads it (the reference's graph
 * instantiation, lu_mustard.cu:543-604 / 633-636); mu_exec_launch replays it
 * (cudaGraphLaunch of the scheduler or of the tiled graph, lu_mustard.cu:609,
 * 640).  The caller's recorded device pointers must stay valid; buffers the
 * library allocated for the DAG (reduced-precision panel copies) live until
 * the DAG and every executable built from it are destroyed, in any order. */
int mu_exec_create(mu_dag_t dag, int mode, const mu_exec_opts_t *opts, mu_exec_t *exec);
int mu_exec_launch(mu_exec_t exec, void *stream);
/* Tiled executables: launch the batched kernel of one graph node alone, and
 * its item count (per-launch timing; nodes are numbered in an order that
 * respects every dependency, so launching 0 .. levels-1 in turn is a full
 * run without overlap).  The "level" functions below index these nodes. */
int mu_exec_launch_level(mu_exec_t exec, int level, void *stream);
int mu_exec_level_items(mu_exec_t exec, int level, long long *items);
/* Items of one level per operation: counts[MU_OP_*] (8 entries). */
int mu_exec_level_ops(mu_exec_t exec, int level, long long *counts);
/* The kernel one level runs on (MU_KERNEL_GENERIC: k_batch, any item kind;
 * MU_KERNEL_GEMM: k_batch_gemm, GEMM/SYRK blocks only; MU_KERNEL_SMALL:
 * k_batch_small, diagonal GETRF/POTRF of order <= 128 only; MU_KERNEL_TRSM: k_batch_trsm, 8-wave panel solves with
 * inverted diagonal blocks only; MU_KERNEL_TRSM_SMALL: k_batch_trsm_small,
 * panel solves of order <= 128 without inverted blocks only; MU_KERNEL_TRSM_256 /
 * MU_KERNEL_TRSM_512: k_batch_trsm_mid<256 / 512>, 4-wave panel solves with
 * inverted diagonal blocks of order <= 256 / 512 and the block inversions
 * before them; MU_KERNEL_GEMM_NARROW: k_batch_gemm_narrow, GEMMs with at most
 * 16 output columns; MU_KERNEL_GEMM_LOW: k_batch_gemm_low, FP32 / BF16 updates
 * reading the converted panels, both operands k-contiguous; MU_KERNEL_CVT:
 * k_batch_cvt, precision conversions only) and the algorithmic
 * flops of its units (GEMM 2mnk, SYRK n(n+1)k, TRSM m^2 n, GETRF 2m^3/3,
 * POTRF m^3/3).  Either output may be NULL. */
#define MU_KERNEL_GENERIC 0
#define MU_KERNEL_GEMM 1
#define MU_KERNEL_SMALL 2
#define MU_KERNEL_TRSM 3
#define MU_KERNEL_TRSM_SMALL 4
#define MU_KERNEL_TRSM_256 5
#define MU_KERNEL_TRSM_512 6
#define MU_KERNEL_GEMM_NARROW 7
#define MU_KERNEL_GEMM_LOW 8
#define MU_KERNEL_CVT 9
int mu_exec_level_info(mu_exec_t exec, int level, int *kernel, double *flops);
/* After the stream is synchronized: MU_ERUN if the scheduler timed out. */
int mu_exec_check(mu_exec_t exec);
int mu_exec_stats(mu_exec_t exec, mu_exec_stats_t *stats);
/* Per-item (start, end) wall-clock stamps of the last persistent run;
 * `stamps` holds 2 * items entries.  Needs opts.trace. */
int mu_exec_trace(mu_exec_t exec, unsigned long long *stamps, size_t count);
/* Unit of each item and task of each unit (for trace export). */
int mu_exec_item_units(mu_exec_t exec, int *unit_of_item, int *task_of_unit, int *op_of_unit);
int mu_exec_destroy(mu_exec_t exec);

/* ---- whole tiled factorizations ------------------------------------------
 * Record the reference's task sequence for an n x n column-major matrix cut
 * into t x t tiles (lu_mustard.cu:343-535, cholesky_mustard.cu:221-400) into
 * `dag`.  `prec` selects the trailing-update precision. */
int mu_record_tiled_lu(mu_dag_t dag, double *A, int n, int t, int lda, int *info_dev, int prec);
int mu_record_tiled_cholesky(mu_dag_t dag, double *A, int n, int t, int lda, int *info_dev,
                             int prec);
/* Triangular solves with the factors of a tiled factorization, recorded on
 * the same kind of DAG (forward then back substitution, one task per tile
 * operation on the right-hand-side block rows): F holds the factors left by
 * mu_record_tiled_lu (kind MU_SOLVE_LU: L unit lower, U upper) or
 * mu_record_tiled_cholesky (MU_SOLVE_CHOL: L lower); X (n x nrhs, ldx) holds
 * B and is overwritten with the solution.  The reference has no solve phase
 * (it stops after the factorization, lu_mustard.cu:650-660); this is the
 * GETRS/POTRS step the refinement below is built on. */
#define MU_SOLVE_LU 0
#define MU_SOLVE_CHOL 1
int mu_record_tiled_solve(mu_dag_t dag, const double *F, int n, int t, int ldf, int kind,
                          double *X, int ldx, int nrhs);

/* Batched uneven factorizations: records nprob independent tiled LU
 * (kinds[p] = MU_SOLVE_LU) or Cholesky (MU_SOLVE_CHOL) problems of order
 * ns[p] at As[p] (leading dimension ldas[p], tiles[p] tiles per dimension,
 * ns[p] % tiles[p] == 0) into one DAG, each under its own tag; one executable
 * then runs them all in one graph launch, every level kernel mixing the
 * items of all problems.  info_dev (nprob ints, or NULL) receives each
 * problem's info; prec applies to every problem's trailing updates.  No
 * counterpart in the reference (one factorization per process there). */
int mu_record_batch(mu_dag_t dag, int nprob, const int *kinds, const int *ns, double *const *As,
                    const int *ldas, const int *tiles, int *info_dev, int prec);
/* Residuals of a batch of packed problems (no reference counterpart: the
 * refinement of batched reduced-precision solves).  Device arrays: ns[p] the
 * order of problem p, a_off[p] the element offset of its n x n matrix
 * (leading dimension n) in A, v_off[p] the offset of its vectors in X, B and
 * R.  Writes r_p = b_p - A_p x_p (FP64) to R and, per problem, norms[4p..4p+3]
 * = ||r||_inf, ||x||_inf, ||A||_inf, ||b||_inf (NaN propagates).  One
 * workgroup per problem, one launch. */
int mu_batch_residual(int nprob, const int *ns, const long long *a_off, const long long *v_off,
                      const double *A, const double *X, const double *B, double *R,
                      double *norms, void *stream);

/* ---- Mixed precision + FP64 iterative refinement ------------------------
 * No counterpart in the reference: the precision form of the tiled
 * factorizations.  Solves A X = B (A n x n, B and X n x nrhs, device
 * pointers; A and B are not modified) by a tiled factorization whose trailing
 * GEMM/SYRK run in `prec` (diagonal factorizations and panel solves in
 * FP64), followed by the refinement of LAPACK dsgesv/dsposv: x = x +
 * solve(b - A x), residual and correction in FP64, the triangular solves on
 * FP32 copies of the factors when `prec` is FP32 or BF16 (FP64 otherwise), until
 * ||r_j||_inf <= ||x_j||_inf ||A||_inf 2^-53 sqrt(n) for every column, at
 * most itermax times (0: 30); with `fallback`, a run whose reduced-precision
 * factorization breaks down (info > 0; iters is then 0) or whose refinement
 * does not converge is refactored and solved in FP64.  The factors live in a
 * library-owned n x n workspace (plus an n x n FP32 copy for reduced
 * precision).  Graph construction happens before the timed region. */
typedef struct {
    int kind;     /* MU_SOLVE_LU or MU_SOLVE_CHOL */
    int tiles;    /* tiles per dimension (n % tiles == 0) */
    int prec;     /* MU_PREC_FP64 / MU_PREC_FP32 / MU_PREC_BF16 trailing updates */
    int itermax;  /* refinement iterations cap; 0 = 30 */
    int fallback; /* 1: refactor in FP64 when the reduced-precision factorization
                     fails or the refinement does not converge */
} mu_ir_opts_t;
typedef struct {
    int info;            /* factorization info (0 = fine) */
    int iters;           /* refinement steps taken */
    int converged;       /* the stopping test held */
    int fell_back;       /* the FP64 fallback ran */
    double anorm;        /* ||A||_inf */
    double berr0;        /* normwise backward error of the unrefined solution */
    double berr;         /* final max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) */
    double berr_hist[32];/* backward error after each step */
    double fact_ms;      /* device time: ||A||, copy and factorization */
    double solve_ms;     /* device time: solves, residuals, corrections */
    double total_ms;     /* both */
} mu_ir_result_t;
int mu_solve_ir(const double *A, int n, int lda, const double *B, int ldb, double *X, int ldx,
                int nrhs, const mu_ir_opts_t *opts, mu_ir_result_t *res, void *stream);

/* ---- Reverse mode of the tiled Cholesky ------------------------------------
 * No counterpart in the reference.  Given the factor L (lower triangle used)
 * of A = L L^T and the gradient Lbar of a scalar loss with respect to the
 * lower triangle of L, writes the symmetric gradient with respect to A,
 * G = (Abar + Abar^T)/2 with Abar = L^-T Phi(L^T Lbar) L^-1 (Phi: lower
 * triangle, diagonal halved; Murray, arXiv:1602.07527), computed by three
 * tiled DAGs of GEMM and TRSM tasks over t x t tiles (n % t == 0).  `prec`
 * sets the precision of the GEMM tasks (MU_PREC_*).  L and Lbar are not
 * modified; G may not alias them.  Workspace: 3 n^2 doubles, library-owned. */
int mu_chol_backward(const double *L, int ldl, const double *Lbar, int ldlb, double *G, int ldg,
                     int n, int t, int prec, void *stream);

/* Test matrices of a given condition number (the mixed-precision inputs):
 * A = Q diag(s) Q^T with s_i = kappa^(-i/(n-1)) (so ||A||_2 = 1 and
 * cond_2(A) = kappa, A symmetric positive definite) and Q the product of
 * `nrefl` Householder reflectors I - 2 v v^T / v^T v, v uniform [-1, 1) from
 * the counter-based stream of mu_gen_uniform (reflector r uses stream seed+r).
 * work: 3n + 2 doubles of device memory. */
int mu_gen_cond(double *A, int n, int lda, double kappa, int nrefl, uint64_t seed, double *work,
                void *stream);

/* ---- generators (include/gen.h) ----------------------------------------- */
/* gen.h:23-38: glibc srand(420) symmetric matrix plus n I, host memory. */
void mu_gen_ref_host(double *A, size_t n);
/* Counter-based uniform stream on the device (SplitMix64): M[k] = lo + (hi -
 * lo) * u(seed, k). */
int mu_gen_uniform(double *M, size_t count, uint64_t seed, double lo, double hi, void *stream);
/* A = M M^T / n + n I with M uniform [-1, 1) from `seed`, built on the device
 * with the library's own GEMM; `work` is an n*n device scratch buffer. */
int mu_gen_spd(double *A, int n, int lda, uint64_t seed, double *work, void *stream);

/* ---- verification (include/verify.h) ------------------------------------ */
/* verify.h:17-71 / 79-129: sum |A - LU| (or |A - L L^T|) by the triple
 * loop, host memory, small n. */
double mu_verify_lu_l1_host(const double *A, const double *F, int n);
double mu_verify_chol_l1_host(const double *A, const double *F, int n);
/* ||A - L U||_F / (||A||_F n eps) on the device (chol != 0: L L^T).  A and F
 * are n x n with leading dimension lda; `work` is 3 n*n doubles of scratch. */
int mu_scaled_residual(const double *A, const double *F, int n, int lda, int chol,
                       double *work, double *result, void *stream);
/* The same product, returning ||A - LU||_F, ||A||_F and the reference's
 * verify quantity sum |A - LU| (include/verify.h:46-53), all on the device. */
int mu_residual_norms(const double *A, const double *F, int n, int lda, int chol, double *work,
                      double *frob_res, double *frob_a, double *l1_res, void *stream);

/* ---- diagnostics ----------------------------------------------------------
 * Run the persistent scheduler's queue/counter protocol for `dag` on host
 * threads (no GPU; items are checked, not computed).  results[0..5] =
 * {items executed once, duplicate executions, dependency-order violations,
 * scheduler error (0 ok, 1 time-out), items counted done, total items}.
 * `bump_unit` >= 0 adds a never-satisfied dependency to that unit (tests the
 * progress watchdog).  No reference counterpart: the reference has no host
 * model of its device queue (mustard/broker_queue.h). */
int mu_dag_emulate(mu_dag_t dag, int threads, double timeout_s, int bump_unit,
                   long long *results);
/* The same with every option: `workgroups` emulated workgroups of four
 * threads (waves) each, running the kernel's own worker loop with a barrier
 * per workgroup; a host watchdog deadline; a stall of every producer between
 * reserving ring positions and storing them; and the CU budget `sm` (percent,
 * as in mu_exec_opts_t).  results[0..8] = the six above, then the most
 * workgroups that ran one task's items at once, the per-task cap, and the
 * most tasks that had items running at once. */
typedef struct {
    int workgroups;
    double timeout_s;       /* device watchdog: no progress for this long */
    double host_deadline_s; /* host watchdog (0 = none) */
    int bump_unit;          /* >= 0: a never-satisfied dependency on this unit */
    int producer_delay_us;
    int sm;
} mu_emu_opts_t;
int mu_dag_emulate_ex(mu_dag_t dag, const mu_emu_opts_t *opts, long long *results);

#ifdef __cplusplus
}
#endif
#endif
