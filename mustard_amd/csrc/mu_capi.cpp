// mu_capi.cpp -- the extern "C" entry points of libmustard_hip.so.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mustard_hip.h"
#include "mu_capi.hpp"
#include "mu_dag.hpp"
#include "mu_runtime.hpp"
#include "mu_sched_host.hpp"


// Device buffers the recorded operations of a DAG use (the reduced-precision
// panel shadows): owned jointly by the DAG and every executable built from
// it, and freed when the last of them is destroyed.
struct OwnedBuffers {
    std::vector<void *> p;
    // buffers of sub-graphs inserted into this DAG (mu_dag_insert_subdag)
    std::vector<std::shared_ptr<OwnedBuffers>> sub;
    ~OwnedBuffers()
    {
        for (void *q : p)
            (void)hipFree(q);
    }
};

struct mu_dag_s {
    mu::Dag dag;
    std::shared_ptr<OwnedBuffers> owned = std::make_shared<OwnedBuffers>();
};


struct mu_exec_s {
    mu::Exec *e = nullptr;
    std::shared_ptr<OwnedBuffers> owned;  // the DAG's buffers its units point into
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

}  // namespace

int mu::set_error(int code, const std::string &msg) { return fail(code, msg); }

namespace {

#define HIPTRY(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail(MU_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
    } while (0)

// Streams being recorded into a DAG, and the update precision per stream.
// Recordings nest: a sub-graph recorded on the stream of a graph that is
// being recorded (the reference's recordSubgraph, lu_mustard.cu:133-254)
// takes the stream until its recording ends, then the outer one resumes.
struct StreamRec {
    std::vector<mu_dag_t> dags;  // innermost last
    int prec = MU_PREC_FP64;
};
struct StreamState {
    mu_dag_t dag = nullptr;  // the innermost recording DAG
    int prec = MU_PREC_FP64;
};
std::mutex g_mu;
std::map<void *, StreamRec> g_streams;

StreamState stream_state(void *s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_streams.find(s);
    StreamState st;
    if (it != g_streams.end()) {
        st.prec = it->second.prec;
        if (!it->second.dags.empty())
            st.dag = it->second.dags.back();
    }
    return st;
}

int prec_flags(int prec)
{
    return prec == MU_PREC_FP32 ? mu::F_LOW_F32 : prec == MU_PREC_BF16 ? mu::F_LOW_BF16 : 0;
}

// Record `op` into the DAG capturing `stream`, or run it now.
int issue(void *stream, const mu::OpRec &op)
{
    StreamState st = stream_state(stream);
    if (st.dag) {
        if (st.dag->dag.add_op(op) != 0)
            return fail(MU_ESTATE, st.dag->dag.err);
        return MU_OK;
    }
    mu::Dag d;
    d.begin_task({0, 0}, {});
    d.add_op(op);
    d.end_task();
    mu::Lowered low;
    std::string err;
    if (!mu::lower(d, mu::LowerOpts(), low, err))
        return fail(MU_EINVAL, err);
    if (mu::run_units(low, (hipStream_t)stream, err) != 0)
        return fail(MU_EHIP, err);
    return MU_OK;
}

bool is_t(char c) { return c == 'T' || c == 't' || c == 'C' || c == 'c'; }
bool is_n(char c) { return c == 'N' || c == 'n'; }
char up(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 'a' + 'A') : c; }

}  // namespace

// mu_dtrsm's arguments as one recorded operation (validated).
int mu::dtrsm_op(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                 const double *A, int lda, double *B, int ldb, OpRec &op)
{
    side = up(side);
    uplo = up(uplo);
    diag = up(diag);
    const bool t = is_t(transa);
    if ((side != 'L' && side != 'R') || (uplo != 'L' && uplo != 'U') ||
        (!t && !is_n(transa)) || (diag != 'U' && diag != 'N'))
        return fail(MU_EINVAL, "mu_dtrsm: bad side/uplo/trans/diag");
    if (m < 0 || n < 0 || ldb < std::max(1, m) || lda < std::max(1, side == 'L' ? m : n))
        return fail(MU_EINVAL, "mu_dtrsm: bad size or leading dimension");
    op = OpRec();
    if (alpha == 0.0) {
        // BLAS: alpha = 0 sets B to zero and never reads A; recorded as a
        // GEMM with k = 0 and beta = 0 (C = 0 without reading C)
        op.op = OP_GEMM;
        op.flags = F_TB;  // the 64 x 64 block path, which skips every operand at k = 0
        op.m = m;
        op.n = n;
        op.k = 0;
        op.alpha = 0.0;
        op.beta = 0.0;
        op.lda = op.ldb = 1;
        op.c = B;
        op.ldc = ldb;
        return MU_OK;
    }
    // Reduce to T Y = alpha Z with T lower triangular:
    //   left:  op(A) X = B      -> T = op(A),    Z = B
    //   right: X op(A) = B      -> T = op(A)^T,  Z = B^T
    // An upper T is handled by reversing the row order (F_REV).
    const bool tt = side == 'L' ? t : !t;  // T = A^T
    const bool t_upper = tt ? (uplo == 'L') : (uplo == 'U');
    op.op = OP_TRSM;
    op.flags = (tt ? F_T_TRANS : 0) | (t_upper ? F_REV : 0) | (side == 'R' ? F_Z_ROWS : 0) |
               (diag == 'U' ? F_UNIT : 0);
    op.m = side == 'L' ? m : n;  // order of the triangle
    op.n = side == 'L' ? n : m;  // right-hand sides
    op.alpha = alpha;
    op.a = A;
    op.lda = lda;
    op.c = B;
    op.ldc = ldb;
    return MU_OK;
}

bool mu::stream_recording(void *stream) { return stream_state(stream).dag != nullptr; }

int mu::gemm_a32(void *stream, int m, int n, int k, double alpha, const float *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc, bool rows)
{
    if (m < 0 || n < 0 || k < 0 || n > 16)
        return fail(MU_EINVAL, "gemm_a32: need m, k >= 0 and 0 <= n <= 16");
    if (lda < std::max(1, m) || ldb < std::max(1, rows ? n : k) ||
        ldc < std::max(1, rows ? n : m))
        return fail(MU_EINVAL, "gemm_a32: leading dimension too small");
    mu::OpRec op;
    op.op = mu::OP_GEMM;
    op.flags = mu::F_A32 | (rows ? mu::F_BROW : 0);
    op.m = m;
    op.n = n;
    op.k = k;
    op.alpha = alpha;
    op.beta = beta;
    op.a = reinterpret_cast<const double *>(A);
    op.lda = lda;
    op.b = B;
    op.ldb = ldb;
    op.c = C;
    op.ldc = ldc;
    return issue(stream, op);
}

int mu::cvt_f32(void *stream, const double *src, int lds, float *dst, int ldd, int m, int n,
                bool transposed, float *plain, int ldp)
{
    if (m < 0 || n < 0 || lds < std::max(1, m) || ldd < std::max(1, transposed ? n : m) ||
        (plain && (!transposed || ldp < std::max(1, m))))
        return fail(MU_EINVAL, "cvt_f32: bad sizes");
    mu::OpRec op;
    op.op = mu::OP_CVT;
    op.flags = mu::F_LOW_F32 | (transposed ? mu::F_CVT_T : 0) | (plain ? mu::F_CVT_BOTH : 0);
    op.a = reinterpret_cast<const double *>(plain);
    op.lda = plain ? ldp : 0;
    op.m = m;
    op.n = n;
    op.c = const_cast<double *>(src);
    op.ldc = lds;
    op.b = reinterpret_cast<const double *>(dst);
    op.ldb = ldd;
    return issue(stream, op);
}

extern "C" {

const char *mu_last_error(void) { return g_err.c_str(); }

int mu_get_device_count(int *count)
{
    if (!count)
        return fail(MU_EINVAL, "count is NULL");
    HIPTRY(hipGetDeviceCount(count));
    return MU_OK;
}

int mu_set_device(int device)
{
    HIPTRY(hipSetDevice(device));
    return MU_OK;
}

int mu_device_props(int device, int *cu_count, int *clock_khz, size_t *lds_per_block,
                    char *name, int name_len)
{
    hipDeviceProp_t p;
    HIPTRY(hipGetDeviceProperties(&p, device));
    if (cu_count)
        *cu_count = p.multiProcessorCount;
    if (clock_khz)
        *clock_khz = p.clockRate;
    if (lds_per_block)
        *lds_per_block = p.sharedMemPerBlock;
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    return MU_OK;
}

int mu_malloc(void **ptr, size_t bytes)
{
    if (!ptr)
        return fail(MU_EINVAL, "ptr is NULL");
    HIPTRY(hipMalloc(ptr, bytes));
    return MU_OK;
}

int mu_free(void *ptr)
{
    HIPTRY(hipFree(ptr));
    return MU_OK;
}

int mu_memcpy(void *dst, const void *src, size_t bytes)
{
    HIPTRY(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return MU_OK;
}

int mu_memcpy_async(void *dst, const void *src, size_t bytes, void *stream)
{
    HIPTRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream));
    return MU_OK;
}

int mu_memset_zero(void *ptr, size_t bytes)
{
    HIPTRY(hipMemset(ptr, 0, bytes));
    return MU_OK;
}

int mu_stream_create(void **stream)
{
    hipStream_t s;
    HIPTRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return MU_OK;
}

int mu_stream_destroy(void *stream)
{
    HIPTRY(hipStreamDestroy((hipStream_t)stream));
    return MU_OK;
}

int mu_stream_synchronize(void *stream)
{
    HIPTRY(hipStreamSynchronize((hipStream_t)stream));
    return MU_OK;
}

int mu_device_synchronize(void)
{
    HIPTRY(hipDeviceSynchronize());
    return MU_OK;
}

int mu_event_create(void **event)
{
    hipEvent_t e;
    HIPTRY(hipEventCreate(&e));
    *event = e;
    return MU_OK;
}

int mu_event_destroy(void *event)
{
    HIPTRY(hipEventDestroy((hipEvent_t)event));
    return MU_OK;
}

int mu_event_record(void *event, void *stream)
{
    HIPTRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return MU_OK;
}

int mu_event_elapsed_ms(float *ms, void *start, void *end)
{
    HIPTRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return MU_OK;
}

int mu_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    HIPTRY(hipMemGetInfo(free_bytes, total_bytes));
    return MU_OK;
}

// ---- tile operations ------------------------------------------------------
int mu_dgemm(void *stream, char transa, char transb, int m, int n, int k, double alpha,
             const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc)
{
    const bool ta = is_t(transa), tb = is_t(transb);
    if ((!ta && !is_n(transa)) || (!tb && !is_n(transb)))
        return fail(MU_EINVAL, "mu_dgemm: trans must be N or T");
    if (m < 0 || n < 0 || k < 0)
        return fail(MU_EINVAL, "mu_dgemm: negative size");
    if (lda < std::max(1, ta ? k : m) || ldb < std::max(1, tb ? n : k) || ldc < std::max(1, m))
        return fail(MU_EINVAL, "mu_dgemm: leading dimension too small");
    mu::OpRec op;
    op.op = mu::OP_GEMM;
    op.flags = (ta ? mu::F_TA : 0) | (tb ? mu::F_TB : 0) | prec_flags(stream_state(stream).prec);
    op.m = m;
    op.n = n;
    op.k = k;
    op.alpha = alpha;
    op.beta = beta;
    op.a = A;
    op.lda = lda;
    op.b = B;
    op.ldb = ldb;
    op.c = C;
    op.ldc = ldc;
    return issue(stream, op);
}

int mu_gemm_a32(void *stream, int m, int n, int k, double alpha, const float *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc, int rows)
{
    return mu::gemm_a32(stream, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rows != 0);
}

int mu_dsyrk(void *stream, char uplo, char trans, int n, int k, double alpha, const double *A,
             int lda, double beta, double *C, int ldc)
{
    if (up(uplo) != 'L' || !is_n(trans))
        return fail(MU_ENOTSUP, "mu_dsyrk: only uplo = L, trans = N");
    if (n < 0 || k < 0 || lda < std::max(1, n) || ldc < std::max(1, n))
        return fail(MU_EINVAL, "mu_dsyrk: bad size or leading dimension");
    mu::OpRec op;
    op.op = mu::OP_SYRK;
    op.flags = prec_flags(stream_state(stream).prec);
    op.m = op.n = n;
    op.k = k;
    op.alpha = alpha;
    op.beta = beta;
    op.a = A;
    op.lda = lda;
    op.c = C;
    op.ldc = ldc;
    return issue(stream, op);
}

int mu_dtrsm(void *stream, char side, char uplo, char transa, char diag, int m, int n,
             double alpha, const double *A, int lda, double *B, int ldb)
{
    mu::OpRec op;
    const int rc = mu::dtrsm_op(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, op);
    return rc != MU_OK ? rc : issue(stream, op);
}

int mu_dgetrf_nopiv(void *stream, int n, double *A, int lda, int *info_dev)
{
    if (n < 0 || lda < std::max(1, n))
        return fail(MU_EINVAL, "mu_dgetrf_nopiv: bad size or leading dimension");
    mu::OpRec op;
    op.op = mu::OP_GETRF;
    op.m = op.n = n;
    op.c = A;
    op.ldc = lda;
    op.info = info_dev;
    op.flags = prec_flags(stream_state(stream).prec);
    return issue(stream, op);
}

int mu_dpotrf(void *stream, char uplo, int n, double *A, int lda, int *info_dev)
{
    if (up(uplo) != 'L')
        return fail(MU_ENOTSUP, "mu_dpotrf: only uplo = L");
    if (n < 0 || lda < std::max(1, n))
        return fail(MU_EINVAL, "mu_dpotrf: bad size or leading dimension");
    mu::OpRec op;
    op.op = mu::OP_POTRF;
    op.m = op.n = n;
    op.c = A;
    op.ldc = lda;
    op.info = info_dev;
    op.flags = prec_flags(stream_state(stream).prec);
    return issue(stream, op);
}

int mu_info_reset(int *info_dev, void *stream)
{
    if (!info_dev)
        return fail(MU_EINVAL, "info is NULL");
    if (mu::fill_int(info_dev, 1, INT_MAX, (hipStream_t)stream) != 0)
        return fail(MU_EHIP, "fill kernel launch failed");
    return MU_OK;
}

int mu_set_update_precision(void *stream, int prec)
{
    if (prec < MU_PREC_FP64 || prec > MU_PREC_BF16)
        return fail(MU_EINVAL, "unknown precision");
    std::lock_guard<std::mutex> lk(g_mu);
    g_streams[stream].prec = prec;
    return MU_OK;
}

// ---- DAG ----------------------------------------------------------------------
int mu_dag_create(mu_dag_t *dag)
{
    if (!dag)
        return fail(MU_EINVAL, "dag is NULL");
    *dag = new mu_dag_s();
    return MU_OK;
}

int mu_dag_destroy(mu_dag_t dag)
{
    if (!dag)
        return MU_OK;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (auto &kv : g_streams) {
            auto &v = kv.second.dags;
            v.erase(std::remove(v.begin(), v.end(), dag), v.end());
        }
    }
    delete dag;  // its buffers go with the last executable that still uses them
    return MU_OK;
}

int mu_dag_capture_begin(mu_dag_t dag, void *stream)
{
    if (!dag)
        return fail(MU_EINVAL, "dag is NULL");
    std::lock_guard<std::mutex> lk(g_mu);
    auto &v = g_streams[stream].dags;
    if (std::find(v.begin(), v.end(), dag) != v.end())
        return fail(MU_ESTATE, "the DAG is already recording this stream");
    v.push_back(dag);
    return MU_OK;
}

int mu_dag_capture_end(mu_dag_t dag, void *stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_streams.find(stream);
    if (it == g_streams.end() || it->second.dags.empty() || it->second.dags.back() != dag)
        return fail(MU_ESTATE, "stream is not being recorded by this DAG (or a sub-graph "
                               "recording on it has not ended)");
    it->second.dags.pop_back();
    return MU_OK;
}

int mu_dag_begin_task(mu_dag_t dag, int w_row, int w_col, const int *read_tiles, int nread,
                      int *task)
{
    if (!dag || nread < 0 || (nread > 0 && !read_tiles))
        return fail(MU_EINVAL, "mu_dag_begin_task: bad arguments");
    std::vector<mu::Tile> r;
    for (int i = 0; i < nread; i++)
        r.push_back({read_tiles[2 * i], read_tiles[2 * i + 1]});
    int t = dag->dag.begin_task({w_row, w_col}, r);
    if (t < 0)
        return fail(MU_ESTATE, dag->dag.err);
    if (task)
        *task = t;
    return MU_OK;
}

int mu_dag_end_task(mu_dag_t dag)
{
    if (!dag)
        return fail(MU_EINVAL, "dag is NULL");
    if (dag->dag.end_task() != 0)
        return fail(MU_ESTATE, dag->dag.err);
    return MU_OK;
}

int mu_dag_num_tasks(mu_dag_t dag, int *ntasks)
{
    if (!dag || !ntasks)
        return fail(MU_EINVAL, "bad arguments");
    *ntasks = (int)dag->dag.tasks.size();
    return MU_OK;
}

int mu_dag_task_deps(mu_dag_t dag, int task, int *deps, int max, int *ndeps)
{
    if (!dag || task < 0 || task >= (int)dag->dag.tasks.size() || !ndeps)
        return fail(MU_EINVAL, "bad arguments");
    const auto &d = dag->dag.tasks[task].deps;
    *ndeps = (int)d.size();
    for (int i = 0; i < (int)d.size() && i < max && deps; i++)
        deps[i] = d[i];
    return MU_OK;
}

int mu_dag_add_edge(mu_dag_t dag, int src, int dst)
{
    if (!dag || src < 0 || dst <= src || dst >= (int)dag->dag.tasks.size())
        return fail(MU_EINVAL, "mu_dag_add_edge: edge must go from an earlier to a later task");
    dag->dag.add_edge(src, dst);
    return MU_OK;
}

int mu_dag_insert_subdag(mu_dag_t dag, mu_dag_t child)
{
    if (!dag || !child || dag == child)
        return fail(MU_EINVAL, "mu_dag_insert_subdag: bad arguments");
    if (dag->dag.insert_sub(child->dag) != 0)
        return fail(MU_ESTATE, dag->dag.err);
    dag->owned->sub.push_back(child->owned);
    return MU_OK;
}

int mu_dag_task_info(mu_dag_t dag, int task, int *w_row, int *w_col, int *ops, int max_ops,
                     int *nops)
{
    if (!dag || task < 0 || task >= (int)dag->dag.tasks.size())
        return fail(MU_EINVAL, "mu_dag_task_info: bad task");
    const mu::TaskRec &t = dag->dag.tasks[task];
    if (w_row)
        *w_row = t.w.first;
    if (w_col)
        *w_col = t.w.second;
    if (nops)
        *nops = (int)t.ops.size();
    for (int i = 0; ops && i < (int)t.ops.size() && i < max_ops; i++)
        ops[i] = t.ops[i].op;
    return MU_OK;
}

int mu_dag_set_tag(mu_dag_t dag, int tag)
{
    if (!dag)
        return fail(MU_EINVAL, "dag is NULL");
    dag->dag.set_tag(tag);
    return MU_OK;
}

// ---- executables --------------------------------------------------------------
int mu_exec_create(mu_dag_t dag, int mode, const mu_exec_opts_t *opts, mu_exec_t *exec)
{
    if (!dag || !exec || mode < MU_MODE_TILED || mode > MU_MODE_STREAM)
        return fail(MU_EINVAL, "mu_exec_create: bad arguments");
    if (dag->dag.current() >= 0)
        return fail(MU_ESTATE, "mu_exec_create: a task is still being recorded");
    mu::LowerOpts lo;
    mu::ExecOpts eo;
    if (opts) {
        if (opts->small_max > 0)
            lo.small_max = std::min(opts->small_max, mu::SMALL_MAX);
        if (opts->trsm_max > 0)
            lo.trsm_max = std::min(opts->trsm_max, mu::TRSM_NMAX);
        if (opts->timeout_s > 0)
            eo.timeout_s = opts->timeout_s;
        eo.grid = opts->grid;
        eo.trace = opts->trace;
        eo.lookahead = opts->lookahead < 0 ? 0 : opts->lookahead == 2 ? 2 : 1;
        if (opts->sm < 0 || opts->sm > 100)
            return fail(MU_EINVAL, "mu_exec_create: sm is a percentage of the GPU (0 = none)");
        eo.sm = opts->sm;
        eo.fuse = opts->no_fuse ? 0 : 1;
    }
    // the persistent kernel runs panel solves of order <= 256 (larger ones nest)
    // and GEMM items of 2 x 2 output blocks
    if (mode == MU_MODE_SUBGRAPH) {
        lo.trsm_max = std::min(lo.trsm_max, mu::PERSIST_TRSM_MAX);
        lo.gemm_group = 2;
    }
    for (const auto &t : dag->dag.tasks)
        for (const auto &op : t.ops)
            if (op.flags & (mu::F_LOW_F32 | mu::F_LOW_BF16))
                if (op.op != mu::OP_GEMM && op.op != mu::OP_SYRK && op.op != mu::OP_GETRF &&
                    op.op != mu::OP_POTRF && op.op != mu::OP_CVT)
                    return fail(MU_ENOTSUP, "reduced precision applies to GEMM/SYRK only");
    mu::Exec *e = nullptr;
    std::string err;
    if (mu::exec_create(dag->dag, mode, lo, eo, &e, err) != 0)
        return fail(MU_EHIP, err);
    *exec = new mu_exec_s();
    (*exec)->e = e;
    (*exec)->owned = dag->owned;
    return MU_OK;
}

int mu_exec_launch(mu_exec_t exec, void *stream)
{
    if (!exec)
        return fail(MU_EINVAL, "exec is NULL");
    std::string err;
    if (mu::exec_launch(exec->e, (hipStream_t)stream, err) != 0)
        return fail(MU_EHIP, err);
    return MU_OK;
}

int mu_exec_launch_level(mu_exec_t exec, int level, void *stream)
{
    if (!exec)
        return fail(MU_EINVAL, "exec is NULL");
    std::string err;
    if (mu::exec_launch_level(exec->e, level, (hipStream_t)stream, err) != 0)
        return fail(MU_EINVAL, err);
    return MU_OK;
}

int mu_exec_level_items(mu_exec_t exec, int level, long long *items)
{
    if (!exec || !items || level < 0 || level >= exec->e->levels)
        return fail(MU_EINVAL, "mu_exec_level_items: bad arguments");
    *items = (long long)(exec->e->level_off[level + 1] - exec->e->level_off[level]);
    return MU_OK;
}

int mu_exec_level_ops(mu_exec_t exec, int level, long long *counts)
{
    const mu::Exec *e = exec ? exec->e : nullptr;
    if (!e || !counts || level < 0 || level >= e->levels)
        return fail(MU_EINVAL, "mu_exec_level_ops: bad arguments");
    for (int i = 0; i < 8; i++)
        counts[i] = 0;
    // the graph's unit table (lowered units, then merged ones) is numbered in
    // item order: find the unit of each item
    const std::vector<mu::UnitDesc> &U = e->gunits.empty() ? e->low.units : e->gunits;
    for (size_t k = e->level_off[level]; k < e->level_off[level + 1]; k++) {
        const uint32_t item = e->level_items[k];
        size_t lo = 0, hi = U.size();
        while (hi - lo > 1) {
            size_t mid = (lo + hi) / 2;
            if (U[mid].item_base <= item)
                lo = mid;
            else
                hi = mid;
        }
        counts[U[lo].op % 8]++;
    }
    return MU_OK;
}

int mu_exec_level_info(mu_exec_t exec, int level, int *kernel, double *flops)
{
    const mu::Exec *e = exec ? exec->e : nullptr;
    if (!e || level < 0 || level >= e->levels)
        return fail(MU_EINVAL, "mu_exec_level_info: tiled executables only, level out of range");
    if (kernel)
        *kernel = e->level_gemm[level] == mu::LEVEL_GEMM    ? MU_KERNEL_GEMM
                  : e->level_gemm[level] == mu::LEVEL_SMALL ? MU_KERNEL_SMALL
                  : e->level_gemm[level] == mu::LEVEL_TRSM  ? MU_KERNEL_TRSM
                  : e->level_gemm[level] == mu::LEVEL_TRSM_SMALL ? MU_KERNEL_TRSM_SMALL
                  : e->level_gemm[level] == mu::LEVEL_TRSM_256   ? MU_KERNEL_TRSM_256
                  : e->level_gemm[level] == mu::LEVEL_TRSM_512   ? MU_KERNEL_TRSM_512
                  : e->level_gemm[level] == mu::LEVEL_GEMM_NARROW ? MU_KERNEL_GEMM_NARROW
                  : e->level_gemm[level] == mu::LEVEL_GEMM_LOW    ? MU_KERNEL_GEMM_LOW
                  : e->level_gemm[level] == mu::LEVEL_CVT         ? MU_KERNEL_CVT
                                                                 : MU_KERNEL_GENERIC;
    if (flops)
        *flops = e->level_flops[level];
    return MU_OK;
}

int mu_exec_check(mu_exec_t exec)
{
    if (!exec)
        return fail(MU_EINVAL, "exec is NULL");
    std::string err;
    if (mu::exec_check(exec->e, err) != 0)
        return fail(MU_ERUN, err);
    return MU_OK;
}

int mu_exec_stats(mu_exec_t exec, mu_exec_stats_t *s)
{
    if (!exec || !s)
        return fail(MU_EINVAL, "bad arguments");
    const mu::Exec *e = exec->e;
    std::memset(s, 0, sizeof *s);
    s->mode = e->mode;
    s->grid = e->grid;
    s->units = (long long)e->low.units.size();
    s->items = (long long)e->low.nitems;
    for (const auto &v : e->low.succ)
        s->edges += (long long)v.size();
    s->flops = e->low.flops;
    s->critical_path_us = e->low.critical_path;
    s->levels = e->levels;
    s->depth = e->depth;
    for (const auto &u : e->low.units)
        s->priority_units[u.prio & 3]++;
    s->wall_clock_khz = (double)e->wall_khz;
    return MU_OK;
}

int mu_exec_trace(mu_exec_t exec, unsigned long long *stamps, size_t count)
{
    if (!exec || !stamps)
        return fail(MU_EINVAL, "bad arguments");
    std::vector<uint64_t> v;
    std::string err;
    if (mu::exec_trace(exec->e, v, err) != 0)
        return fail(MU_ESTATE, err);
    if (count < v.size())
        return fail(MU_EINVAL, "stamps buffer too small");
    for (size_t i = 0; i < v.size(); i++)
        stamps[i] = v[i];
    return MU_OK;
}

int mu_exec_item_units(mu_exec_t exec, int *unit_of_item, int *task_of_unit, int *op_of_unit)
{
    if (!exec)
        return fail(MU_EINVAL, "exec is NULL");
    const mu::Lowered &L = exec->e->low;
    for (size_t u = 0; u < L.units.size(); u++) {
        if (task_of_unit)
            task_of_unit[u] = L.units[u].task;
        if (op_of_unit)
            op_of_unit[u] = L.units[u].op;
        if (unit_of_item)
            for (int i = 0; i < L.units[u].nitems; i++)
                unit_of_item[L.units[u].item_base + i] = (int)u;
    }
    return MU_OK;
}

int mu_exec_destroy(mu_exec_t exec)
{
    if (!exec)
        return MU_OK;
    mu::exec_destroy(exec->e);
    delete exec;
    return MU_OK;
}

static int emulate_into(mu_dag_t dag, const mu::EmuOpts &eo, long long *results, int nres)
{
    mu::EmuResult r;
    std::string err;
    mu::LowerOpts lo;
    lo.trsm_max = std::min(lo.trsm_max, mu::PERSIST_TRSM_MAX);
    lo.gemm_group = 2;
    if (!mu::emulate(dag->dag, lo, eo, r, err))
        return fail(MU_EINVAL, err);
    const long long v[9] = {(long long)r.executed, (long long)r.duplicates,
                            (long long)r.order_violations, (long long)r.err,
                            (long long)r.done, (long long)r.nitems,
                            (long long)r.max_task_workgroups, (long long)r.task_cap,
                            (long long)r.max_tasks_running};
    for (int i = 0; i < nres; i++)
        results[i] = v[i];
    return MU_OK;
}

int mu_dag_emulate(mu_dag_t dag, int threads, double timeout_s, int bump_unit,
                   long long *results)
{
    if (!dag || !results || threads < 1 || timeout_s <= 0)
        return fail(MU_EINVAL, "mu_dag_emulate: bad arguments");
    mu::EmuOpts eo;
    eo.workgroups = threads;
    eo.timeout_s = timeout_s;
    eo.bump_unit = bump_unit;
    return emulate_into(dag, eo, results, 6);
}

int mu_dag_emulate_ex(mu_dag_t dag, const mu_emu_opts_t *opts, long long *results)
{
    if (!dag || !opts || !results || opts->workgroups < 1 || opts->timeout_s <= 0 ||
        opts->sm < 0 || opts->sm > 100 || opts->producer_delay_us < 0)
        return fail(MU_EINVAL, "mu_dag_emulate_ex: bad arguments");
    mu::EmuOpts eo;
    eo.workgroups = opts->workgroups;
    eo.timeout_s = opts->timeout_s;
    eo.host_deadline_s = opts->host_deadline_s;
    eo.bump_unit = opts->bump_unit;
    eo.producer_delay_us = opts->producer_delay_us;
    eo.sm = opts->sm;
    return emulate_into(dag, eo, results, 9);
}

// ---- tiled factorizations -----------------------------------------------------
// Reduced-precision factorizations keep a shadow of the matrix in the update
// precision: every panel tile is rounded into it once, right after its TRSM
// (an OP_CVT op in the same task), and the trailing GEMM/SYRK read their
// operands from the shadow (F_SRC_LOW) -- half (FP32) or a quarter (BF16) of
// the FP64 operand traffic.  C stays FP64.
namespace {
// Reduced-precision copies of the solved panel tiles, read by the trailing
// updates.  Row-panel tiles (LU's U_kj, the B operand) keep the matrix's
// layout: tile (i, j) at at(i, j), leading dimension n.  Column-panel tiles
// (L_ik, the A operand; for Cholesky also B^T) are stored transposed, panel
// by panel: element (global row R, k) of column panel k at
// pt + k n b + R b + k, leading dimension b.  Both operands of an update are
// then k-contiguous, and the tiles of one panel follow each other the way
// one merged update reads them.
struct Shadow {
    char *p = nullptr, *pt = nullptr;
    size_t elem = 0;
    int ld = 0, b = 0;
    const double *at(int i, int j) const
    {
        return reinterpret_cast<const double *>(p + ((size_t)i * b + (size_t)j * b * ld) * elem);
    }
    const double *at_t(int i, int k) const
    {
        return reinterpret_cast<const double *>(pt + ((size_t)k * ld * b + (size_t)i * b * b) *
                                                         elem);
    }
};

int make_shadow(mu_dag_t dag, int n, int b, int prec, bool rows, Shadow &sh)
{
    if (prec == MU_PREC_FP64)
        return MU_OK;
    sh.elem = prec == MU_PREC_FP32 ? 4 : 2;
    sh.ld = n;
    sh.b = b;
    void *p = nullptr;
    if (rows) {
        HIPTRY(hipMalloc(&p, (size_t)n * n * sh.elem));
        dag->owned->p.push_back(p);
        sh.p = (char *)p;
    }
    HIPTRY(hipMalloc(&p, (size_t)n * n * sh.elem));
    dag->owned->p.push_back(p);
    sh.pt = (char *)p;
    return MU_OK;
}

// Conversion of tile (i, j) to its shadow: a row-panel tile as it is, a
// column-panel tile (transposed) into its panel's transposed storage.
int issue_cvt(void *key, const double *src, int lds, const Shadow &sh, int i, int j, int prec,
              bool transposed)
{
    mu::OpRec op;
    op.op = mu::OP_CVT;
    op.flags = prec_flags(prec) | (transposed ? mu::F_CVT_T : 0);
    op.m = op.n = sh.b;
    op.c = const_cast<double *>(src);
    op.ldc = lds;
    op.b = transposed ? sh.at_t(i, j) : sh.at(i, j);
    op.ldb = transposed ? sh.b : sh.ld;
    return issue(key, op);
}

// C -= op(A) op(B) from the shadows: op(A) = the column-panel tile (ai, k)
// read from its transposed storage (F_TA); op(B) = the row-panel tile
// (k, bj) (LU, tb false) or the transpose of column-panel tile (bi, k)
// (Cholesky, tb true), both k-contiguous.
int issue_gemm_low(void *key, bool tb, const Shadow &sh, int ai, int aj, int bi, int bj,
                   double *C, int ldc, int prec)
{
    mu::OpRec op;
    op.op = mu::OP_GEMM;
    op.flags = mu::F_TA | prec_flags(prec) | mu::F_SRC_LOW;
    op.m = op.n = op.k = sh.b;
    op.alpha = -1.0;
    op.beta = 1.0;
    op.a = sh.at_t(ai, aj);
    op.lda = sh.b;
    op.b = tb ? sh.at_t(bi, bj) : sh.at(bi, bj);
    op.ldb = tb ? sh.b : sh.ld;
    op.c = C;
    op.ldc = ldc;
    return issue(key, op);
}

int issue_syrk_low(void *key, const Shadow &sh, int ai, int aj, double *C, int ldc, int prec)
{
    mu::OpRec op;
    op.op = mu::OP_SYRK;
    op.flags = mu::F_TA | prec_flags(prec) | mu::F_SRC_LOW;
    op.m = op.n = op.k = sh.b;
    op.alpha = -1.0;
    op.beta = 1.0;
    op.a = sh.at_t(ai, aj);
    op.lda = sh.b;
    op.c = C;
    op.ldc = ldc;
    return issue(key, op);
}
}  // namespace

int mu_record_tiled_lu(mu_dag_t dag, double *A, int n, int t, int lda, int *info_dev, int prec)
{
    if (!dag || !A || n <= 0 || t <= 0 || n % t != 0 || lda < n)
        return fail(MU_EINVAL, "mu_record_tiled_lu: need n > 0, t > 0, n % t == 0, lda >= n");
    const int b = n / t;
    void *key = dag;  // record through the same path as a captured stream
    int rc;
    Shadow sh;
    if ((rc = make_shadow(dag, n, b, prec, true, sh)) != 0)
        return rc;
    const bool low = sh.pt != nullptr;
    if ((rc = mu_dag_capture_begin(dag, key)) != 0)
        return rc;
    auto blk = [&](int i, int j) { return A + (size_t)i * b + (size_t)j * b * lda; };
    auto rd = [](std::initializer_list<int> v) { return std::vector<int>(v); };
    // the shadow of tile (i, j) is written by a task of its own, tile key
    // (i + t, j), outside the matrix's keys; L tiles (below the diagonal)
    // are stored transposed
    auto record_cvt = [&](int i, int j) {
        const int r1[2] = {i, j};
        int c = mu_dag_begin_task(dag, i + t, j, r1, 1, nullptr);
        if (!c)
            c = issue_cvt(key, blk(i, j), lda, sh, i, j, prec, i > j);
        if (!c)
            c = mu_dag_end_task(dag);
        return c;
    };
    // lu_mustard.cu:343-535
    for (int k = 0; k < t && rc == 0; k++) {
        auto r0 = rd({k, k});
        rc = mu_dag_begin_task(dag, k, k, r0.data(), 1, nullptr);
        if (!rc) {
            mu_set_update_precision(key, MU_PREC_FP64);
            rc = mu_dgetrf_nopiv(key, b, blk(k, k), lda, info_dev);
            if (!rc && !dag->dag.tasks.back().ops.empty())
                dag->dag.tasks.back().ops.back().info_base = k * b;
        }
        if (!rc)
            rc = mu_dag_end_task(dag);
        for (int i = k + 1; i < t && rc == 0; i++) {
            auto r = rd({k, k, k, i});
            rc = mu_dag_begin_task(dag, k, i, r.data(), 2, nullptr);
            if (!rc)
                rc = mu_dtrsm(key, 'L', 'L', 'N', 'U', b, b, 1.0, blk(k, k), lda, blk(k, i), lda);
            if (!rc)
                rc = mu_dag_end_task(dag);
            if (!rc && low)
                rc = record_cvt(k, i);
        }
        for (int i = k + 1; i < t && rc == 0; i++) {
            auto r = rd({k, k, i, k});
            rc = mu_dag_begin_task(dag, i, k, r.data(), 2, nullptr);
            if (!rc)
                rc = mu_dtrsm(key, 'R', 'U', 'N', 'N', b, b, 1.0, blk(k, k), lda, blk(i, k), lda);
            if (!rc)
                rc = mu_dag_end_task(dag);
            if (!rc && low)
                rc = record_cvt(i, k);
            for (int j = k + 1; j < t && rc == 0; j++) {
                // with shadows, every update reads the converted panel tiles
                // (the conversions of panel k run beside the next panel's
                // chain, so its updates wait only for them)
                auto r3 = low ? rd({i, j, i + t, k, k + t, j}) : rd({i, k, k, j, i, j});
                rc = mu_dag_begin_task(dag, i, j, r3.data(), 3, nullptr);
                if (!rc && low)
                    rc = issue_gemm_low(key, false, sh, i, k, k, j, blk(i, j), lda, prec);
                else if (!rc)
                    rc = mu_dgemm(key, 'N', 'N', b, b, b, -1.0, blk(i, k), lda, blk(k, j), lda,
                                  1.0, blk(i, j), lda);
                if (!rc)
                    rc = mu_dag_end_task(dag);
            }
        }
    }
    int rc2 = mu_dag_capture_end(dag, key);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_streams.erase(key);
    }
    return rc ? rc : rc2;
}

int mu_record_tiled_cholesky(mu_dag_t dag, double *A, int n, int t, int lda, int *info_dev,
                             int prec)
{
    if (!dag || !A || n <= 0 || t <= 0 || n % t != 0 || lda < n)
        return fail(MU_EINVAL,
                    "mu_record_tiled_cholesky: need n > 0, t > 0, n % t == 0, lda >= n");
    const int b = n / t;
    void *key = dag;
    int rc;
    Shadow sh;
    if ((rc = make_shadow(dag, n, b, prec, false, sh)) != 0)
        return rc;
    const bool low = sh.pt != nullptr;
    if ((rc = mu_dag_capture_begin(dag, key)) != 0)
        return rc;
    auto blk = [&](int i, int j) { return A + (size_t)i * b + (size_t)j * b * lda; };
    auto rd = [](std::initializer_list<int> v) { return std::vector<int>(v); };
    // the panel tiles' shadows, stored transposed
    auto record_cvt = [&](int i, int j) {
        const int r1[2] = {i, j};
        int c = mu_dag_begin_task(dag, i + t, j, r1, 1, nullptr);
        if (!c)
            c = issue_cvt(key, blk(i, j), lda, sh, i, j, prec, true);
        if (!c)
            c = mu_dag_end_task(dag);
        return c;
    };
    // cholesky_mustard.cu:221-400
    for (int k = 0; k < t && rc == 0; k++) {
        auto r0 = rd({k, k});
        rc = mu_dag_begin_task(dag, k, k, r0.data(), 1, nullptr);
        if (!rc) {
            rc = mu_dpotrf(key, 'L', b, blk(k, k), lda, info_dev);
            if (!rc && !dag->dag.tasks.back().ops.empty())
                dag->dag.tasks.back().ops.back().info_base = k * b;
        }
        if (!rc)
            rc = mu_dag_end_task(dag);
        for (int i = k + 1; i < t && rc == 0; i++) {
            auto r = rd({k, k, i, k});
            rc = mu_dag_begin_task(dag, i, k, r.data(), 2, nullptr);
            if (!rc)
                rc = mu_dtrsm(key, 'R', 'L', 'T', 'N', b, b, 1.0, blk(k, k), lda, blk(i, k), lda);
            if (!rc)
                rc = mu_dag_end_task(dag);
            if (!rc && low)
                rc = record_cvt(i, k);
        }
        for (int i = k + 1; i < t && rc == 0; i++) {
            // with shadows, every update reads the converted panel tiles
            auto r = low ? rd({i, i, i + t, k}) : rd({i, i, i, k});
            rc = mu_dag_begin_task(dag, i, i, r.data(), 2, nullptr);
            if (!rc && low)
                rc = issue_syrk_low(key, sh, i, k, blk(i, i), lda, prec);
            else if (!rc)
                rc = mu_dsyrk(key, 'L', 'N', b, b, -1.0, blk(i, k), lda, 1.0, blk(i, i), lda);
            if (!rc)
                rc = mu_dag_end_task(dag);
            for (int j = i + 1; j < t && rc == 0; j++) {
                auto r3 = low ? rd({j, i, j + t, k, i + t, k}) : rd({j, i, j, k, i, k});
                rc = mu_dag_begin_task(dag, j, i, r3.data(), 3, nullptr);
                if (!rc && low)
                    rc = issue_gemm_low(key, true, sh, j, k, i, k, blk(j, i), lda, prec);
                else if (!rc)
                    rc = mu_dgemm(key, 'N', 'T', b, b, b, -1.0, blk(j, k), lda, blk(i, k), lda,
                                  1.0, blk(j, i), lda);
                if (!rc)
                    rc = mu_dag_end_task(dag);
            }
        }
    }
    int rc2 = mu_dag_capture_end(dag, key);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_streams.erase(key);
    }
    return rc ? rc : rc2;
}

int mu_record_tiled_solve(mu_dag_t dag, const double *F, int n, int t, int ldf, int kind,
                          double *X, int ldx, int nrhs)
{
    if (!dag || !F || !X || n <= 0 || t <= 0 || n % t != 0 || ldf < n || ldx < n || nrhs <= 0 ||
        (kind != MU_SOLVE_LU && kind != MU_SOLVE_CHOL))
        return fail(MU_EINVAL, "mu_record_tiled_solve: need n % t == 0, ldf, ldx >= n, nrhs > 0, "
                               "kind MU_SOLVE_LU or MU_SOLVE_CHOL");
    const int b = n / t;
    void *key = (char *)dag + 1;  // a recording key distinct from the factorization's
    int rc;
    if ((rc = mu_dag_capture_begin(dag, key)) != 0)
        return rc;
    double *Fm = const_cast<double *>(F);
    auto fb = [&](int i, int j) { return Fm + (size_t)i * b + (size_t)j * b * ldf; };
    auto xb = [&](int i) { return X + (size_t)i * b; };
    // right-hand-side block i is tile (i, t): it never aliases a factor tile
    auto task = [&](int w, std::initializer_list<int> reads) {
        std::vector<int> r(reads);
        return mu_dag_begin_task(dag, w, t, r.data(), (int)r.size() / 2, nullptr);
    };
    // forward substitution: L Y = B (L unit lower for LU, lower for Cholesky)
    const char fdiag = kind == MU_SOLVE_LU ? 'U' : 'N';
    for (int k = 0; k < t && rc == 0; k++) {
        rc = task(k, {k, k, k, t});
        if (!rc)
            rc = mu_dtrsm(key, 'L', 'L', 'N', fdiag, b, nrhs, 1.0, fb(k, k), ldf, xb(k), ldx);
        if (!rc)
            rc = mu_dag_end_task(dag);
        for (int i = k + 1; i < t && rc == 0; i++) {
            rc = task(i, {i, k, k, t, i, t});
            if (!rc)
                rc = mu_dgemm(key, 'N', 'N', b, nrhs, b, -1.0, fb(i, k), ldf, xb(k), ldx, 1.0,
                              xb(i), ldx);
            if (!rc)
                rc = mu_dag_end_task(dag);
        }
    }
    // back substitution: U X = Y (LU) or L^T X = Y (Cholesky)
    for (int k = t - 1; k >= 0 && rc == 0; k--) {
        rc = task(k, {k, k, k, t});
        if (!rc)
            rc = kind == MU_SOLVE_LU
                     ? mu_dtrsm(key, 'L', 'U', 'N', 'N', b, nrhs, 1.0, fb(k, k), ldf, xb(k), ldx)
                     : mu_dtrsm(key, 'L', 'L', 'T', 'N', b, nrhs, 1.0, fb(k, k), ldf, xb(k), ldx);
        if (!rc)
            rc = mu_dag_end_task(dag);
        for (int i = 0; i < k && rc == 0; i++) {
            rc = kind == MU_SOLVE_LU ? task(i, {i, k, k, t, i, t}) : task(i, {k, i, k, t, i, t});
            if (!rc)
                rc = kind == MU_SOLVE_LU
                         ? mu_dgemm(key, 'N', 'N', b, nrhs, b, -1.0, fb(i, k), ldf, xb(k), ldx,
                                    1.0, xb(i), ldx)
                         : mu_dgemm(key, 'T', 'N', b, nrhs, b, -1.0, fb(k, i), ldf, xb(k), ldx,
                                    1.0, xb(i), ldx);
            if (!rc)
                rc = mu_dag_end_task(dag);
        }
    }
    int rc2 = mu_dag_capture_end(dag, key);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_streams.erase(key);
    }
    return rc ? rc : rc2;
}

int mu_record_batch(mu_dag_t dag, int nprob, const int *kinds, const int *ns, double *const *As,
                    const int *ldas, const int *tiles, int *info_dev, int prec)
{
    if (!dag || nprob < 0 || (nprob > 0 && (!kinds || !ns || !As || !ldas || !tiles)))
        return fail(MU_EINVAL, "mu_record_batch: bad arguments");
    for (int p = 0; p < nprob; p++) {
        if (kinds[p] != MU_SOLVE_LU && kinds[p] != MU_SOLVE_CHOL)
            return fail(MU_EINVAL, "mu_record_batch: problem " + std::to_string(p) +
                                       ": kind must be MU_SOLVE_LU or MU_SOLVE_CHOL");
        int rc = mu_dag_set_tag(dag, p + 1);  // tiles of different problems never alias
        if (rc)
            return rc;
        rc = kinds[p] == MU_SOLVE_LU
                 ? mu_record_tiled_lu(dag, As[p], ns[p], tiles[p], ldas[p],
                                      info_dev ? info_dev + p : nullptr, prec)
                 : mu_record_tiled_cholesky(dag, As[p], ns[p], tiles[p], ldas[p],
                                            info_dev ? info_dev + p : nullptr, prec);
        if (rc)
            return fail(rc, "mu_record_batch: problem " + std::to_string(p) + ": " + g_err);
    }
    return mu_dag_set_tag(dag, 0);
}

int mu_batch_residual(int nprob, const int *ns, const long long *a_off, const long long *v_off,
                      const double *A, const double *X, const double *B, double *R,
                      double *norms, void *stream)
{
    if (nprob < 0 || (nprob > 0 && (!ns || !a_off || !v_off || !A || !X || !B || !R || !norms)))
        return fail(MU_EINVAL, "mu_batch_residual: bad arguments");
    if (mu::batch_residual(nprob, ns, a_off, v_off, A, X, B, R, norms, (hipStream_t)stream) != 0)
        return fail(MU_EHIP, "mu_batch_residual: kernel launch failed");
    return MU_OK;
}

// ---- generators -------------------------------------------------------------
void mu_gen_ref_host(double *h_A, size_t n)
{
    // include/gen.h:23-38: same seed, same loop order, same int indices
    srand(420);
    for (int i = 0; i < (int)n; i++)
        for (int j = i; j < (int)n; j++)
            h_A[(size_t)i * n + j] = (double)rand() / (double)RAND_MAX;
    for (int i = 0; i < (int)n; i++)
        for (int j = i; j >= 0; j--)
            h_A[(size_t)i * n + j] = h_A[(size_t)j * n + i];
    for (int i = 0; i < (int)n; i++)
        h_A[(size_t)i * n + i] = h_A[(size_t)i * n + i] + (double)n;
}

int mu_gen_uniform(double *M, size_t count, uint64_t seed, double lo, double hi, void *stream)
{
    if (!M)
        return fail(MU_EINVAL, "M is NULL");
    if (mu::gen_uniform_device(M, count, seed, lo, hi, (hipStream_t)stream) != 0)
        return fail(MU_EHIP, "generator launch failed");
    return MU_OK;
}

int mu_gen_spd(double *A, int n, int lda, uint64_t seed, double *work, void *stream)
{
    if (!A || !work || n <= 0 || lda < n)
        return fail(MU_EINVAL, "mu_gen_spd: bad arguments");
    int rc = mu_gen_uniform(work, (size_t)n * n, seed, -1.0, 1.0, stream);
    if (rc)
        return rc;
    rc = mu_dgemm(stream, 'N', 'T', n, n, n, 1.0, work, n, work, n, 0.0, A, lda);
    if (rc)
        return rc;
    if (mu::add_diag_scale(A, n, lda, (double)n, (double)n, (hipStream_t)stream) != 0 ||
        mu::symmetrize_lower(A, n, lda, (hipStream_t)stream) != 0)
        return fail(MU_EHIP, "kernel launch failed");
    return MU_OK;
}

int mu_gen_cond(double *A, int n, int lda, double kappa, int nrefl, uint64_t seed, double *work,
                void *stream)
{
    if (!A || !work || n <= 0 || lda < n || !(kappa >= 1.0) || nrefl < 0)
        return fail(MU_EINVAL, "mu_gen_cond: bad arguments (kappa >= 1, nrefl >= 0)");
    hipStream_t s = (hipStream_t)stream;
    double *v = work, *w = work + n, *sc = work + 2 * (size_t)n;
    if (mu::gen_cond_start(A, n, lda, kappa, s) != 0)
        return fail(MU_EHIP, "kernel launch failed");
    for (int r = 0; r < nrefl; r++) {
        int rc = mu_gen_uniform(v, (size_t)n, seed + (uint64_t)r, -1.0, 1.0, stream);
        if (rc)
            return rc;
        // w = A v (the current, symmetric A), sc = (v.v, v.w)
        rc = mu_dgemm(stream, 'N', 'N', n, 1, n, 1.0, A, lda, v, n, 0.0, w, n);
        if (rc)
            return rc;
        HIPTRY(hipMemsetAsync(sc, 0, 2 * sizeof(double), s));
        if (mu::dot_acc(v, v, n, sc, s) != 0 || mu::dot_acc(v, w, n, sc + 1, s) != 0 ||
            mu::reflect2(A, lda, n, v, w, sc, s) != 0)
            return fail(MU_EHIP, "kernel launch failed");
    }
    return MU_OK;
}

// ---- verification -----------------------------------------------------------
double mu_verify_lu_l1_host(const double *A, const double *F, int n)
{
    // include/verify.h:17-71: row-major unit-lower L and upper U, L*U by
    // triple loop, sum |A - LU|.
    const size_t nn = (size_t)n * n;
    std::vector<double> L(F, F + nn), U(nn, 0.0), P(nn, 0.0);
    for (int i = 0; i < n; i++) {
        for (int j = i; j < n; j++) {
            std::swap(L[i + (size_t)j * n], L[(size_t)i * n + j]);
            U[(size_t)i * n + j] = L[(size_t)i * n + j];
            L[(size_t)i * n + j] = 0;
        }
        L[(size_t)i * n + i] = 1;
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++)
                P[(size_t)i * n + j] += L[(size_t)i * n + k] * U[(size_t)k * n + j];
    double e = 0;
    for (size_t x = 0; x < nn; x++)
        e += std::fabs(A[x] - P[x]);
    return e;
}

double mu_verify_chol_l1_host(const double *A, const double *F, int n)
{
    // include/verify.h:79-129
    const size_t nn = (size_t)n * n;
    std::vector<double> L(F, F + nn), P(nn, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            L[i + (size_t)j * n] = 0;
            std::swap(L[i + (size_t)j * n], L[(size_t)i * n + j]);
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++)
                P[(size_t)i * n + j] += L[(size_t)i * n + k] * L[k + (size_t)j * n];
    double e = 0;
    for (size_t x = 0; x < nn; x++)
        e += std::fabs(A[x] - P[x]);
    return e;
}

int mu_residual_norms(const double *A, const double *F, int n, int lda, int chol, double *work,
                      double *frob_res, double *frob_a, double *l1_res, void *stream)
{
    if (!A || !F || !work || n <= 0 || lda < n)
        return fail(MU_EINVAL, "mu_residual_norms: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const size_t nn = (size_t)n * n;
    double *L = work, *U = work + nn, *R = work + 2 * nn;
    double *acc = nullptr;
    HIPTRY(hipMalloc(&acc, 3 * sizeof(double)));
    HIPTRY(hipMemsetAsync(acc, 0, 3 * sizeof(double), s));
    int rc = 0;
    if (mu::split_lu(F, n, lda, L, U, n, chol, s) || mu::copy_matrix(A, lda, R, n, n, n, s))
        rc = fail(MU_EHIP, "kernel launch failed");
    if (!rc)
        rc = chol ? mu_dgemm(stream, 'N', 'T', n, n, n, -1.0, L, n, L, n, 1.0, R, n)
                  : mu_dgemm(stream, 'N', 'N', n, n, n, -1.0, L, n, U, n, 1.0, R, n);
    if (!rc && (mu::sub_frob2(R, n, nullptr, 0, n, n, acc, s) ||
                mu::sub_frob2(A, lda, nullptr, 0, n, n, acc + 1, s) ||
                mu::sum_abs(R, n, n, n, acc + 2, s)))
        rc = fail(MU_EHIP, "kernel launch failed");
    double h[3] = {0, 0, 0};
    if (!rc) {
        hipError_t e = hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s);
        if (e != hipSuccess)
            rc = fail(MU_EHIP, hipGetErrorString(e));
    }
    (void)hipFree(acc);
    if (!rc) {
        if (frob_res)
            *frob_res = std::sqrt(h[0]);
        if (frob_a)
            *frob_a = std::sqrt(h[1]);
        if (l1_res)
            *l1_res = h[2];
    }
    return rc;
}

int mu_scaled_residual(const double *A, const double *F, int n, int lda, int chol, double *work,
                       double *result, void *stream)
{
    if (!result)
        return fail(MU_EINVAL, "result is NULL");
    double fr = 0, fa = 0;
    int rc = mu_residual_norms(A, F, n, lda, chol, work, &fr, &fa, nullptr, stream);
    if (!rc)
        *result = fr / (fa * n * DBL_EPSILON);
    return rc;
}

}  // extern "C"
