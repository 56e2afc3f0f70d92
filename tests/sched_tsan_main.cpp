// Host driver for the persistent scheduler's emulation under ThreadSanitizer,
// and for the recorder and lowering under AddressSanitizer + UBSan (built and
// run by tests/test_sched_tsan.py with g++ -fsanitize=...).
//
// Records the reference's tiled LU and Cholesky task sequences
// (lu_mustard.cu:343-535, cholesky_mustard.cu:221-400) on fake device
// pointers -- nothing is computed -- and runs mu::emulate on them: the
// kernel's own worker loop (mu_sched.hpp) on one thread per wave, four per
// workgroup.  Prints one line per case and exits non-zero when an item ran
// twice, ran before a predecessor, was lost, or a CU budget was exceeded.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "mu_dag.hpp"
#include "mu_sched_host.hpp"

using mu::OpRec;

static double *fake(size_t off) { return reinterpret_cast<double *>((1ull << 40) + 8 * off); }

static void task(mu::Dag &d, mu::Tile w, std::vector<mu::Tile> r, const OpRec &op)
{
    d.begin_task(w, r);
    d.add_op(op);
    d.end_task();
}

// Tile (i, j) of an n x n column-major matrix with tiles of b.
static double *tile(int n, int b, int i, int j) { return fake((size_t)i * b + (size_t)j * b * n); }

static void record_lu(mu::Dag &d, int n, int t, int ld = 0, size_t off = 0, int subt = 0);

// Diagonal GETRF task k of a tiled LU as an empty task with the tiled LU of
// the tile (subt x subt sub-tiles) inserted as its nested sub-DAG.
static void nested_getrf(mu::Dag &d, int n, int b, int k, int ld, size_t off, int subt)
{
    d.begin_task({k, k}, {{k, k}});
    d.end_task();
    mu::Dag child;
    record_lu(child, b, subt, ld, off + (size_t)k * b + (size_t)k * b * ld);
    d.insert_sub(child);
}

// ld: leading dimension (0: n); off: element offset of the matrix; subt > 0
// nests every diagonal tile as an inserted sub-DAG of subt x subt tiles.
static void record_lu(mu::Dag &d, int n, int t, int ld, size_t off, int subt)
{
    const int b = n / t;
    if (!ld)
        ld = n;
    auto tile = [&](int n_, int b_, int i, int j) {
        (void)n_;
        return fake(off + (size_t)i * b_ + (size_t)j * b_ * ld);
    };
    for (int k = 0; k < t; k++) {
        if (subt > 0) {
            nested_getrf(d, n, b, k, ld, off, subt);
        } else {
            OpRec g;
            g.op = mu::OP_GETRF;
            g.m = g.n = b;
            g.c = tile(n, b, k, k);
            g.ldc = ld;
            task(d, {k, k}, {{k, k}}, g);
        }
        for (int i = k + 1; i < t; i++) {  // A_ki <- L_kk^-1 A_ki
            OpRec s;
            s.op = mu::OP_TRSM;
            s.flags = mu::F_UNIT;
            s.m = b;
            s.n = b;
            s.alpha = 1.0;
            s.a = tile(n, b, k, k);
            s.lda = ld;
            s.c = tile(n, b, k, i);
            s.ldc = ld;
            task(d, {k, i}, {{k, k}, {k, i}}, s);
        }
        for (int i = k + 1; i < t; i++) {
            OpRec s;  // A_ik <- A_ik U_kk^-1
            s.op = mu::OP_TRSM;
            s.flags = mu::F_T_TRANS | mu::F_Z_ROWS;
            s.m = b;
            s.n = b;
            s.alpha = 1.0;
            s.a = tile(n, b, k, k);
            s.lda = ld;
            s.c = tile(n, b, i, k);
            s.ldc = ld;
            task(d, {i, k}, {{k, k}, {i, k}}, s);
            for (int j = k + 1; j < t; j++) {
                OpRec m;  // A_ij -= A_ik A_kj
                m.op = mu::OP_GEMM;
                m.m = m.n = m.k = b;
                m.a = tile(n, b, i, k);
                m.b = tile(n, b, k, j);
                m.c = tile(n, b, i, j);
                m.lda = m.ldb = m.ldc = ld;
                task(d, {i, j}, {{i, k}, {k, j}, {i, j}}, m);
            }
        }
    }
}

static void record_cholesky(mu::Dag &d, int n, int t)
{
    const int b = n / t;
    for (int k = 0; k < t; k++) {
        OpRec p;
        p.op = mu::OP_POTRF;
        p.m = p.n = b;
        p.c = tile(n, b, k, k);
        p.ldc = n;
        task(d, {k, k}, {{k, k}}, p);
        for (int i = k + 1; i < t; i++) {
            OpRec s;  // A_ik <- A_ik L_kk^-T
            s.op = mu::OP_TRSM;
            s.flags = mu::F_Z_ROWS;
            s.m = b;
            s.n = b;
            s.alpha = 1.0;
            s.a = tile(n, b, k, k);
            s.lda = n;
            s.c = tile(n, b, i, k);
            s.ldc = n;
            task(d, {i, k}, {{k, k}, {i, k}}, s);
        }
        for (int i = k + 1; i < t; i++) {
            OpRec y;  // A_ii -= A_ik A_ik^T
            y.op = mu::OP_SYRK;
            y.m = y.n = b;
            y.k = b;
            y.a = tile(n, b, i, k);
            y.lda = n;
            y.c = tile(n, b, i, i);
            y.ldc = n;
            task(d, {i, i}, {{i, k}, {i, i}}, y);
            for (int j = i + 1; j < t; j++) {
                OpRec m;  // A_ji -= A_jk A_ik^T
                m.op = mu::OP_GEMM;
                m.flags = mu::F_TB;
                m.m = m.n = m.k = b;
                m.a = tile(n, b, j, k);
                m.b = tile(n, b, i, k);
                m.c = tile(n, b, j, i);
                m.lda = m.ldb = m.ldc = n;
                task(d, {j, i}, {{j, k}, {i, k}, {j, i}}, m);
            }
        }
    }
}

int main()
{
    struct Case {
        const char *kind;
        int n, t, wgs, delay_us, sm, subt;
    };
    const Case cases[] = {
        {"lu", 600, 2, 3, 0, 0, 0},    {"lu", 1200, 4, 6, 0, 0, 0},
        {"chol", 1200, 4, 6, 0, 0, 0}, {"lu", 1200, 3, 4, 50, 0, 0},
        {"chol", 1500, 3, 5, 0, 20, 0}, {"lu", 1000, 4, 8, 0, 100, 0},
        {"lu", 2400, 2, 6, 20, 0, 3},  // diagonal tiles as inserted sub-DAGs
    };
    int bad = 0;
    for (const Case &c : cases) {
        mu::Dag d;
        if (std::string(c.kind) == "lu")
            record_lu(d, c.n, c.t, 0, 0, c.subt);
        else
            record_cholesky(d, c.n, c.t);
        mu::LowerOpts lo;
        lo.trsm_max = 256;
        lo.gemm_group = 2;
        mu::EmuOpts eo;
        eo.workgroups = c.wgs;
        eo.timeout_s = 60.0;
        eo.host_deadline_s = 120.0;
        eo.producer_delay_us = c.delay_us;
        eo.sm = c.sm;
        mu::EmuResult r;
        std::string err;
        if (!mu::emulate(d, lo, eo, r, err)) {
            std::printf("%s %d/%d: %s\n", c.kind, c.n, c.t, err.c_str());
            return 2;
        }
        const bool ok = r.executed == r.nitems && r.duplicates == 0 && r.order_violations == 0 &&
                        r.err == 0 && r.done == r.nitems &&
                        (c.sm == 0 || r.max_task_workgroups <= r.task_cap) &&
                        (c.sm != 100 || r.max_tasks_running == 1);
        std::printf("%s n=%d t=%d sub=%d workgroups=%d delay=%dus sm=%d: items %u executed "
                    "%llu dup %llu order %llu err %u task max %u cap %u tasks %u %s\n",
                    c.kind, c.n, c.t, c.subt, c.wgs, c.delay_us, c.sm, r.nitems,
                    (unsigned long long)r.executed, (unsigned long long)r.duplicates,
                    (unsigned long long)r.order_violations, r.err, r.max_task_workgroups,
                    r.task_cap, r.max_tasks_running, ok ? "ok" : "FAILED");
        bad += ok ? 0 : 1;
    }
    return bad ? 1 : 0;
}
