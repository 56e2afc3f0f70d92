// mu_dag.cpp -- task recording and lowering to device units (host only).
#include "mu_dag.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <tuple>

namespace mu {

// ---------------------------------------------------------------------------
// recording
// ---------------------------------------------------------------------------
int Dag::begin_task(Tile w, const std::vector<Tile> &r)
{
    if (cur_ >= 0) {
        err = "beginCaptureOperation: the previous operation was not ended";
        return -1;
    }
    TaskRec t;
    t.w = w;
    t.r = r;
    t.tag = tag_;
    std::vector<Tile> tiles(r);
    tiles.push_back(w);
    std::vector<int> d;
    for (const Tile &x : tiles) {
        auto it = last_writer_.find(std::make_tuple(tag_, x.first, x.second));
        if (it != last_writer_.end())
            d.push_back(it->second);
    }
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    t.deps = d;
    tasks.push_back(t);
    cur_ = (int)tasks.size() - 1;
    return cur_;
}

int Dag::end_task()
{
    if (cur_ < 0) {
        err = "endCaptureOperation: no operation in progress";
        return -1;
    }
    const TaskRec &t = tasks[cur_];
    last_writer_[std::make_tuple(tag_, t.w.first, t.w.second)] = cur_;
    cur_ = -1;
    return 0;
}

int Dag::add_op(const OpRec &op)
{
    if (cur_ < 0) {
        err = "tile operation recorded outside begin/end of a task";
        return -1;
    }
    tasks[cur_].ops.push_back(op);
    return 0;
}

void Dag::add_edge(int src, int dst)
{
    if (dst < 0 || dst >= (int)tasks.size() || src < 0 || src >= (int)tasks.size())
        return;
    auto &d = tasks[dst].deps;
    if (!std::binary_search(d.begin(), d.end(), src)) {
        d.push_back(src);
        std::sort(d.begin(), d.end());
    }
}

int Dag::insert_sub(const Dag &child)
{
    if (cur_ >= 0) {
        err = "insertSubgraph: an operation is being recorded";
        return -1;
    }
    if (tasks.empty()) {
        err = "insertSubgraph: no task recorded yet";
        return -1;
    }
    if (child.cur_ >= 0) {
        err = "insertSubgraph: the sub-graph is still recording an operation";
        return -1;
    }
    std::vector<TaskRec> &sub = tasks.back().sub;
    // a second insertion into the same task runs after the first one: its
    // tasks follow the earlier ones, each also waiting for all of them
    const int base = (int)sub.size();
    for (const TaskRec &c : child.tasks) {
        TaskRec t = c;
        for (int &d : t.deps)
            d += base;
        for (int e = 0; e < base; e++)
            t.deps.push_back(e);
        std::sort(t.deps.begin(), t.deps.end());
        t.deps.erase(std::unique(t.deps.begin(), t.deps.end()), t.deps.end());
        sub.push_back(std::move(t));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// lowering
// ---------------------------------------------------------------------------
static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

int unit_items(const UnitDesc &u)
{
    switch (u.op) {
    case OP_GEMM:
        return u.grp > 1 ? cdiv(u.gm, u.grp) * cdiv(u.gn, u.grp) : u.gm * u.gn;
    case OP_SYRK: return u.gm * (u.gm + 1) / 2;
    case OP_TRSM: return u.gn;
    case OP_GETRF:
    case OP_POTRF: return 1;
    case OP_TRINV: return u.gm;
    case OP_CVT: return u.gm * u.gn;
    default: return 0;
    }
}

size_t ws_total_doubles(const Lowered &low)
{
    size_t t = 0;
    for (size_t v : low.ws_doubles)
        t += v;
    return t;
}

void patch_workspace(std::vector<UnitDesc> &units, const Lowered &low, double *ws)
{
    std::vector<size_t> off(low.ws_doubles.size() + 1, 0);
    for (size_t i = 0; i < low.ws_doubles.size(); i++)
        off[i + 1] = off[i] + low.ws_doubles[i];
    for (UnitDesc &d : units) {
        if (d.ws_slot < 0)
            continue;
        double *p = ws + off[d.ws_slot];
        if (d.op == OP_TRINV)
            d.c = p;
        else if (d.op == OP_GEMM)
            d.part = reinterpret_cast<float *>(p);
        else
            d.b = p;
    }
}

double op_flops(const OpRec &o)
{
    const double m = o.m, n = o.n, k = o.k;
    switch (o.op) {
    case OP_GEMM: return 2.0 * m * n * k;
    case OP_SYRK: return n * (n + 1) * k;
    case OP_TRSM: return m * m * n;  // order m, n right-hand sides
    case OP_GETRF: return 2.0 * m * m * m / 3.0;
    case OP_POTRF: return m * m * m / 3.0;
    default: return 0.0;
    }
}

// Operand element pairs are loaded as one vector: 16 bytes for FP64 storage,
// 8 for FP32, 4 for BF16.
static bool pair_aligned(const void *p, int flags)
{
    const uintptr_t a = (uintptr_t)p;
    if (!(flags & F_SRC_LOW))
        return (a & 15) == 0;
    return (flags & F_LOW_BF16) ? (a & 3) == 0 : (a & 7) == 0;
}

// Reduced-precision operands both k-contiguous (op(A) = A^T of a transposed
// panel, op(B) = B) with 16-byte aligned rows and k a whole number of
// 16-byte chunks: the long-k-slice kernel's layout (F_KVEC).
static bool kvec(int flags, const void *a, int lda, const void *b, int ldb, int k)
{
    if (!(flags & F_SRC_LOW) || !(flags & F_TA) || (flags & F_TB) ||
        !(flags & (F_LOW_F32 | F_LOW_BF16)))
        return false;
    const int ve = (flags & F_LOW_BF16) ? 8 : 4;  // elements per 16 bytes
    return k % ve == 0 && lda % ve == 0 && ldb % ve == 0 && ((uintptr_t)a & 15) == 0 &&
           ((uintptr_t)b & 15) == 0;
}

namespace {

struct Frag {
    std::vector<int> src, snk;
    // sources that read only the triangle of a solve (TRINV units): inside a
    // nested sub-DAG they wait for the triangle's producer, not the
    // right-hand side's
    std::vector<int> tri;
};

struct Builder {
    const LowerOpts &o;
    std::vector<UnitDesc> units;
    std::vector<std::vector<uint32_t>> succ;
    std::vector<int> has_pred;
    std::vector<size_t> ws_doubles;  // per workspace slot
    std::vector<char> upd;           // per unit: an outer-level trailing update
    bool cur_upd = false;            // the flag of the units lower_op adds now
    int task = 0;
    std::string err;

    explicit Builder(const LowerOpts &opts) : o(opts) {}

    void edge(int a, int b)
    {
        succ[a].push_back((uint32_t)b);
        has_pred[b] = 1;
    }

    int add_unit(UnitDesc d)
    {
        d.task = task;
        d.nitems = unit_items(d);
        units.push_back(d);
        succ.emplace_back();
        has_pred.push_back(0);
        upd.push_back(cur_upd ? 1 : 0);
        return (int)units.size() - 1;
    }

    static UnitDesc base(const OpRec &op)
    {
        UnitDesc d;
        std::memset(&d, 0, sizeof d);
        d.ws_slot = -1;
        d.op = op.op;
        d.flags = op.flags;
        d.m = op.m;
        d.n = op.n;
        d.k = op.k;
        d.lda = op.lda;
        d.ldb = op.ldb;
        d.ldc = op.ldc;
        d.alpha = op.alpha;
        d.beta = op.beta;
        d.a = op.a;
        d.b = op.b;
        d.c = op.c;
        d.info = op.info;
        d.info_base = op.info_base;
        return d;
    }

    // A sub-DAG recorded with the last-writer rule over integer block keys.
    struct SubRec {
        Builder &b;
        std::map<std::pair<int, int>, Frag> last;
        std::vector<Frag> frags;
        std::vector<int> is_pred;
        explicit SubRec(Builder &bb) : b(bb) {}
        // tri: key of the block holding a solve's triangle ({-1, -1}: not
        // written inside this sub-DAG)
        bool add(std::pair<int, int> w, std::initializer_list<std::pair<int, int>> r,
                 const OpRec &op, std::pair<int, int> tri = {-1, -1})
        {
            Frag f;
            if (!b.lower_op(op, f))
                return false;
            std::vector<const Frag *> preds;
            auto look = [&](std::pair<int, int> key) {
                auto it = last.find(key);
                if (it != last.end())
                    preds.push_back(&it->second);
            };
            for (auto key : r)
                look(key);
            look(w);
            auto is_tri = [&](int u) {
                return std::find(f.tri.begin(), f.tri.end(), u) != f.tri.end();
            };
            for (const Frag *p : preds)
                for (int s : p->snk)
                    for (int d : f.src)
                        if (!is_tri(d))
                            b.edge(s, d);
            if (!f.tri.empty() && tri.first >= 0) {
                auto it = last.find(tri);
                if (it != last.end())
                    for (int s : it->second.snk)
                        for (int d : f.tri)
                            b.edge(s, d);
            }
            last[w] = f;
            frags.push_back(f);
            return true;
        }
        Frag result() const
        {
            // sources: units without a predecessor inside the sub-DAG;
            // sinks: units without a successor inside it.
            Frag out;
            int lo = 1 << 30, hi = -1;
            for (const Frag &f : frags)
                for (int u : f.src) {
                    lo = std::min(lo, u);
                    hi = std::max(hi, u);
                }
            for (const Frag &f : frags)
                for (int u : f.snk)
                    hi = std::max(hi, u);
            if (hi < 0)
                return out;
            for (int u = lo; u <= hi; u++) {
                if (!b.has_pred[u])
                    out.src.push_back(u);
                bool inner = false;
                for (uint32_t s : b.succ[u])
                    if ((int)s <= hi)
                        inner = true;
                if (!inner)
                    out.snk.push_back(u);
            }
            return out;
        }
    };

    // Blocks of a nested operation of order n whose leaves hold at most
    // `leaf`: two levels (blocks of about nest_outer, nested again) above
    // 2 * nest_outer, one level below.
    int nest_parts(int n, int leaf) const
    {
        if (o.nest_outer > leaf && n > 2 * o.nest_outer)
            return cdiv(n, o.nest_outer);
        return cdiv(n, leaf);
    }

    static std::vector<int> split(int n, int parts)
    {
        std::vector<int> b(parts + 1);
        for (int i = 0; i <= parts; i++)
            b[i] = (int)((int64_t)i * n / parts);
        return b;
    }

    bool lower_op(const OpRec &op, Frag &f)
    {
        cur_upd = op.update && (op.op == OP_GEMM || op.op == OP_SYRK);
        switch (op.op) {
        case OP_GEMM: {
            if (op.m <= 0 || op.n <= 0)
                return true;
            UnitDesc d = base(op);
            d.gm = cdiv(op.m, GEMM_BM);
            d.gn = cdiv(op.n, GEMM_BN);
            if (op.flags & F_A32) {
                // FP32 A, op(A) = A, op(B) = B, n <= 16 (checked where recorded)
                d.flags |= F_NARROW;
                // 64-row items (NARROW32_ROWS); with fewer row blocks than
                // NARROW32_SPLIT_BELOW, split k into slices (items) as well
                d.gm = cdiv(op.m, 64);
                d.gn = d.gm < NARROW32_SPLIT_BELOW ? std::max(1, cdiv(op.k, NARROW32_KSLICE)) : 1;
                if (reinterpret_cast<uintptr_t>(op.a) % 16 == 0 && op.lda % 4 == 0 && op.m % 4 == 0)
                    d.flags |= F_VEC;
                if (d.gn > 1) {
                    // partial tiles (1024 floats per item) and one counter per row block
                    d.ws_slot = (int)ws_doubles.size();
                    ws_doubles.push_back((size_t)d.gm * d.gn * 512 + ((size_t)d.gm + 1) / 2);
                }
                int u = add_unit(d);
                f.src = {u};
                f.snk = {u};
                return true;
            }
            if (op.n <= 16 && !(op.flags & (F_TB | F_LOW_F32 | F_LOW_BF16))) {
                d.flags |= F_NARROW;
                d.gm = cdiv(op.m, 16);  // 16-row items (NARROW_ROWS)
                int u = add_unit(d);
                f.src = {u};
                f.snk = {u};
                return true;
            }
            if (pair_aligned(op.a, op.flags) && pair_aligned(op.b, op.flags) && op.lda % 2 == 0 &&
                op.ldb % 2 == 0 && op.m % 2 == 0 && op.n % 2 == 0 && op.k % 2 == 0)
                d.flags |= F_VEC;
            if (kvec(op.flags, op.a, op.lda, op.b, op.ldb, op.k))
                d.flags |= F_KVEC;
            d.grp = o.gemm_group > 1 ? o.gemm_group : 1;
            int u = add_unit(d);
            f.src = {u};
            f.snk = {u};
            return true;
        }
        case OP_SYRK: {
            if (op.n <= 0)
                return true;
            UnitDesc d = base(op);
            d.m = op.n;
            d.b = op.a;
            d.ldb = op.lda;
            d.gm = cdiv(op.n, GEMM_BM);
            d.gn = d.gm;
            // A A^T: op(B) = A^T, or, for a converted panel stored transposed
            // (F_TA: op(A) = A^T of the stored A^T), op(B) = the stored A^T
            if (!(op.flags & F_TA))
                d.flags |= F_TB;
            if (pair_aligned(op.a, op.flags) && op.lda % 2 == 0 && op.n % 2 == 0 && op.k % 2 == 0)
                d.flags |= F_VEC;
            if (kvec(op.flags, op.a, op.lda, op.a, op.lda, op.k))
                d.flags |= F_KVEC;
            int u = add_unit(d);
            f.src = {u};
            f.snk = {u};
            return true;
        }
        case OP_CVT: {
            if (op.m <= 0 || op.n <= 0)
                return true;
            UnitDesc d = base(op);
            d.gm = cdiv(op.m, GEMM_BM);
            d.gn = cdiv(op.n, GEMM_BN);
            int u = add_unit(d);
            f.src = {u};
            f.snk = {u};
            return true;
        }
        case OP_TRSM:
            if (op.m <= 0 || op.n <= 0)
                return true;
            if ((int64_t)(std::min(op.m, o.trsm_max) + 1) * op.lda >= ((int64_t)1 << 31)) {
                // the panel kernels address the triangle with 32-bit offsets
                err = "TRSM: leading dimension of the triangle too large (order x lda >= 2^31)";
                return false;
            }
            if (op.m <= o.trsm_max) {
                UnitDesc d = base(op);
                d.gm = 1;
                d.gn = cdiv(op.n, TRSM_W);
                if (op.m >= o.trsm_inv_min) {
                    // the triangle's 16 x 16 diagonal blocks are inverted once,
                    // by a TRINV unit, for all panels of the solve
                    const int nblk = cdiv(op.m, 16), slot = (int)ws_doubles.size();
                    ws_doubles.push_back((size_t)nblk * 256);
                    UnitDesc t = d;
                    t.op = OP_TRINV;
                    t.gm = cdiv(nblk, TRINV_BLOCKS);
                    t.gn = 1;
                    t.b = nullptr;
                    t.c = nullptr;
                    t.ws_slot = slot;
                    d.flags |= F_INV;
                    d.ws_slot = slot;
                    const int ut = add_unit(t);
                    const int u = add_unit(d);
                    edge(ut, u);
                    f.src = {ut, u};
                    f.snk = {u};
                    f.tri = {ut};
                    return true;
                }
                int u = add_unit(d);
                f.src = {u};
                f.snk = {u};
                return true;
            }
            return lower_trsm_nested(op, f);
        case OP_GETRF:
        case OP_POTRF:
            if (op.m <= 0)
                return true;
            if (op.m <= o.small_max) {
                UnitDesc d = base(op);
                d.gm = d.gn = 1;
                int u = add_unit(d);
                f.src = {u};
                f.snk = {u};
                return true;
            }
            return op.op == OP_GETRF ? lower_getrf_nested(op, f) : lower_potrf_nested(op, f);
        default:
            err = "unknown operation";
            return false;
        }
    }

    // Tiled LU of one large diagonal tile, in the reference's task order
    // (mustard/lu_mustard.cu:133-254, recordSubgraph).
    bool lower_getrf_nested(const OpRec &op, Frag &f)
    {
        const int n = op.m, P = nest_parts(n, o.small_max);
        const bool outer = P < cdiv(n, o.small_max);  // blocks that nest again
        std::vector<int> bd = split(n, P);
        const size_t ld = op.ldc;
        auto blk = [&](int i, int j) { return op.c + bd[i] + (size_t)bd[j] * ld; };
        auto sz = [&](int i) { return bd[i + 1] - bd[i]; };
        SubRec rec(*this);
        for (int k = 0; k < P; k++) {
            OpRec g;
            g.op = OP_GETRF;
            g.m = g.n = sz(k);
            g.c = blk(k, k);
            g.ldc = (int)ld;
            g.info = op.info;
            g.info_base = op.info_base + bd[k];
            if (!rec.add({k, k}, {{k, k}}, g))
                return false;
            for (int i = k + 1; i < P; i++) {
                OpRec t;
                t.op = OP_TRSM;
                t.flags = F_UNIT;
                t.m = sz(k);
                t.n = sz(i);
                t.a = blk(k, k);
                t.lda = (int)ld;
                t.c = blk(k, i);
                t.ldc = (int)ld;
                t.alpha = 1.0;
                if (!rec.add({k, i}, {{k, k}, {k, i}}, t, {k, k}))
                    return false;
            }
            for (int i = k + 1; i < P; i++) {
                OpRec t;
                t.op = OP_TRSM;
                t.flags = F_T_TRANS | F_Z_ROWS;
                t.m = sz(k);
                t.n = sz(i);
                t.a = blk(k, k);
                t.lda = (int)ld;
                t.c = blk(i, k);
                t.ldc = (int)ld;
                t.alpha = 1.0;
                if (!rec.add({i, k}, {{k, k}, {i, k}}, t, {k, k}))
                    return false;
                for (int j = k + 1; j < P; j++) {
                    OpRec g2;
                    g2.op = OP_GEMM;
                    g2.m = sz(i);
                    g2.n = sz(j);
                    g2.k = sz(k);
                    g2.a = blk(i, k);
                    g2.lda = (int)ld;
                    g2.b = blk(k, j);
                    g2.ldb = (int)ld;
                    g2.c = blk(i, j);
                    g2.ldc = (int)ld;
                    g2.flags = op.flags & (F_LOW_F32 | F_LOW_BF16);
                    g2.update = outer;
                    if (!rec.add({i, j}, {{i, k}, {k, j}, {i, j}}, g2))
                        return false;
                }
            }
        }
        f = rec.result();
        return true;
    }

    // Tiled Cholesky of one large diagonal tile
    // (order of mustard/cholesky_mustard.cu:221-400).
    bool lower_potrf_nested(const OpRec &op, Frag &f)
    {
        const int n = op.m, P = nest_parts(n, o.small_max);
        const bool outer = P < cdiv(n, o.small_max);  // blocks that nest again
        std::vector<int> bd = split(n, P);
        const size_t ld = op.ldc;
        auto blk = [&](int i, int j) { return op.c + bd[i] + (size_t)bd[j] * ld; };
        auto sz = [&](int i) { return bd[i + 1] - bd[i]; };
        SubRec rec(*this);
        for (int k = 0; k < P; k++) {
            OpRec p;
            p.op = OP_POTRF;
            p.m = p.n = sz(k);
            p.c = blk(k, k);
            p.ldc = (int)ld;
            p.info = op.info;
            p.info_base = op.info_base + bd[k];
            if (!rec.add({k, k}, {{k, k}}, p))
                return false;
            for (int i = k + 1; i < P; i++) {
                OpRec t;
                t.op = OP_TRSM;
                t.flags = F_Z_ROWS;
                t.m = sz(k);
                t.n = sz(i);
                t.a = blk(k, k);
                t.lda = (int)ld;
                t.c = blk(i, k);
                t.ldc = (int)ld;
                t.alpha = 1.0;
                if (!rec.add({i, k}, {{k, k}, {i, k}}, t, {k, k}))
                    return false;
            }
            for (int i = k + 1; i < P; i++) {
                OpRec s;
                s.op = OP_SYRK;
                s.n = s.m = sz(i);
                s.k = sz(k);
                s.a = blk(i, k);
                s.lda = (int)ld;
                s.c = blk(i, i);
                s.ldc = (int)ld;
                s.flags = op.flags & (F_LOW_F32 | F_LOW_BF16);
                s.update = outer;
                if (!rec.add({i, i}, {{i, i}, {i, k}}, s))
                    return false;
                for (int j = i + 1; j < P; j++) {
                    OpRec g;
                    g.op = OP_GEMM;
                    g.flags = F_TB | (op.flags & (F_LOW_F32 | F_LOW_BF16));
                    g.update = outer;
                    g.m = sz(j);
                    g.n = sz(i);
                    g.k = sz(k);
                    g.a = blk(j, k);
                    g.lda = (int)ld;
                    g.b = blk(i, k);
                    g.ldb = (int)ld;
                    g.c = blk(j, i);
                    g.ldc = (int)ld;
                    if (!rec.add({j, i}, {{j, i}, {j, k}, {i, k}}, g))
                        return false;
                }
            }
        }
        f = rec.result();
        return true;
    }

    // Blocked solve of T Y = alpha Z for a triangle above trsm_max, right
    // looking over Z-space row blocks: Y_p = T_pp^-1 Z_p, then Z_q -= T_qp Y_p.
    bool lower_trsm_nested(const OpRec &op, Frag &f)
    {
        const int n = op.m, P = nest_parts(n, o.trsm_max);
        std::vector<int> bd = split(n, P);
        const bool rev = op.flags & F_REV, tt = op.flags & F_T_TRANS, zr = op.flags & F_Z_ROWS;
        const size_t lda = op.lda, ldz = op.ldc;
        // memory range of Z-space block p
        auto m0 = [&](int p) { return rev ? n - bd[p + 1] : bd[p]; };
        auto sz = [&](int p) { return bd[p + 1] - bd[p]; };
        SubRec rec(*this);
        for (int p = 0; p < P; p++) {
            OpRec t = op;
            t.m = sz(p);
            t.a = op.a + m0(p) + (size_t)m0(p) * lda;
            t.c = zr ? op.c + (size_t)m0(p) * ldz : op.c + m0(p);
            if (!rec.add({p, 0}, {{p, 0}}, t))
                return false;
            for (int q = p + 1; q < P; q++) {
                OpRec g;
                g.op = OP_GEMM;
                g.alpha = -1.0 / op.alpha;
                g.beta = 1.0;
                // T_qp in memory: S(mq, mp) or, read transposed, S(mp, mq)^T
                const double *s_qp = tt ? op.a + m0(p) + (size_t)m0(q) * lda
                                        : op.a + m0(q) + (size_t)m0(p) * lda;
                if (!zr) {
                    // Z[q rows] -= T_qp Z[p rows]
                    g.m = sz(q);
                    g.n = op.n;
                    g.k = sz(p);
                    g.a = s_qp;
                    g.lda = (int)lda;
                    g.flags = tt ? F_TA : 0;
                    g.b = op.c + m0(p);
                    g.ldb = (int)ldz;
                    g.c = op.c + m0(q);
                    g.ldc = (int)ldz;
                } else {
                    // B[:, q] -= B[:, p] T_qp^T
                    g.m = op.n;
                    g.n = sz(q);
                    g.k = sz(p);
                    g.a = op.c + (size_t)m0(p) * ldz;
                    g.lda = (int)ldz;
                    g.b = s_qp;
                    g.ldb = (int)lda;
                    g.flags = tt ? 0 : F_TB;
                    g.c = op.c + (size_t)m0(q) * ldz;
                    g.ldc = (int)ldz;
                }
                if (!rec.add({q, 0}, {{p, 0}, {q, 0}}, g))
                    return false;
            }
        }
        f = rec.result();
        return true;
    }
};

// Estimated time of one item in microseconds (drives the priorities only).
double item_cost_us(const UnitDesc &u)
{
    const double per_cu_flops = 0.25e6;  // flops per microsecond per CU, FP64 MFMA
    switch (u.op) {
    case OP_GEMM:
    case OP_SYRK:
        return 3.0 + 2.0 * std::min(u.m, GEMM_BM) * std::min(u.n, GEMM_BN) * (double)u.k /
                         (0.8 * per_cu_flops);
    case OP_TRSM: return 3.0 + (double)u.m * u.m * TRSM_W / (0.2 * per_cu_flops) + u.m * 0.05;
    case OP_GETRF: return 3.0 + 2.0 * u.m * u.m * (double)u.m / 3.0 / (0.05 * per_cu_flops);
    case OP_POTRF: return 3.0 + u.m * u.m * (double)u.m / 3.0 / (0.05 * per_cu_flops);
    case OP_TRINV: return 5.0;
    case OP_CVT: return 3.0;
    default: return 1.0;
    }
}

}  // namespace

// Lower one task: its operations in the order they were issued, then its
// nested sub-DAG (if any).  `have` is false for a task without device work.
static bool lower_task(Builder &b, const TaskRec &task, Frag &whole, bool &have, double &flops);

// Lower a list of tasks whose deps index the same list; tf[i] is task i's
// fragment (a task without device work passes on its predecessors' sinks).
static bool lower_tasks(Builder &b, const std::vector<TaskRec> &tasks, std::vector<Frag> &tf,
                        std::vector<int> *first_unit, std::vector<int> *nunits, double &flops,
                        bool top)
{
    tf.assign(tasks.size(), Frag());
    for (size_t t = 0; t < tasks.size(); t++) {
        const TaskRec &task = tasks[t];
        if (top)
            b.task = (int)t;
        const int first = (int)b.units.size();
        Frag whole;
        bool have = false;
        if (!lower_task(b, task, whole, have, flops))
            return false;
        for (int p : task.deps)
            if (p < 0 || p >= (int)t) {
                b.err = "task dependency is not an earlier task";
                return false;
            }
        if (!have) {
            for (int p : task.deps)
                whole.snk.insert(whole.snk.end(), tf[p].snk.begin(), tf[p].snk.end());
            std::sort(whole.snk.begin(), whole.snk.end());
            whole.snk.erase(std::unique(whole.snk.begin(), whole.snk.end()), whole.snk.end());
        } else {
            for (int p : task.deps)
                for (int s : tf[p].snk)
                    for (int d : whole.src)
                        b.edge(s, d);
        }
        tf[t] = whole;
        if (first_unit) {
            (*first_unit)[t] = first;
            (*nunits)[t] = (int)b.units.size() - first;
        }
    }
    return true;
}

static bool lower_task(Builder &b, const TaskRec &task, Frag &whole, bool &have, double &flops)
{
    have = false;
    auto chain = [&](const Frag &f) {
        if (!have) {
            whole = f;
            have = true;
            return;
        }
        // ops of one task run in the order they were issued
        for (int s : whole.snk)
            for (int d : f.src)
                b.edge(s, d);
        whole.snk = f.snk;
    };
    for (const OpRec &op : task.ops) {
        flops += op_flops(op);
        Frag f;
        if (!b.lower_op(op, f))
            return false;
        if (!f.src.empty())
            chain(f);
    }
    if (!task.sub.empty()) {
        // the sub-DAG's own sources and sinks: units of its range without a
        // predecessor inside it, and without any successor (nothing later
        // exists yet)
        const int lo = (int)b.units.size();
        std::vector<Frag> stf;
        if (!lower_tasks(b, task.sub, stf, nullptr, nullptr, flops, false))
            return false;
        Frag f;
        for (int u = lo; u < (int)b.units.size(); u++) {
            if (!b.has_pred[u])
                f.src.push_back(u);
            if (b.succ[u].empty())
                f.snk.push_back(u);
        }
        if (!f.src.empty())
            chain(f);
    }
    return true;
}

bool lower(const Dag &dag, const LowerOpts &opts, Lowered &out, std::string &err)
{
    Builder b(opts);
    std::vector<Frag> tf;
    out.task_first_unit.assign(dag.tasks.size(), -1);
    out.task_nunits.assign(dag.tasks.size(), 0);
    out.flops = 0.0;
    if (!lower_tasks(b, dag.tasks, tf, &out.task_first_unit, &out.task_nunits, out.flops, true)) {
        err = b.err;
        return false;
    }

    const size_t nu = b.units.size();
    // de-duplicate edges, count dependencies, check the order is topological
    for (size_t u = 0; u < nu; u++)
        b.units[u].ndeps = 0;
    for (size_t u = 0; u < nu; u++) {
        auto &s = b.succ[u];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        for (uint32_t v : s) {
            if (v <= u) {
                err = "lowered DAG is not topologically ordered";
                return false;
            }
            b.units[v].ndeps++;
        }
    }
    // priorities: slack of each unit against the critical path
    std::vector<double> cost(nu), tl(nu, 0.0), bl(nu, 0.0);
    for (size_t u = 0; u < nu; u++)
        cost[u] = b.units[u].nitems ? item_cost_us(b.units[u]) : 0.5;
    for (size_t u = 0; u < nu; u++)
        for (uint32_t v : b.succ[u])
            tl[v] = std::max(tl[v], tl[u] + cost[u]);
    for (size_t u = nu; u-- > 0;) {
        double m = 0.0;
        for (uint32_t v : b.succ[u])
            m = std::max(m, bl[v]);
        bl[u] = cost[u] + m;
    }
    double cp = 0.0;
    for (size_t u = 0; u < nu; u++)
        cp = std::max(cp, tl[u] + bl[u]);
    out.critical_path = cp;
    uint32_t base = 0;
    for (size_t u = 0; u < nu; u++) {
        UnitDesc &d = b.units[u];
        const double slack = cp > 0 ? (cp - tl[u] - bl[u]) / cp : 0.0;
        int pr = slack < 0.01 ? 0 : slack < 0.05 ? 1 : slack < 0.2 ? 2 : 3;
        d.prio = std::min(pr, opts.nprio - 1);
        d.item_base = base;
        base += (uint32_t)d.nitems;
    }
    out.nitems = base;
    out.units = std::move(b.units);
    out.succ = std::move(b.succ);
    out.ws_doubles = std::move(b.ws_doubles);
    out.update_unit = std::move(b.upd);
    return true;
}

}  // namespace mu
