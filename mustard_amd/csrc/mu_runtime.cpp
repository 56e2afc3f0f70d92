// mu_runtime.cpp -- host side of the execution of lowered unit DAGs on one
// gfx950 GPU: the tiled mode's hipGraph of level kernels, the stream mode,
// and the persistent scheduler's state, launch and host watchdog.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <map>

#include "mu_kernels.hpp"
#include "mu_runtime.hpp"

namespace mu {

#define HIPCHK(x)                                                        \
    do {                                                                 \
        hipError_t e_ = (x);                                             \
        if (e_ != hipSuccess) {                                          \
            err = std::string(#x) + ": " + hipGetErrorString(e_);        \
            return -1;                                                   \
        }                                                                \
    } while (0)

static bool gemm_only_op(int op) { return op == OP_GEMM || op == OP_SYRK; }

static bool small_only_op(int op) { return op == OP_GETRF || op == OP_POTRF || op == OP_CVT; }

// Algorithmic flops of one lowered unit (the whole operation its items cover).
static double unit_flops(const UnitDesc &u)
{
    const double m = u.m, n = u.n, k = u.k;
    switch (u.op) {
    case OP_GEMM: return 2.0 * m * n * k;
    case OP_SYRK: return m * (m + 1) * k;
    case OP_TRSM: return m * m * n;  // triangle of order m, n right-hand sides
    case OP_GETRF: return 2.0 * m * m * m / 3.0;
    case OP_POTRF: return m * m * m / 3.0;
    default: return 0.0;
    }
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ---------------------------------------------------------------------------
// Merging the GEMM/SYRK units of one graph node.  The units of a level node
// are independent, and the trailing update of a tiled factorization is one
// GEMM (LU) or SYRK (Cholesky) cut into 1000-wide tiles: merged, its 64 x 64
// blocks run over the whole region instead of restarting at every tile edge,
// where a 1000-wide tile leaves a 40-wide ragged block row and column (5.9 %
// of the MFMA work of a tile).  A merged unit computes exactly what its parts
// computed, element for element and in the same order (bit-identical), as
// long as every part's operand and result addresses follow the merged
// unit's address arithmetic, which is what the checks below establish.
// ---------------------------------------------------------------------------
static bool mergeable(const UnitDesc &d)
{
    // FP64 storage with op(A) = A, or converted panels with op(A) = A^T (F_KVEC)
    if (!(d.op == OP_GEMM || d.op == OP_SYRK) || d.nitems <= 0 || d.grp > 1 || (d.flags & F_NARROW))
        return false;
    if (d.flags & F_SRC_LOW)
        return (d.flags & F_KVEC) != 0;
    return !(d.flags & F_TA);
}

// storage bytes of one operand element
static int64_t elem_bytes(const UnitDesc &d)
{
    if (!(d.flags & F_SRC_LOW))
        return 8;
    return (d.flags & F_LOW_BF16) ? 2 : 4;
}

static bool same_shape(const UnitDesc &x, const UnitDesc &y)
{
    return x.op == y.op && x.flags == y.flags && x.m == y.m && x.n == y.n && x.k == y.k &&
           x.lda == y.lda && x.ldb == y.ldb && x.ldc == y.ldc && x.alpha == y.alpha &&
           x.beta == y.beta;
}

// offset q - p in elements of `es` bytes (-1 when not a whole number of them;
// callers compare offsets only between operands of one storage type)
static int64_t eoff(const double *p, const double *q, int64_t es = 8)
{
    const int64_t d = (int64_t)((const char *)q - (const char *)p);
    return d % es ? INT64_MIN : d / es;
}
static const double *eadd(const double *p, int64_t elems, int64_t es)
{
    return (const double *)((const char *)p + elems * es);
}

struct MergeOut {
    std::vector<UnitDesc> &table;  // graph unit table; merged units are appended
    uint32_t &next_item;           // next free item number
    std::vector<uint32_t> &item_unit;
    uint32_t add(UnitDesc d)
    {
        d.nitems = unit_items(d);
        d.item_base = next_item;
        next_item += (uint32_t)d.nitems;
        const uint32_t id = (uint32_t)table.size();
        table.push_back(d);
        item_unit.resize(next_item, id);
        return id;
    }
};

// GEMM units (same shape) laid out as an R x Cn grid of C blocks: unit
// (r, c) has a = a0 + r sa with sa = m (op(A) = A) or m lda (op(A) = A^T),
// b = b0 + c sb with sb = n ldb (op(B) = B) or n (op(B) = B^T), and
// c = c0 + r m + c n ldc, offsets in elements of the operands' storage.
// Returns false unless `ids` fill the grid exactly once.
static bool merge_rect(const std::vector<UnitDesc> &U, const std::vector<uint32_t> &ids,
                       MergeOut &mo, uint32_t &merged)
{
    if (ids.size() < 2)
        return false;
    const UnitDesc &f = U[ids[0]];
    if (f.op != OP_GEMM)
        return false;
    const int64_t es = elem_bytes(f);
    const int64_t sa = (f.flags & F_TA) ? (int64_t)f.m * f.lda : (int64_t)f.m;
    const int64_t sb = (f.flags & F_TB) ? (int64_t)f.n : (int64_t)f.n * f.ldb;
    const int64_t sm = f.m, sc = (int64_t)f.n * f.ldc;
    const double *a0 = f.a, *b0 = f.b;
    for (uint32_t u : ids) {
        const int64_t da = eoff(a0, U[u].a, es), db = eoff(b0, U[u].b, es);
        if (da == INT64_MIN || db == INT64_MIN)
            return false;
        if (da < 0)
            a0 = U[u].a;
        if (db < 0)
            b0 = U[u].b;
    }
    int64_t R = 0, C = 0;
    std::vector<std::pair<int64_t, int64_t>> rc;
    for (uint32_t u : ids) {
        const int64_t da = eoff(a0, U[u].a, es), db = eoff(b0, U[u].b, es);
        if (da == INT64_MIN || db == INT64_MIN || da % sa || db % sb)
            return false;
        rc.push_back({da / sa, db / sb});
        R = std::max(R, da / sa + 1);
        C = std::max(C, db / sb + 1);
    }
    if (R * C != (int64_t)ids.size() || R * f.m > INT_MAX / 2 || C * f.n > INT_MAX / 2)
        return false;
    std::vector<char> seen((size_t)(R * C), 0);
    const double *c0 = U[ids[0]].c - (rc[0].first * sm + rc[0].second * sc);
    for (size_t q = 0; q < ids.size(); q++) {
        char &s = seen[(size_t)(rc[q].first * C + rc[q].second)];
        if (s || U[ids[q]].c != c0 + rc[q].first * sm + rc[q].second * sc)
            return false;
        s = 1;
    }
    UnitDesc d = f;
    d.m = (int32_t)(R * f.m);
    d.n = (int32_t)(C * f.n);
    d.a = a0;
    d.b = b0;
    d.c = const_cast<double *>(c0);
    d.gm = (d.m + GEMM_BM - 1) / GEMM_BM;
    d.gn = (d.n + GEMM_BN - 1) / GEMM_BN;
    merged = mo.add(d);
    return true;
}

// The Cholesky trailing update of rows/columns r0..r0+R-1 (tile order B):
// SYRK units c_r = c0 + r B (1 + ldc), a_r = a0 + r sa (sa = B, or B lda
// for a panel stored transposed, op(A) = A^T), and for every j > i the GEMM
// with a = a_j, b = a_i, c = c0 + j B + i B ldc; merged into one SYRK of
// order R B.  Removes what it merges from `ids`.
static void merge_triangles(const std::vector<UnitDesc> &U, std::vector<uint32_t> &ids,
                            MergeOut &mo, std::vector<uint32_t> &out)
{
    std::vector<uint32_t> syrk;
    for (uint32_t u : ids)
        if (U[u].op == OP_SYRK)
            syrk.push_back(u);
    if (syrk.size() < 2)
        return;
    const int64_t es = elem_bytes(U[syrk[0]]);
    std::sort(syrk.begin(), syrk.end(),
              [&](uint32_t x, uint32_t y) { return eoff(U[x].a, U[y].a, es) > 0; });
    const UnitDesc &f = U[syrk[0]];
    const int64_t B = f.m, R = (int64_t)syrk.size();
    const int64_t sa = (f.flags & F_TA) ? B * f.lda : B;
    for (int64_t r = 0; r < R; r++) {
        const UnitDesc &s = U[syrk[r]];
        if (!same_shape(s, f) || s.a != eadd(f.a, r * sa, es) || s.b != s.a ||
            s.c != f.c + r * B * (1 + (int64_t)f.ldc) || f.lda != f.ldb)
            return;
    }
    // GEMM (j, i) for every j > i, looked up by its C address
    std::map<const double *, uint32_t> by_c;
    for (uint32_t u : ids)
        if (U[u].op == OP_GEMM)
            by_c[U[u].c] = u;
    std::vector<uint32_t> parts(syrk);
    for (int64_t i = 0; i < R; i++)
        for (int64_t j = i + 1; j < R; j++) {
            auto it = by_c.find(f.c + j * B + i * B * (int64_t)f.ldc);
            if (it == by_c.end())
                return;
            const UnitDesc &g = U[it->second];
            if (g.flags != f.flags || g.m != B || g.n != B || g.k != f.k || g.lda != f.lda ||
                g.ldb != f.lda || g.ldc != f.ldc || g.alpha != f.alpha || g.beta != f.beta ||
                g.a != eadd(f.a, j * sa, es) || g.b != eadd(f.a, i * sa, es))
                return;
            parts.push_back(it->second);
        }
    UnitDesc d = f;
    d.m = d.n = (int32_t)(R * B);
    d.gm = d.gn = (d.m + GEMM_BM - 1) / GEMM_BM;
    out.push_back(mo.add(d));
    std::sort(parts.begin(), parts.end());
    std::vector<uint32_t> rest;
    for (uint32_t u : ids)
        if (!std::binary_search(parts.begin(), parts.end(), u))
            rest.push_back(u);
    ids.swap(rest);
}

// Merge what merges among the units `ids` of one GEMM node; returns the
// node's unit list (merged units plus the ones left alone).
static std::vector<uint32_t> merge_node(const std::vector<UnitDesc> &lowered,
                                        std::vector<uint32_t> ids, MergeOut &mo)
{
    std::vector<uint32_t> out, other;
    std::vector<uint32_t> cand;
    for (uint32_t u : ids)
        (mergeable(lowered[u]) ? cand : other).push_back(u);
    merge_triangles(mo.table, cand, mo, out);
    // groups of identical shape, then the whole group as one grid, else its
    // rows (same a) and then its columns (same b) as strips
    std::vector<char> used(cand.size(), 0);
    for (size_t g = 0; g < cand.size(); g++) {
        if (used[g])
            continue;
        std::vector<uint32_t> grp;
        for (size_t h = g; h < cand.size(); h++)
            if (!used[h] && same_shape(mo.table[cand[h]], mo.table[cand[g]])) {
                grp.push_back(cand[h]);
                used[h] = 1;
            }
        uint32_t m = 0;
        if (merge_rect(mo.table, grp, mo, m)) {
            out.push_back(m);
            continue;
        }
        for (int pass = 0; pass < 2; pass++) {
            std::map<const double *, std::vector<uint32_t>> lines;
            for (uint32_t u : grp)
                lines[pass == 0 ? mo.table[u].a : mo.table[u].b].push_back(u);
            std::vector<uint32_t> rest;
            for (auto &kv : lines) {
                if (kv.second.size() >= 2 && merge_rect(mo.table, kv.second, mo, m))
                    out.push_back(m);
                else
                    rest.insert(rest.end(), kv.second.begin(), kv.second.end());
            }
            grp.swap(rest);
        }
        out.insert(out.end(), grp.begin(), grp.end());
    }
    out.insert(out.end(), other.begin(), other.end());
    return out;
}

int run_units(const Lowered &low, hipStream_t s, std::string &err)
{
    // units of an executable come patched; an eager call gets a
    // stream-ordered workspace for the life of its launches
    const size_t wsd = ws_total_doubles(low);
    bool patched = true;
    for (const UnitDesc &d : low.units)
        if (d.ws_slot >= 0 && ((d.op == OP_TRINV && !d.c) || (d.op == OP_TRSM && !d.b) ||
                               (d.op == OP_GEMM && !d.part)))
            patched = false;
    if (wsd && !patched) {
        double *ws = nullptr;
        HIPCHK(hipMallocAsync((void **)&ws, wsd * sizeof(double), s));
        // the split GEMMs' arrival counters start at zero (a fill kernel, not
        // a memset node, in case the stream is being captured)
        if (fill_int(reinterpret_cast<int32_t *>(ws), wsd * 2, 0, s) != 0) {
            (void)hipFreeAsync(ws, s);
            err = "workspace fill failed";
            return -1;
        }
        std::vector<UnitDesc> units = low.units;
        patch_workspace(units, low, ws);
        Lowered copy;
        copy.units = std::move(units);
        int rc = run_units(copy, s, err);
        hipError_t fe = hipFreeAsync(ws, s);
        if (rc != 0)
            return rc;
        HIPCHK(fe);
        return 0;
    }
    for (const UnitDesc &d : low.units) {
        if (d.nitems <= 0)
            continue;
        if (launch_unit(d, s) != 0)
            HIPCHK(hipGetLastError());
    }
    return 0;
}

static int exec_build(Exec *e, const Dag &dag, int mode, const LowerOpts &lopts,
                      const ExecOpts &eopts, Exec **out, std::string &err);

int exec_create(const Dag &dag, int mode, const LowerOpts &lopts, const ExecOpts &eopts,
                Exec **out, std::string &err)
{
    Exec *e = new Exec();
    e->mode = mode;
    if (exec_build(e, dag, mode, lopts, eopts, out, err) != 0) {
        exec_destroy(e);  // frees whatever the failed build had allocated
        return -1;
    }
    return 0;
}

static int exec_build(Exec *e, const Dag &dag, int mode, const LowerOpts &lopts,
                      const ExecOpts &eopts, Exec **out, std::string &err)
{
    if (!lower(dag, lopts, e->low, err))
        return -1;
    Lowered &L = e->low;
    const uint32_t nu = (uint32_t)L.units.size();
    HIPCHK(hipGetDevice(&e->device));
    if (const size_t wsd = ws_total_doubles(L)) {
        HIPCHK(hipMalloc((void **)&e->ws, wsd * sizeof(double)));
        // zeroed once: the split GEMMs' arrival counters, which every launch
        // leaves at zero again
        HIPCHK(hipMemset(e->ws, 0, wsd * sizeof(double)));
        patch_workspace(L.units, L, e->ws);
    }

    // CSR successors and per-priority ring capacities
    std::vector<uint32_t> succ;
    uint32_t cap[NPRIO] = {0, 0, 0, 0};
    for (uint32_t u = 0; u < nu; u++) {
        L.units[u].succ_begin = (uint32_t)succ.size();
        succ.insert(succ.end(), L.succ[u].begin(), L.succ[u].end());
        L.units[u].succ_end = (uint32_t)succ.size();
        cap[L.units[u].prio] += (uint32_t)L.units[u].nitems;
    }
    std::vector<uint32_t> item_unit(L.nitems);
    for (uint32_t u = 0; u < nu; u++)
        for (int i = 0; i < L.units[u].nitems; i++)
            item_unit[L.units[u].item_base + i] = u;

    if (mode == MODE_GRAPH) {
        // Level of each unit = longest chain of units before it; units of one
        // level are independent.  Lane: a unit of a trailing update that
        // feeds only further updates goes to the bulk lane 1, every unit on
        // the way to the next panel to lane 0.  Levels increase along every
        // unit edge, so the node graph below is acyclic, and the bulk update
        // of one step runs beside the next step's panel work (lookahead).
        std::vector<int> level(nu, 0);
        int nlev = nu ? 1 : 0;
        for (uint32_t u = 0; u < nu; u++)
            for (uint32_t v : L.succ[u]) {
                level[v] = std::max(level[v], level[u] + 1);
                nlev = std::max(nlev, level[v] + 1);
            }
        // a unit belongs to a trailing update when its recorded task is a
        // GEMM/SYRK task (not a piece of a nested factorization or solve),
        // when it is an outer-level update of a factorization nested in two
        // levels (a whole-matrix task), or the conversion of a panel tile to
        // its reduced-precision shadow;
        // such a unit whose successors are all update units is bulk work ...
        auto update_task = [&](uint32_t u) {
            if (u < L.update_unit.size() && L.update_unit[u])
                return true;  // outer-level update of a two-level nested factorization
            const int tk = L.units[u].task;
            if (tk < 0 || tk >= (int)dag.tasks.size() || dag.tasks[tk].ops.empty())
                return false;
            for (const OpRec &op : dag.tasks[tk].ops)
                if (!gemm_only_op(op.op) && op.op != OP_CVT)
                    return false;
            return true;
        };
        // ... and so does every unit of a task no other task depends on (a
        // result nothing in this graph reads, such as the inverse of a
        // factored diagonal tile) whose successors stay inside that task
        std::vector<char> sink_task(dag.tasks.size(), 1);
        for (const TaskRec &tr : dag.tasks)
            for (int d : tr.deps)
                if (d >= 0 && d < (int)sink_task.size())
                    sink_task[d] = 0;
        std::vector<int> lane(nu, 0);
        if (eopts.lookahead == 2) {
            // Narrow lookahead, three lanes.  Lane 0 carries only the chain
            // from one diagonal factorization to the next: the factorizations,
            // the panel solves of the tiles at most two off the diagonal, and
            // the updates (within two updates) and conversions feeding them;
            // the other panel solves, with what feeds them as closely, run on
            // lane 2; every other update on lane 1.  A step then waits for the
            // solves of the tiles next to the diagonal instead of the whole
            // panel's, and for the updates of those tiles instead of the whole
            // previous trailing update.
            auto rank = [](int l) { return l == 0 ? 0 : l == 2 ? 1 : 2; };
            auto urgent = [&](int a, int b) { return rank(a) <= rank(b) ? a : b; };
            for (uint32_t u = 0; u < nu; u++) {
                if (update_task(u))
                    continue;
                const int tk = L.units[u].task;
                if (tk < 0 || tk >= (int)dag.tasks.size())
                    continue;
                if (sink_task[tk]) {
                    bool inner = true;
                    for (uint32_t v : L.succ[u])
                        inner = inner && L.units[v].task == tk;
                    lane[u] = inner ? 1 : 0;
                    continue;
                }
                const TaskRec &tr = dag.tasks[tk];
                bool factor = false, solve = false;
                for (const OpRec &op : tr.ops) {
                    factor = factor || op.op == OP_GETRF || op.op == OP_POTRF;
                    solve = solve || op.op == OP_TRSM;
                }
                const int dist = std::abs(tr.w.first - tr.w.second);
                lane[u] = (factor || !solve || dist <= 2) ? 0 : 2;
            }
            // updates: the most urgent lane among the solves / factorizations
            // they feed directly or through one more update
            std::vector<int> direct(nu, 1);
            for (uint32_t u = 0; u < nu; u++)
                if (update_task(u) && L.units[u].op != OP_CVT)
                    for (uint32_t v : L.succ[u])
                        if (!update_task(v))
                            direct[u] = urgent(direct[u], lane[v]);
            for (uint32_t u = 0; u < nu; u++)
                if (update_task(u) && L.units[u].op != OP_CVT) {
                    int l = direct[u];
                    for (uint32_t v : L.succ[u])
                        if (update_task(v) && L.units[v].op != OP_CVT)
                            l = urgent(l, direct[v]);
                    lane[u] = l;
                }
            for (uint32_t u = 0; u < nu; u++)
                if (update_task(u) && L.units[u].op == OP_CVT) {
                    int l = 1;
                    for (uint32_t v : L.succ[u])
                        l = urgent(l, lane[v]);
                    lane[u] = l;
                }
        } else if (eopts.lookahead)
            for (uint32_t u = 0; u < nu; u++) {
                const int tk = L.units[u].task;
                if (update_task(u)) {
                    bool bulk = true;
                    for (uint32_t v : L.succ[u])
                        bulk = bulk && update_task(v);
                    lane[u] = bulk ? 1 : 0;
                } else if (tk >= 0 && tk < (int)sink_task.size() && sink_task[tk]) {
                    bool inner = true;
                    for (uint32_t v : L.succ[u])
                        inner = inner && L.units[v].task == tk;
                    lane[u] = inner ? 1 : 0;
                }
            }
        // a panel tile's conversion to its reduced-precision shadow feeds the
        // next panel's updates too: it stays on the panel lane unless every
        // update that reads it is bulk work (unit order is topological, so
        // its successors' lanes are settled)
        if (eopts.lookahead == 1)
            for (uint32_t u = 0; u < nu; u++)
                if (L.units[u].op == OP_CVT && lane[u] == 1)
                    for (uint32_t v : L.succ[u])
                        if (lane[v] == 0)
                            lane[u] = 0;
        // the kernel a unit's items run on: GEMM/SYRK blocks on k_batch_gemm,
        // small diagonal factorizations / conversions on k_batch_small, panel
        // solves with inverted diagonal blocks (and the inversions, TRINV) on
        // the 4-wave k_batch_trsm_mid<256 / 512> or, for larger triangles, the
        // 8-wave k_batch_trsm, small solves without inverses on
        // k_batch_trsm_small; anything else on the generic k_batch
        auto ukind = [&](const UnitDesc &d) -> int {
            if (d.op == OP_GEMM && (d.flags & F_NARROW))
                return LEVEL_GEMM_NARROW;
            if (gemm_only_op(d.op) && (d.flags & F_KVEC))
                return LEVEL_GEMM_LOW;
            if (gemm_only_op(d.op))
                return LEVEL_GEMM;
            if (d.op == OP_CVT)
                return LEVEL_CVT;
            if (small_only_op(d.op))
                return LEVEL_SMALL;
            if (d.op == OP_TRINV)
                return LEVEL_TRSM_256;
            if (d.op == OP_TRSM && (d.flags & F_INV))
                return d.m <= 256 ? LEVEL_TRSM_256 : d.m <= 512 ? LEVEL_TRSM_512 : LEVEL_TRSM;
            if (d.op == OP_TRSM && d.m <= SMALL_MAX)
                return LEVEL_TRSM_SMALL;
            return LEVEL_GENERIC;
        };
        // one kernel node per (level, lane, kernel) with work, in that order
        // (a topological order of the node graph).  The nodes of a lane form a
        // chain of levels: each depends on every node of the lane's previous
        // level; a unit whose predecessor sits in another lane adds an edge
        // between the two nodes.
        const size_t NK = LEVEL_KINDS, NL = 3;
        const size_t nkey = NL * NK * (size_t)std::max(nlev, 1);
        auto ukey = [&](uint32_t u) {
            return (NL * (size_t)level[u] + lane[u]) * NK + ukind(L.units[u]);
        };
        std::vector<int> node_of(nkey, -1);
        std::vector<std::vector<uint32_t>> lv;
        std::vector<int> node_lane, node_level;
        // the graph's unit table: the lowered units, then the merged ones
        e->gunits = L.units;
        uint32_t next_item = (uint32_t)L.nitems;
        MergeOut mo{e->gunits, next_item, item_unit};
        {
            std::vector<std::vector<uint32_t>> by(nkey);
            for (uint32_t u = 0; u < nu; u++)
                if (L.units[u].nitems > 0)
                    by[ukey(u)].push_back(u);
            for (size_t k = 0; k < by.size(); k++)
                if (!by[k].empty()) {
                    std::vector<uint32_t> us = by[k];
                    if (eopts.fuse && (k % NK == LEVEL_GEMM || k % NK == LEVEL_GEMM_LOW))
                        us = merge_node(L.units, us, mo);
                    // the reduced-precision updates run on 128 x 128 blocks here
                    if (k % NK == LEVEL_GEMM_LOW)
                        for (uint32_t &u : us) {
                            UnitDesc d = e->gunits[u];
                            d.flags |= F_BLK128;
                            d.gm = (d.m + 127) / 128;
                            d.gn = d.op == OP_SYRK ? d.gm : (d.n + 127) / 128;
                            u = mo.add(d);
                        }
                    std::vector<uint32_t> its;
                    for (uint32_t u : us)
                        for (int i = 0; i < e->gunits[u].nitems; i++)
                            its.push_back(e->gunits[u].item_base + i);
                    node_of[k] = (int)lv.size();
                    node_lane.push_back((int)((k / NK) % NL));
                    node_level.push_back((int)(k / NK / NL));
                    e->level_gemm.push_back((char)(k % NK));
                    lv.push_back(std::move(its));
                }
        }
        const uint32_t ngu = (uint32_t)e->gunits.size();
        const int nnode = (int)lv.size();
        // node of a unit; units without items (none are lowered) map to -1
        auto unode = [&](uint32_t u) { return node_of[ukey(u)]; };
        std::vector<std::vector<int>> node_deps(nnode);
        {
            // per lane: the nodes of the last level seen, and of the level before
            std::vector<int> prev[NL], cur[NL];
            int cur_level[NL] = {-1, -1, -1};
            for (int k = 0; k < nnode; k++) {
                const int ln = node_lane[k];
                if (node_level[k] != cur_level[ln]) {
                    prev[ln] = cur[ln];
                    cur[ln].clear();
                    cur_level[ln] = node_level[k];
                }
                node_deps[k] = prev[ln];
                cur[ln].push_back(k);
            }
        }
        for (uint32_t u = 0; u < nu; u++)
            for (uint32_t v : L.succ[u])
                if (lane[u] != lane[v]) {
                    const int a = unode(u), b = unode(v);
                    if (a >= 0 && b >= 0 && a != b)
                        node_deps[b].push_back(a);
                }
        for (auto &d : node_deps) {
            std::sort(d.begin(), d.end());
            d.erase(std::unique(d.begin(), d.end()), d.end());
        }
        std::vector<uint32_t> items;
        std::vector<size_t> off;
        for (auto &v : lv) {
            off.push_back(items.size());
            items.insert(items.end(), v.begin(), v.end());
        }
        size_t o = 0;
        auto take = [&](size_t bytes) {
            size_t r = align_up(o, 256);
            o = r + bytes;
            return r;
        };
        const size_t o_units = take(sizeof(UnitDesc) * std::max<uint32_t>(ngu, 1));
        const size_t o_iu = take(sizeof(uint32_t) * std::max<size_t>(next_item, 1));
        const size_t o_items = take(sizeof(uint32_t) * std::max<size_t>(items.size(), 1));
        e->dbytes = o;
        HIPCHK(hipMalloc(&e->dbuf, e->dbytes));
        if (ngu)
            HIPCHK(hipMemcpy(e->dbuf + o_units, e->gunits.data(), sizeof(UnitDesc) * ngu,
                             hipMemcpyHostToDevice));
        if (next_item)
            HIPCHK(hipMemcpy(e->dbuf + o_iu, item_unit.data(), sizeof(uint32_t) * next_item,
                             hipMemcpyHostToDevice));
        if (!items.empty())
            HIPCHK(hipMemcpy(e->dbuf + o_items, items.data(), sizeof(uint32_t) * items.size(),
                             hipMemcpyHostToDevice));
        const UnitDesc *du = reinterpret_cast<const UnitDesc *>(e->dbuf + o_units);
        const uint32_t *diu = reinterpret_cast<const uint32_t *>(e->dbuf + o_iu);
        const uint32_t *ditems = reinterpret_cast<const uint32_t *>(e->dbuf + o_items);
        e->level_flops.assign(nnode, 0.0);
        for (uint32_t u = 0; u < nu; u++)
            if (L.units[u].nitems > 0)
                e->level_flops[unode(u)] += unit_flops(L.units[u]);
        HIPCHK(hipGraphCreate(&e->graph, 0));
        std::vector<hipGraphNode_t> gnodes(nnode);
        for (int l = 0; l < nnode; l++) {
            const uint32_t *first = ditems + off[l];
            void *args[] = {(void *)&du, (void *)&diu, (void *)&first};
            hipKernelNodeParams kp;
            std::memset(&kp, 0, sizeof kp);
            kp.func = const_cast<void *>(level_kernel(e->level_gemm[l]));
            kp.gridDim = dim3((unsigned)lv[l].size());
            kp.blockDim = dim3((unsigned)level_threads(e->level_gemm[l]));
            kp.kernelParams = args;
            std::vector<hipGraphNode_t> deps;
            for (int d : node_deps[l])
                deps.push_back(gnodes[d]);
            HIPCHK(hipGraphAddKernelNode(&gnodes[l], e->graph, deps.empty() ? nullptr : deps.data(),
                                         deps.size(), &kp));
        }
        e->levels = nnode;
        e->depth = nlev;
        e->level_items = std::move(items);
        e->level_off.assign(off.begin(), off.end());
        e->level_off.push_back(e->level_items.size());
        e->level_args[0] = (const void *)du;
        e->level_args[1] = (const void *)diu;
        e->level_args[2] = (const void *)ditems;
        if (nnode)
            HIPCHK(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
        *out = e;
        return 0;
    }
    if (mode == MODE_STREAM) {
        *out = e;
        return 0;
    }

    // persistent mode: every unit must fit the persistent kernel's item
    // bodies (the lowering options of this mode guarantee it)
    for (uint32_t u = 0; u < nu; u++) {
        const UnitDesc &d = L.units[u];
        const bool ok = (d.op == OP_TRSM && (d.flags & F_INV)) ? d.m <= PERSIST_TRSM_MAX
                        : d.op == OP_TRSM                      ? d.m <= SMALL_MAX
                        : (d.op == OP_GETRF || d.op == OP_POTRF) ? d.m <= SMALL_MAX
                                                                 : true;
        if (!ok) {
            err = "persistent mode: a unit exceeds the persistent kernel's item sizes "
                  "(nest triangular solves at <= 256, factorizations at <= 128)";
            return -1;
        }
    }
    // one device allocation for descriptors and state
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = align_up(off, 256);
        off = o + bytes;
        return o;
    };
    const uint32_t ntasks = (uint32_t)std::max<size_t>(dag.tasks.size(), 1);
    const size_t o_units = take(sizeof(UnitDesc) * nu);
    const size_t o_iu = take(sizeof(uint32_t) * L.nitems);
    const size_t o_succ = take(sizeof(uint32_t) * std::max<size_t>(succ.size(), 1));
    const size_t o_dep = take(sizeof(int32_t) * nu);
    const size_t o_left = take(sizeof(int32_t) * nu);
    uint32_t total = 0;
    for (int p = 0; p < NPRIO; p++)
        total += cap[p];
    const size_t o_slots = take(sizeof(uint32_t) * std::max<uint32_t>(total, 1));
    const size_t o_init = take(sizeof(uint32_t) * std::max<uint32_t>(total, 1));
    const size_t o_st = take(sizeof(SchedState));
    const size_t o_run = take(sizeof(int32_t) * ntasks);
    const size_t o_trace = eopts.trace ? take(sizeof(uint64_t) * 2 * L.nitems) : 0;
    e->dbytes = off;
    HIPCHK(hipMalloc(&e->dbuf, e->dbytes));
    // the host-written abort word: pinned, mapped host memory
    HIPCHK(hipHostMalloc((void **)&e->host_abort, 64, hipHostMallocMapped | hipHostMallocCoherent));
    __atomic_store_n(e->host_abort, 0u, __ATOMIC_SEQ_CST);
    uint32_t *dev_abort = nullptr;
    HIPCHK(hipHostGetDevicePointer((void **)&dev_abort, e->host_abort, 0));
    HIPCHK(hipEventCreateWithFlags(&e->done_ev, hipEventDisableTiming));

    // initial ring contents: items of the units without dependencies
    std::vector<uint32_t> init(std::max<uint32_t>(total, 1), 0);
    uint32_t ring_base[NPRIO], fill[NPRIO] = {0, 0, 0, 0};
    for (int p = 0, b = 0; p < NPRIO; p++) {
        ring_base[p] = b;
        b += cap[p];
    }
    for (uint32_t u = 0; u < nu; u++) {
        const UnitDesc &d = L.units[u];
        if (d.ndeps == 0)
            for (int i = 0; i < d.nitems; i++)
                init[ring_base[d.prio] + fill[d.prio]++] = d.item_base + i + 1;
    }
    HIPCHK(hipMemcpy(e->dbuf + o_units, L.units.data(), sizeof(UnitDesc) * nu,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->dbuf + o_iu, item_unit.data(), sizeof(uint32_t) * L.nitems,
                     hipMemcpyHostToDevice));
    if (!succ.empty())
        HIPCHK(hipMemcpy(e->dbuf + o_succ, succ.data(), sizeof(uint32_t) * succ.size(),
                         hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->dbuf + o_init, init.data(), sizeof(uint32_t) * init.size(),
                     hipMemcpyHostToDevice));

    int ncu = 0, per_cu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, e->device));
    if (persist_blocks_per_cu(&per_cu) != 0)
        HIPCHK(hipGetLastError());
    if (per_cu < 1)
        per_cu = 1;
    e->grid = eopts.grid > 0 ? eopts.grid : ncu * per_cu;

    ExecView &ev = e->ev;
    std::memset(&ev, 0, sizeof ev);
    ev.units = reinterpret_cast<const UnitDesc *>(e->dbuf + o_units);
    ev.item_unit = reinterpret_cast<const uint32_t *>(e->dbuf + o_iu);
    ev.succ = reinterpret_cast<const uint32_t *>(e->dbuf + o_succ);
    ev.dep_count = reinterpret_cast<int32_t *>(e->dbuf + o_dep);
    ev.items_left = reinterpret_cast<int32_t *>(e->dbuf + o_left);
    for (int p = 0; p < NPRIO; p++) {
        ev.slots[p] = reinterpret_cast<uint32_t *>(e->dbuf + o_slots) + ring_base[p];
        ev.cap[p] = cap[p];
    }
    e->slots_init = reinterpret_cast<uint32_t *>(e->dbuf + o_init);
    e->slots_total = total;
    e->ntasks = ntasks;
    ev.st = reinterpret_cast<SchedState *>(e->dbuf + o_st);
    ev.nunits = nu;
    ev.nitems = (uint32_t)L.nitems;
    e->trace = eopts.trace ? reinterpret_cast<uint64_t *>(e->dbuf + o_trace) : nullptr;
    ev.trace = e->trace;
    ev.host_abort = dev_abort;
    ev.task_running = reinterpret_cast<int32_t *>(e->dbuf + o_run);
    // CU budget per recorded task (--sm, percent of the GPU): workgroups per
    // task, and the task's share in the running-task sum that admits tasks
    // while it is below 100
    if (eopts.sm > 0 && eopts.sm < 100) {
        ev.task_budget = eopts.sm;
        ev.task_cap = std::max(1, (int)(((long long)eopts.sm * e->grid + 99) / 100));
    } else if (eopts.sm >= 100) {
        ev.task_budget = 100;
        ev.task_cap = e->grid;
    }
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->device));
    // the constant wall clock runs at 1 MHz .. 1 GHz; anything else is a
    // misreported rate, and 100 MHz (the CDNA value) is used instead
    e->wall_khz = (khz >= 1000 && khz <= 1000000) ? (uint64_t)khz : 100000;
    ev.timeout_ticks = (uint64_t)(eopts.timeout_s * 1000.0 * (double)e->wall_khz);
    // backstop independent of the clock: one empty poll takes well over 50 ns
    ev.poll_limit = (uint64_t)(eopts.timeout_s * 2.0e7) + 1000;
    e->host_timeout_s = 3.0 * eopts.timeout_s + 1.0;
    std::memcpy(e->init_cnt, fill, sizeof fill);
    *out = e;
    return 0;
}

// The host watchdog of a persistent run: waits for the run's end event and,
// past the deadline, sets the abort word every polling loop of the kernel
// reads, so a run that stopped making progress always ends.
static void watchdog_join(Exec *e)
{
    if (e->watchdog.joinable()) {
        e->wd_stop.store(true);
        e->watchdog.join();
    }
}

int exec_launch(Exec *e, hipStream_t s, std::string &err)
{
    if (e->mode == MODE_GRAPH) {
        if (e->gexec)
            HIPCHK(hipGraphLaunch(e->gexec, s));
        return 0;
    }
    if (e->mode == MODE_STREAM)
        return run_units(e->low, s, err);
    watchdog_join(e);
    __atomic_store_n(e->host_abort, 0u, __ATOMIC_SEQ_CST);
    e->host_aborted = false;
    ResetArgs ra;
    std::memset(&ra, 0, sizeof ra);
    for (int p = 0; p < NPRIO; p++)
        ra.init_cnt[p] = (int32_t)e->init_cnt[p];
    ra.slots_init = e->slots_init;
    ra.slots_flat = e->ev.slots[0];
    ra.slots_total = e->slots_total;
    ra.ntasks = e->ntasks;
    if (persist_launch(e->ev, ra, e->grid, s) != 0)
        HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e->done_ev, s));
    e->wd_stop.store(false);
    const double deadline = e->host_timeout_s;
    e->watchdog = std::thread([e, deadline]() {
        const auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            if (e->wd_stop.load())
                return;
            if (hipEventQuery(e->done_ev) != hipErrorNotReady)
                return;
            const double el =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el > deadline) {
                e->host_aborted = true;
                __atomic_store_n(e->host_abort, 1u, __ATOMIC_SEQ_CST);
                return;
            }
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
    });
    return 0;
}

int exec_launch_level(Exec *e, int level, hipStream_t s, std::string &err)
{
    if (e->mode != MODE_GRAPH || level < 0 || level >= e->levels) {
        err = "exec_launch_level: tiled executables only, level out of range";
        return -1;
    }
    const UnitDesc *du = (const UnitDesc *)e->level_args[0];
    const uint32_t *diu = (const uint32_t *)e->level_args[1];
    const uint32_t *first = (const uint32_t *)e->level_args[2] + e->level_off[level];
    const unsigned n = (unsigned)(e->level_off[level + 1] - e->level_off[level]);
    void *args[] = {(void *)&du, (void *)&diu, (void *)&first};
    HIPCHK(hipLaunchKernel(level_kernel(e->level_gemm[level]), dim3(n),
                           dim3((unsigned)level_threads(e->level_gemm[level])),
                           args, 0, s));
    HIPCHK(hipGetLastError());
    return 0;
}

int exec_check(Exec *e, std::string &err)
{
    if (e->mode != MODE_PERSISTENT)
        return 0;
    HIPCHK(hipEventSynchronize(e->done_ev));
    watchdog_join(e);
    SchedState st;
    HIPCHK(hipMemcpy(&st, e->ev.st, sizeof st, hipMemcpyDeviceToHost));
    if (st.err || st.done != e->ev.nitems || e->host_aborted) {
        const char *why = st.err == 1   ? "no progress for the time-out"
                          : st.err == 2 ? "poll limit without progress"
                          : st.err == 3 ? "more claims than queued items"
                          : e->host_aborted ? "aborted by the host watchdog"
                                            : "stopped early";
        err = "persistent scheduler stopped after " + std::to_string(st.done) + " of " +
              std::to_string(e->ev.nitems) + " items (" + why + ")";
        return -1;
    }
    return 0;
}

int exec_trace(Exec *e, std::vector<uint64_t> &stamps, std::string &err)
{
    if (!e->trace) {
        err = "tracing was not enabled";
        return -1;
    }
    stamps.resize(2 * (size_t)e->ev.nitems);
    HIPCHK(hipMemcpy(stamps.data(), e->trace, sizeof(uint64_t) * stamps.size(),
                     hipMemcpyDeviceToHost));
    return 0;
}

void exec_destroy(Exec *e)
{
    if (!e)
        return;
    watchdog_join(e);
    if (e->done_ev)
        (void)hipEventDestroy(e->done_ev);
    if (e->gexec)
        (void)hipGraphExecDestroy(e->gexec);
    if (e->graph)
        (void)hipGraphDestroy(e->graph);
    if (e->dbuf)
        (void)hipFree(e->dbuf);
    if (e->ws)
        (void)hipFree(e->ws);
    if (e->host_abort)
        (void)hipHostFree(e->host_abort);
    delete e;
}

}  // namespace mu
