// mu_runtime.hip -- kernels and host runtime for lowered unit DAGs on gfx950.
//
// Replaces the reference's device runtime (mustard/mustard.h:57-129,
// mustard/broker_queue.h): instead of a one-thread scheduler kernel that
// device-launches one CUDA sub-graph per task (HIP has no device-side graph
// launch), every workgroup of one persistent kernel pulls ready work items
// from HBM-resident queue rings, runs them inline, and on completion
// decrements the dependency counters of the successor units, queueing the
// items of every unit whose counter reaches zero.
#include <algorithm>
#include <climits>
#include <cstring>

#include "mu_ops.hpp"
#include "mu_runtime.hpp"
#include "mu_sched.hpp"

namespace mu {

#define HIPCHK(x)                                                        \
    do {                                                                 \
        hipError_t e_ = (x);                                             \
        if (e_ != hipSuccess) {                                          \
            err = std::string(#x) + ": " + hipGetErrorString(e_);        \
            return -1;                                                   \
        }                                                                \
    } while (0)

#define AGENT __HIP_MEMORY_SCOPE_AGENT

// ---------------------------------------------------------------------------
// per-unit grid kernels (graph and stream modes): item = blockIdx.x.  The
// GEMM/SYRK variants need only the GEMM block's LDS, so four workgroups fit
// on a CU (the generic kernel's LDS, sized for the TRSM panel, allows one).
// ---------------------------------------------------------------------------
#define MU_GEMM_KERNEL __launch_bounds__(WG_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4)))

__global__ __launch_bounds__(WG_THREADS) void k_unit(UnitDesc d)
{
    __shared__ double smem[LDS_DOUBLES];
    run_item(d, blockIdx.x, smem);
}

__global__ MU_GEMM_KERNEL void k_unit_gemm(UnitDesc d)
{
    __shared__ double smem[LDS_GEMM];
    run_gemm_item(d, blockIdx.x, smem);
}

// ---------------------------------------------------------------------------
// level-batched kernels (graph mode): one launch runs every item of every
// unit at one level of the unit DAG; block b runs items[b]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG_THREADS) void k_batch(const UnitDesc *units,
                                                      const uint32_t *item_unit,
                                                      const uint32_t *items)
{
    __shared__ double smem[LDS_DOUBLES];
    const uint32_t item = items[blockIdx.x];
    // a private copy of the descriptor lives in registers; a reference into
    // global memory would be re-read after every barrier of the item body
    const UnitDesc d = units[item_unit[item]];
    run_item(d, (int)(item - d.item_base), smem);
}

__global__ MU_GEMM_KERNEL void k_batch_gemm(const UnitDesc *units, const uint32_t *item_unit,
                                            const uint32_t *items)
{
    __shared__ double smem[LDS_GEMM];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_gemm_item(d, (int)(item - d.item_base), smem);
}

__global__ __launch_bounds__(WG_THREADS) void k_batch_small(const UnitDesc *units,
                                                            const uint32_t *item_unit,
                                                            const uint32_t *items)
{
    __shared__ double smem[LDS_SMALL];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_small_item(d, (int)(item - d.item_base), smem);
}

// GEMMs with at most 16 output columns (substitution steps of a solve).
__global__ __launch_bounds__(WG_THREADS) void k_batch_gemm_narrow(const UnitDesc *units,
                                                                  const uint32_t *item_unit,
                                                                  const uint32_t *items)
{
    __shared__ double smem[LDS_GEMM];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_gemm_narrow_item(d, (int)(item - d.item_base), smem);
}

// Panel solves of order <= SMALL_MAX (inside nested diagonal factorizations).
__global__ __launch_bounds__(WG_THREADS) void k_batch_trsm_small(const UnitDesc *units,
                                                                 const uint32_t *item_unit,
                                                                 const uint32_t *items)
{
    __shared__ double smem[LDS_TRSM_SMALL];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_trsm_small_item(d, (int)(item - d.item_base), smem);
}

constexpr int TRSM_KERNEL_WAVES = 8;
__global__ __launch_bounds__(64 * TRSM_KERNEL_WAVES) void k_batch_trsm(const UnitDesc *units,
                                                                      const uint32_t *item_unit,
                                                                      const uint32_t *items)
{
    __shared__ double smem[TRSM_NMAX * 16];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_trsm_inv_item<TRSM_KERNEL_WAVES>(d, (int)(item - d.item_base), smem);
}

// Panel solves with inverted diagonal blocks of order <= NMAX, and the TRINV
// items before them, on 4 waves and at most NMAX * 128 bytes of LDS
// (k_batch_trsm_256: 34 KB, k_batch_trsm_512: 64 KB), so GEMM workgroups keep
// running beside them on the same CU.
template <int NMAX>
__global__ __launch_bounds__(WG_THREADS) void k_batch_trsm_mid(const UnitDesc *units,
                                                               const uint32_t *item_unit,
                                                               const uint32_t *items)
{
    constexpr int LDSD = NMAX * 16 > TRINV_BLOCKS * 16 * 17 ? NMAX * 16 : TRINV_BLOCKS * 16 * 17;
    __shared__ double smem[LDSD];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    if (d.op == OP_TRINV)
        run_trinv_item(d, (int)(item - d.item_base), smem);
    else
        run_trsm_inv_item<4, NMAX>(d, (int)(item - d.item_base), smem);
}

static bool gemm_only_op(int op) { return op == OP_GEMM || op == OP_SYRK; }

static const void *level_kernel(int kind)
{
    switch (kind) {
    case LEVEL_GEMM: return reinterpret_cast<const void *>(k_batch_gemm);
    case LEVEL_SMALL: return reinterpret_cast<const void *>(k_batch_small);
    case LEVEL_TRSM: return reinterpret_cast<const void *>(k_batch_trsm);
    case LEVEL_TRSM_SMALL: return reinterpret_cast<const void *>(k_batch_trsm_small);
    case LEVEL_GEMM_NARROW: return reinterpret_cast<const void *>(k_batch_gemm_narrow);
    case LEVEL_TRSM_256: return reinterpret_cast<const void *>(k_batch_trsm_mid<256>);
    case LEVEL_TRSM_512: return reinterpret_cast<const void *>(k_batch_trsm_mid<512>);
    default: return reinterpret_cast<const void *>(k_batch);
    }
}
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

// ---------------------------------------------------------------------------
// persistent scheduler (protocol in mu_sched.hpp)
// ---------------------------------------------------------------------------
struct DevAtom {
    static __device__ __forceinline__ uint32_t load(uint32_t *p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT);
    }
    static __device__ __forceinline__ uint32_t load_acquire(uint32_t *p)
    {
        return __hip_atomic_load(p, __ATOMIC_ACQUIRE, AGENT);
    }
    static __device__ __forceinline__ void store(uint32_t *p, uint32_t v)
    {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, AGENT);
    }
    static __device__ __forceinline__ uint32_t add(uint32_t *p, uint32_t v)
    {
        return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, AGENT);
    }
    static __device__ __forceinline__ uint32_t add_release(uint32_t *p, uint32_t v)
    {
        return __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, AGENT);
    }
    static __device__ __forceinline__ int32_t add_acq_rel(int32_t *p, int32_t v)
    {
        return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, AGENT);
    }
    // compare-and-swap; returns the value found (== expected on success)
    static __device__ __forceinline__ uint32_t cas(uint32_t *p, uint32_t expected,
                                                   uint32_t desired)
    {
        __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, AGENT);
        return expected;
    }
    static __device__ __forceinline__ void fence_release()
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    }
    static __device__ __forceinline__ uint64_t now() { return (uint64_t)wall_clock64(); }
    static __device__ __forceinline__ void pause() { __builtin_amdgcn_s_sleep(2); }
};

__global__ __launch_bounds__(WG_THREADS) void k_persistent(ExecView ev)
{
    __shared__ double smem[LDS_DOUBLES];
    __shared__ int s_item, s_unit, s_last;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
            const int it = sched_pop<DevAtom>(ev);
            s_item = it;
            s_unit = it >= 0 ? (int)ev.item_unit[it] : -1;
        }
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (item < 0)
            break;
        const UnitDesc d = ev.units[u];
        if (ev.trace && tid == 0)
            ev.trace[2 * (size_t)item] = wall_clock64();
        run_item(d, item - (int)d.item_base, smem);
        // make this item's tile writes visible device-wide before it is
        // counted as complete
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            if (ev.trace)
                ev.trace[2 * (size_t)item + 1] = wall_clock64();
            s_last = sched_item_done<DevAtom>(ev, (uint32_t)u);
        }
        __syncthreads();
        if (s_last)
            sched_release<DevAtom>(ev, (uint32_t)u, (uint32_t)tid, WG_THREADS);
        __syncthreads();
        if (tid == 0)
            sched_count<DevAtom>(ev);
    }
}

struct ResetArgs {
    uint32_t init_tail[NPRIO];
    const uint32_t *slots_init;
    uint32_t *slots_flat;
    uint32_t slots_total;
};

__global__ void k_reset(ExecView ev, ResetArgs ra)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t u = i0; u < ev.nunits; u += stride) {
        ev.dep_count[u] = ev.units[u].ndeps;
        ev.items_left[u] = ev.units[u].nitems;
    }
    for (size_t s = i0; s < ra.slots_total; s += stride)
        ra.slots_flat[s] = ra.slots_init[s];
    if (i0 == 0) {
        for (int p = 0; p < NPRIO; p++) {
            ev.st->head[p] = 0;
            ev.st->tail[p] = ra.init_tail[p];
        }
        ev.st->done = 0;
        ev.st->abort = 0;
        ev.st->err = 0;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int run_units(const Lowered &low, hipStream_t s, std::string &err)
{
    // units of an executable come patched; an eager call gets a
    // stream-ordered workspace for the life of its launches
    const size_t wsd = ws_total_doubles(low);
    bool patched = true;
    for (const UnitDesc &d : low.units)
        if (d.ws_slot >= 0 && ((d.op == OP_TRINV && !d.c) || (d.op == OP_TRSM && !d.b)))
            patched = false;
    if (wsd && !patched) {
        double *ws = nullptr;
        HIPCHK(hipMallocAsync((void **)&ws, wsd * sizeof(double), s));
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
        if (gemm_only_op(d.op) && !(d.flags & F_NARROW))
            hipLaunchKernelGGL(k_unit_gemm, dim3(d.nitems), dim3(WG_THREADS), 0, s, d);
        else
            hipLaunchKernelGGL(k_unit, dim3(d.nitems), dim3(WG_THREADS), 0, s, d);
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
        // GEMM/SYRK task (not a piece of a nested factorization or solve) or
        // the conversion of a panel tile to its reduced-precision shadow;
        // such a unit whose successors are all update units is bulk work ...
        auto update_task = [&](uint32_t u) {
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
        if (eopts.lookahead)
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
        // the kernel a unit's items run on: GEMM/SYRK blocks on k_batch_gemm,
        // small diagonal factorizations / conversions on k_batch_small, panel
        // solves with inverted diagonal blocks (and the inversions, TRINV) on
        // the 4-wave k_batch_trsm_mid<256 / 512> or, for larger triangles, the
        // 8-wave k_batch_trsm, small solves without inverses on
        // k_batch_trsm_small; anything else on the generic k_batch
        auto ukind = [&](const UnitDesc &d) -> int {
            if (d.op == OP_GEMM && (d.flags & F_NARROW))
                return LEVEL_GEMM_NARROW;
            if (gemm_only_op(d.op))
                return LEVEL_GEMM;
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
        // level; a unit whose predecessor sits in the other lane adds an edge
        // between the two nodes.
        const size_t NK = LEVEL_KINDS;
        const size_t nkey = 2 * NK * (size_t)std::max(nlev, 1);
        auto ukey = [&](uint32_t u) {
            return (2 * (size_t)level[u] + lane[u]) * NK + ukind(L.units[u]);
        };
        std::vector<int> node_of(nkey, -1);
        std::vector<std::vector<uint32_t>> lv;
        std::vector<int> node_lane, node_level;
        {
            std::vector<std::vector<uint32_t>> by(nkey);
            for (uint32_t u = 0; u < nu; u++)
                for (int i = 0; i < L.units[u].nitems; i++)
                    by[ukey(u)].push_back(L.units[u].item_base + i);
            for (size_t k = 0; k < by.size(); k++)
                if (!by[k].empty()) {
                    node_of[k] = (int)lv.size();
                    node_lane.push_back((int)((k / NK) & 1));
                    node_level.push_back((int)(k / NK / 2));
                    e->level_gemm.push_back((char)(k % NK));
                    lv.push_back(std::move(by[k]));
                }
        }
        const int nnode = (int)lv.size();
        // node of a unit; units without items (none are lowered) map to -1
        auto unode = [&](uint32_t u) { return node_of[ukey(u)]; };
        std::vector<std::vector<int>> node_deps(nnode);
        {
            // per lane: the nodes of the last level seen, and of the level before
            std::vector<int> prev[2], cur[2];
            int cur_level[2] = {-1, -1};
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
        const size_t o_units = take(sizeof(UnitDesc) * std::max<uint32_t>(nu, 1));
        const size_t o_iu = take(sizeof(uint32_t) * std::max<size_t>(L.nitems, 1));
        const size_t o_items = take(sizeof(uint32_t) * std::max<size_t>(items.size(), 1));
        e->dbytes = o;
        HIPCHK(hipMalloc(&e->dbuf, e->dbytes));
        if (nu)
            HIPCHK(hipMemcpy(e->dbuf + o_units, L.units.data(), sizeof(UnitDesc) * nu,
                             hipMemcpyHostToDevice));
        if (L.nitems)
            HIPCHK(hipMemcpy(e->dbuf + o_iu, item_unit.data(), sizeof(uint32_t) * L.nitems,
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
            kp.blockDim = dim3(e->level_gemm[l] == LEVEL_TRSM ? 64 * TRSM_KERNEL_WAVES : WG_THREADS);
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

    // persistent mode: one device allocation for descriptors and state
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = align_up(off, 256);
        off = o + bytes;
        return o;
    };
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
    const size_t o_trace = eopts.trace ? take(sizeof(uint64_t) * 2 * L.nitems) : 0;
    e->dbytes = off;
    HIPCHK(hipMalloc(&e->dbuf, e->dbytes));

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

    ExecView &ev = e->ev;
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
    ev.st = reinterpret_cast<SchedState *>(e->dbuf + o_st);
    ev.nunits = nu;
    ev.nitems = (uint32_t)L.nitems;
    e->trace = eopts.trace ? reinterpret_cast<uint64_t *>(e->dbuf + o_trace) : nullptr;
    ev.trace = e->trace;
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->device));
    // the constant wall clock runs at 1 MHz .. 1 GHz; anything else is a
    // misreported rate, and 100 MHz (the CDNA value) is used instead
    e->wall_khz = (khz >= 1000 && khz <= 1000000) ? (uint64_t)khz : 100000;
    ev.timeout_ticks = (uint64_t)(eopts.timeout_s * 1000.0 * (double)e->wall_khz);
    // backstop independent of the clock: one empty poll takes well over 50 ns
    ev.poll_limit = (uint64_t)(eopts.timeout_s * 2.0e7) + 1000;

    int ncu = 0, per_cu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, e->device));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persistent, WG_THREADS, 0));
    if (per_cu < 1)
        per_cu = 1;
    e->grid = eopts.grid > 0 ? eopts.grid : ncu * per_cu;
    // the tail values the reset kernel restores before every run
    std::memcpy(e->init_tail, fill, sizeof fill);
    *out = e;
    return 0;
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
    ResetArgs ra;
    std::memcpy(ra.init_tail, e->init_tail, sizeof ra.init_tail);
    ra.slots_init = e->slots_init;
    ra.slots_flat = e->ev.slots[0];
    ra.slots_total = e->slots_total;
    const size_t work = std::max<size_t>(e->ev.nunits, e->slots_total);
    const int rg = (int)std::min<size_t>(1024, (work + 255) / 256 + 1);
    hipLaunchKernelGGL(k_reset, dim3(rg), dim3(256), 0, s, e->ev, ra);
    HIPCHK(hipGetLastError());
    if (e->ev.nitems == 0)
        return 0;
    hipLaunchKernelGGL(k_persistent, dim3(e->grid), dim3(WG_THREADS), 0, s, e->ev);
    HIPCHK(hipGetLastError());
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
                           dim3(e->level_gemm[level] == LEVEL_TRSM ? 64 * TRSM_KERNEL_WAVES
                                                                   : WG_THREADS),
                           args, 0, s));
    HIPCHK(hipGetLastError());
    return 0;
}

int exec_check(Exec *e, std::string &err)
{
    if (e->mode != MODE_PERSISTENT)
        return 0;
    SchedState st;
    HIPCHK(hipMemcpy(&st, e->ev.st, sizeof st, hipMemcpyDeviceToHost));
    if (st.err || st.done != e->ev.nitems) {
        err = "persistent scheduler stopped after " + std::to_string(st.done) + " of " +
              std::to_string(e->ev.nitems) + " items (no progress for the time-out)";
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
    if (e->gexec)
        (void)hipGraphExecDestroy(e->gexec);
    if (e->graph)
        (void)hipGraphDestroy(e->graph);
    if (e->dbuf)
        (void)hipFree(e->dbuf);
    if (e->ws)
        (void)hipFree(e->ws);
    delete e;
}

// ---------------------------------------------------------------------------
// generators, residuals and small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void k_gen_uniform(double *M, size_t count, uint64_t seed, double lo, double hi)
{
    const uint64_t s0 = sm64(seed);
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (size_t)gridDim.x * blockDim.x) {
        const uint64_t z = sm64(s0 + k);
        M[k] = lo + (hi - lo) * ((double)(z >> 11) * (1.0 / 9007199254740992.0));
    }
}

int gen_uniform_device(double *M, size_t count, uint64_t seed, double lo, double hi,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_gen_uniform, dim3(2048), dim3(256), 0, s, M, count, seed, lo, hi);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// A <- A / div + shift I
__global__ void k_add_diag_scale(double *A, int n, size_t lda, double div, double shift)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        double *p = A + i + j * lda;
        *p = *p / div + (i == j ? shift : 0.0);
    }
}

int add_diag_scale(double *A, int n, int lda, double div, double shift, hipStream_t s)
{
    hipLaunchKernelGGL(k_add_diag_scale, dim3(2048), dim3(256), 0, s, A, n, (size_t)lda, div,
                       shift);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_symmetrize_lower(double *A, int n, size_t lda)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        if (i < j)
            A[i + j * lda] = A[j + i * lda];
    }
}

int symmetrize_lower(double *A, int n, int lda, hipStream_t s)
{
    hipLaunchKernelGGL(k_symmetrize_lower, dim3(2048), dim3(256), 0, s, A, n, (size_t)lda);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_fill_int(int32_t *p, size_t count, int32_t v)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = v;
}

int fill_int(int32_t *p, size_t count, int32_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, s, p, count, v);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// F packed (LU or Cholesky) -> explicit L (unit lower for LU) and U (upper
// for LU; unused for Cholesky).
__global__ void k_split_lu(const double *F, int n, size_t ldf, double *L, double *U, size_t ld,
                           int chol)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        const double f = F[i + j * ldf];
        if (chol) {
            L[i + j * ld] = i >= j ? f : 0.0;
        } else {
            L[i + j * ld] = i > j ? f : (i == j ? 1.0 : 0.0);
            U[i + j * ld] = i <= j ? f : 0.0;
        }
    }
}

int split_lu(const double *F, int n, int ldf, double *L, double *U, int ld, int chol,
             hipStream_t s)
{
    hipLaunchKernelGGL(k_split_lu, dim3(4096), dim3(256), 0, s, F, n, (size_t)ldf, L, U,
                       (size_t)ld, chol);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_sub_frob2(const double *A, size_t lda, const double *B, size_t ldb, int m,
                            int n, double *acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % m, j = k / m;
        const double d = B ? A[i + j * lda] - B[i + j * ldb] : A[i + j * lda];
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(acc, red[0]);
}

int sub_frob2(const double *A, int lda, const double *B, int ldb, int m, int n, double *acc,
              hipStream_t s)
{
    hipLaunchKernelGGL(k_sub_frob2, dim3(2048), dim3(256), 0, s, A, (size_t)lda, B, (size_t)ldb,
                       m, n, acc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_sum_abs(const double *A, size_t lda, int m, int n, double *acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * n;
         k += (size_t)gridDim.x * blockDim.x)
        s += fabs(A[k % m + (k / m) * lda]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(acc, red[0]);
}

int sum_abs(const double *A, int lda, int m, int n, double *acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_abs, dim3(2048), dim3(256), 0, s, A, (size_t)lda, m, n, acc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Column-major copy: workgroup (row block of 1024, column chunk) walks its
// columns; each thread moves 4 rows of a column per step (no index division
// per element, 8 KB contiguous per wave and column).
__global__ void k_copy_matrix(const double *src, size_t lds, double *dst, size_t ldd, int m,
                              int n)
{
    const int i0 = blockIdx.x * 1024 + threadIdx.x;
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        const double *s = src + (size_t)j * lds;
        double *d = dst + (size_t)j * ldd;
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = i0 + 256 * q;
            v[q] = i < m ? s[i] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = i0 + 256 * q;
            if (i < m)
                d[i] = v[q];
        }
    }
}

int copy_matrix(const double *src, int lds, double *dst, int ldd, int m, int n, hipStream_t s)
{
    if (m <= 0 || n <= 0)
        return 0;
    const int gx = (m + 1023) / 1024;
    const int gy = std::min(n, std::max(1, 16384 / gx));
    hipLaunchKernelGGL(k_copy_matrix, dim3(gx, gy), dim3(256), 0, s, src, (size_t)lds, dst,
                       (size_t)ldd, m, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Column maxima of |M| (m x n): out[j] = max_i |M(i, j)|.  Non-negative
// doubles order like their bit patterns, so an integer atomic max combines
// the per-thread maxima; out must be zeroed first.
__global__ void k_colmax_abs(const double *M, size_t ld, int m, int n, unsigned long long *out)
{
    const int j = blockIdx.y;
    double v = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        v = fmax(v, fabs(M[i + j * ld]));
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(out + j, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

int colmax_abs(const double *M, int ld, int m, int n, double *out, hipStream_t s)
{
    if (n <= 0)
        return 0;
    unsigned long long *o = reinterpret_cast<unsigned long long *>(out);
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, o, n, 0ull);
    const int gx = std::max(1, std::min(256, (m + 255) / 256));
    hipLaunchKernelGGL(k_colmax_abs, dim3(gx, n), dim3(256), 0, s, M, (size_t)ld, m, n, o);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ||A||_inf = max_i sum_j |A(i, j)|.  Pass 1: thread (row i, column chunk c)
// sums |A(i, j)| over the chunk's INF_KC columns into W[c m + i] (a wave
// reads 64 consecutive rows of a column: coalesced; eight loads in flight).
// Pass 2: per row, the chunks in order, then the maximum (an atomic max of
// non-negative doubles, order-independent, on their bit patterns).
constexpr int INF_KC = 256;

__global__ void k_rowsum_partial(const double *A, size_t lda, int m, int n, double *W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j0 = blockIdx.y * INF_KC, j1 = min(n, j0 + INF_KC);
    if (i >= m)
        return;
    const double *a = A + i;
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int j = j0;
    for (; j + 8 <= j1; j += 8)
#pragma unroll
        for (int u = 0; u < 8; u++)
            s[u] += fabs(a[(size_t)(j + u) * lda]);
    for (; j < j1; j++)
        s[0] += fabs(a[(size_t)j * lda]);
    W[(size_t)blockIdx.y * m + i] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

__global__ void k_rowsum_max(const double *W, int m, int nchunk, unsigned long long *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sum = 0.0;
    if (i < m)
        for (int c = 0; c < nchunk; c++)
            sum += W[(size_t)c * m + i];
    __shared__ double red[256];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

size_t inf_norm_work_doubles(int m, int n) { return (size_t)((n + INF_KC - 1) / INF_KC) * m; }

int inf_norm(const double *A, int lda, int m, int n, double *out, double *work, hipStream_t s)
{
    unsigned long long *o = reinterpret_cast<unsigned long long *>(out);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, s, o, 1, 0ull);
    if (m <= 0 || n <= 0)
        return hipGetLastError() == hipSuccess ? 0 : -1;
    const int nchunk = (n + INF_KC - 1) / INF_KC;
    hipLaunchKernelGGL(k_rowsum_partial, dim3((m + 255) / 256, nchunk), dim3(256), 0, s, A,
                       (size_t)lda, m, n, work);
    hipLaunchKernelGGL(k_rowsum_max, dim3((m + 255) / 256), dim3(256), 0, s, work, m, nchunk, o);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// X += D (m x n)
__global__ void k_add_matrix(double *X, size_t ldx, const double *D, size_t ldd, int m, int n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % m, j = k / m;
        X[i + j * ldx] += D[i + j * ldd];
    }
}

int add_matrix(double *X, int ldx, const double *D, int ldd, int m, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_add_matrix, dim3(1024), dim3(256), 0, s, X, (size_t)ldx, D, (size_t)ldd,
                       m, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Residual R = B - A X for up to 16 right-hand sides, split over K so the
// n x n read of A (the whole cost) runs on thousands of workgroups.  Pass 1:
// workgroup (256 rows, K chunk c) forms P_c = A[rows, chunk] X[chunk, :]:
// thread r reads row r of A down the chunk (a wave reads 64 consecutive rows
// of a column: 512 contiguous bytes) and keeps 16 sums, one per right-hand
// side, against X rows staged in LDS 128 at a time (read as broadcasts).
// The stores go to the partial buffer P[c] (column-major, leading dimension
// m).  Pass 2 sums the chunks in a fixed order, so the result does not
// depend on scheduling.
constexpr int RES_KC = 1024;  // K chunk of one workgroup
constexpr int RES_KS = 128;   // X rows staged in LDS at a time

__global__ __launch_bounds__(256) void k_residual_partial(const double *A, size_t lda,
                                                          const double *X, size_t ldx, int m,
                                                          int n, int nrhs, double *P)
{
    __shared__ double xs[RES_KS * 16];
    const int tid = threadIdx.x, c = blockIdx.y;
    const int r = blockIdx.x * 256 + tid;
    const bool rok = r < m;
    const double *a = A + min(r, m - 1);
    const int k0 = c * RES_KC, k1 = min(n, k0 + RES_KC);
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; j++)
        acc[j] = 0.0;
    for (int ks = k0; ks < k1; ks += RES_KS) {
        // stage X[ks .. ks + RES_KS)[0 .. 16) (zero past the chunk / nrhs)
#pragma unroll
        for (int q = 0; q < RES_KS * 16 / 256; q++) {
            const int idx = tid + 256 * q, kq = idx >> 4, j = idx & 15;
            const int kk = ks + kq;
            xs[idx] = (kk < k1 && j < nrhs) ? X[(size_t)kk + (size_t)j * ldx] : 0.0;
        }
        lds_barrier();
        const int kn = min(RES_KS, k1 - ks);
        for (int kq = 0; kq < kn; kq += 8) {
            double av[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int kk = ks + kq + u;
                av[u] = (rok && kq + u < kn) ? a[(size_t)min(kk, n - 1) * lda] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const double *xr = xs + (kq + u) * 16;
#pragma unroll
                for (int j = 0; j < 16; j++)
                    acc[j] = fma(av[u], xr[j], acc[j]);
            }
        }
        lds_barrier();
    }
    if (rok) {
        double *p = P + (size_t)c * m * 16 + r;
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (j < nrhs)
                p[(size_t)j * m] = acc[j];
    }
}

__global__ void k_residual_reduce(const double *B, size_t ldb, const double *P, int m, int nrhs,
                                  int nchunk, double *R, size_t ldr)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * nrhs;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % m, j = k / m;
        double sum = 0.0;
        for (int c = 0; c < nchunk; c++)
            sum += P[(size_t)c * m * 16 + i + j * m];
        R[i + j * ldr] = B[i + j * ldb] - sum;
    }
}

size_t residual_partial_doubles(int m, int n)
{
    return (size_t)((n + RES_KC - 1) / RES_KC) * m * 16;
}

int residual(const double *A, int lda, const double *X, int ldx, const double *B, int ldb,
             double *R, int ldr, int m, int n, int nrhs, double *P, hipStream_t s)
{
    if (m <= 0 || nrhs <= 0)
        return 0;
    const int nchunk = (n + RES_KC - 1) / RES_KC;
    for (int j0 = 0; j0 < nrhs; j0 += 16) {
        const int w = std::min(16, nrhs - j0);
        if (n > 0)
            hipLaunchKernelGGL(k_residual_partial, dim3((m + 255) / 256, nchunk), dim3(256), 0,
                               s, A, (size_t)lda, X + (size_t)j0 * ldx, (size_t)ldx, m, n, w, P);
        hipLaunchKernelGGL(k_residual_reduce, dim3(1024), dim3(256), 0, s,
                           B + (size_t)j0 * ldb, (size_t)ldb, P, m, w, n > 0 ? nchunk : 0,
                           R + (size_t)j0 * ldr, (size_t)ldr);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Conditioned test matrices: A = diag(s), then A <- H A H for each reflector.
__global__ void k_diag_geometric(double *A, size_t lda, int n, double kappa)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        A[i + j * lda] = i == j ? (n > 1 ? pow(kappa, -(double)i / (double)(n - 1)) : 1.0) : 0.0;
    }
}

// acc[0] += x . y
__global__ void k_dot(const double *x, const double *y, int n, double *acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += x[i] * y[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(acc, red[0]);
}

// A <- H A H with H = I - tau v v^T, tau = 2 / v.v, given w = A v and
// c = v.w (sc[0] = v.v, sc[1] = v.w):
//   A - tau (v w^T + w v^T) + tau^2 c v v^T.
// The products are rounded separately (contraction into fused multiply-adds
// is off here), so A(i, j) and A(j, i) stay bit-for-bit equal.
__global__ void k_reflect2(double *A, size_t lda, int n, const double *v, const double *w,
                           const double *sc)
{
#pragma clang fp contract(off)
    const double tau = 2.0 / sc[0], c = sc[1];
    const double t2c = (tau * tau) * c;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        const double p1 = v[i] * w[j], p2 = w[i] * v[j];
        const double s1 = p1 + p2;
        const double s2 = v[i] * v[j];
        const double d1 = tau * s1, d2 = t2c * s2;
        A[i + j * lda] = (A[i + j * lda] - d1) + d2;
    }
}

// dst = tril(src) with the diagonal scaled by dscale and zeros above it
__global__ void k_tril_scaled(const double *src, size_t lds, double *dst, size_t ldd, int n,
                              double dscale)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        const double v = src[i + j * lds];
        dst[i + j * ldd] = i > j ? v : i == j ? dscale * v : 0.0;
    }
}

int tril_scaled(const double *src, int lds, double *dst, int ldd, int n, double dscale,
                hipStream_t s)
{
    hipLaunchKernelGGL(k_tril_scaled, dim3(4096), dim3(256), 0, s, src, (size_t)lds, dst,
                       (size_t)ldd, n, dscale);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// dst = (src + src^T) / 2 (out of place)
__global__ void k_sym_avg(const double *src, size_t lds, double *dst, size_t ldd, int n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        dst[i + j * ldd] = 0.5 * (src[i + j * lds] + src[j + i * lds]);
    }
}

int sym_avg(const double *src, int lds, double *dst, int ldd, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_sym_avg, dim3(4096), dim3(256), 0, s, src, (size_t)lds, dst, (size_t)ldd,
                       n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// t consecutive b x b identity tiles (tile k at D + k b^2, leading dimension b)
__global__ void k_identity_tiles(double *D, int b, int t)
{
    const size_t total = (size_t)b * b * t;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t e = k % ((size_t)b * b);
        D[k] = (e % b) == (e / b) ? 1.0 : 0.0;
    }
}

int identity_tiles(double *D, int b, int t, hipStream_t s)
{
    hipLaunchKernelGGL(k_identity_tiles, dim3(2048), dim3(256), 0, s, D, b, t);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gen_cond_start(double *A, int n, int lda, double kappa, hipStream_t s)
{
    hipLaunchKernelGGL(k_diag_geometric, dim3(4096), dim3(256), 0, s, A, (size_t)lda, n, kappa);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dot_acc(const double *x, const double *y, int n, double *acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot, dim3(std::max(1, std::min(1024, (n + 255) / 256))), dim3(256), 0, s,
                       x, y, n, acc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int reflect2(double *A, int lda, int n, const double *v, const double *w, const double *sc,
             hipStream_t s)
{
    hipLaunchKernelGGL(k_reflect2, dim3(4096), dim3(256), 0, s, A, (size_t)lda, n, v, w, sc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mu
