// mu_persist.hip -- the persistent scheduler kernel (reference --subgraph).
//
// Replaces the reference's device runtime (mustard/mustard.h:57-129,
// mustard/broker_queue.h): instead of a one-thread scheduler kernel that
// device-launches one CUDA sub-graph per task (HIP has no device-side graph
// launch), every workgroup of one persistent kernel pulls ready work items
// from HBM-resident queue rings, runs them inline, and on completion
// decrements the dependency counters of the successor units, queueing the
// items of every unit whose counter reaches zero.  The protocol and the
// worker loop are in mu_sched.hpp; this file binds them to gfx950 atomics,
// LDS and barriers, and holds the item dispatch.
#include <algorithm>

#include "mu_kernels.hpp"
#include "mu_ops.hpp"
#include "mu_sched.hpp"

namespace mu {

#define AGENT __HIP_MEMORY_SCOPE_AGENT

// Per-lane memory operations at agent scope (the whole GPU: all eight XCDs).
struct DevAtom {
    static __device__ __forceinline__ bool lane0() { return (threadIdx.x & 63) == 0; }
    static __device__ __forceinline__ uint32_t bcast(uint32_t v)
    {
        return __builtin_amdgcn_readfirstlane(v);
    }
    static __device__ __forceinline__ uint32_t load(uint32_t *p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT);
    }
    static __device__ __forceinline__ uint32_t load_acquire(uint32_t *p)
    {
        return __hip_atomic_load(p, __ATOMIC_ACQUIRE, AGENT);
    }
    // the host-written abort word lives in pinned host memory
    static __device__ __forceinline__ uint32_t load_system(uint32_t *p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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
    static __device__ __forceinline__ void producer_delay() {}
};

// Item bodies the persistent kernel runs.  The lowering for this mode keeps
// every item within them: GEMM/SYRK blocks, narrow GEMMs, GETRF/POTRF/CVT of
// order <= SMALL_MAX, panel solves with inverted diagonal blocks of order <=
// PERSIST_TRSM_MAX (and their TRINV items), plain panel solves of order <=
// SMALL_MAX.
constexpr int LDS_PERSIST_A = PERSIST_TRSM_MAX * 16 > TRINV_BLOCKS * 16 * 17
                                  ? PERSIST_TRSM_MAX * 16
                                  : TRINV_BLOCKS * 16 * 17;
constexpr int LDS_PERSIST_B = LDS_GEMM > LDS_TRSM_SMALL ? LDS_GEMM : LDS_TRSM_SMALL;
constexpr int LDS_PERSIST = LDS_PERSIST_A > LDS_PERSIST_B ? LDS_PERSIST_A : LDS_PERSIST_B;
static_assert(LDS_PERSIST >= LDS_SMALL, "persistent LDS holds the small factorizations");

// The kernel's LDS, at file scope so the item bodies below address it as LDS
// without a pointer argument.
__shared__ double p_smem[LDS_PERSIST];

// One function per item kind, each kept out of line: its registers are
// allocated for that body alone, and nothing of it is hoisted into the
// worker loop (inlined into the loop, the loop-invariant parts of every
// body stayed live across items and the GEMM loop spilled).
__device__ __noinline__ void persist_gemm(const UnitDesc *ud, int it)
{
    const UnitDesc d = *ud;
    run_gemm_item<true>(d, it, p_smem);
}
__device__ __noinline__ void persist_gemm_narrow(const UnitDesc *ud, int it)
{
    const UnitDesc d = *ud;
    run_gemm_narrow_item(d, it, p_smem);
}
__device__ __noinline__ void persist_trsm_inv(const UnitDesc *ud, int it)
{
    const UnitDesc d = *ud;
    run_trsm_inv_item<4, PERSIST_TRSM_MAX>(d, it, p_smem);
}
__device__ __noinline__ void persist_trsm_small(const UnitDesc *ud, int it)
{
    const UnitDesc d = *ud;
    run_trsm_small_item(d, it, p_smem);
}
__device__ __noinline__ void persist_trinv(const UnitDesc *ud, int it)
{
    const UnitDesc d = *ud;
    run_trinv_item(d, it, p_smem);
}
__device__ __noinline__ void persist_small(const UnitDesc *ud, int it)
{
    const UnitDesc d = *ud;
    run_small_item(d, it, p_smem);  // GETRF, POTRF, CVT
}

__device__ __forceinline__ void run_persist_item(const UnitDesc *ud, int op, int flags, int it)
{
    switch (op) {
    case OP_GEMM:
        if (flags & F_NARROW)
            persist_gemm_narrow(ud, it);
        else
            persist_gemm(ud, it);
        break;
    case OP_SYRK: persist_gemm(ud, it); break;
    case OP_TRSM:
        if (flags & F_INV)
            persist_trsm_inv(ud, it);
        else
            persist_trsm_small(ud, it);
        break;
    case OP_TRINV: persist_trinv(ud, it); break;
    default: persist_small(ud, it); break;
    }
}

// The per-wave context of sched_worker on the device.
struct DevWave {
    const ExecView &ev;
    int *s_item;
    int *s_last;
    __device__ __forceinline__ int wave() const
    {
        return (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    }
    __device__ __forceinline__ void barrier() const { __syncthreads(); }
    // every lane of wave 0 stores the same (broadcast) value
    __device__ __forceinline__ void set_item(int v) const { *s_item = v; }
    __device__ __forceinline__ int item() const
    {
        return (int)__builtin_amdgcn_readfirstlane(*s_item);
    }
    __device__ __forceinline__ void set_last(int v) const { *s_last = v; }
    __device__ __forceinline__ int last() const
    {
        return (int)__builtin_amdgcn_readfirstlane(*s_last);
    }
    // this wave's stores have completed (reached the L2); the L2 write-back
    // to device scope is wave 0's release that follows the barrier
    __device__ __forceinline__ void fence_release() const
    {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    template <class F>
    __device__ __forceinline__ void threads(F f) const
    {
        f(threadIdx.x, WG_THREADS);
    }
    __device__ __forceinline__ void run(uint32_t u, int item) const
    {
        const UnitDesc *ud = ev.units + u;
        const int op = ud->op, flags = ud->flags, base = (int)ud->item_base;
        if (ev.trace && threadIdx.x == 0)
            ev.trace[2 * (size_t)item] = wall_clock64();
        run_persist_item(ud, op, flags, item - base);
        if (ev.trace && threadIdx.x == 0)
            ev.trace[2 * (size_t)item + 1] = wall_clock64();
    }
};

// Two workgroups per CU: the loop keeps the loop-invariant parts of every
// item body live across items, so the kernel is held to 256 registers per
// lane (the spills this costs sit outside the bodies' inner loops).
__global__ __launch_bounds__(WG_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void
k_persistent(ExecView ev)
{
    __shared__ int s_item, s_last;
    DevWave w{ev, &s_item, &s_last};
    sched_worker<DevAtom>(ev, w);
}

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
    for (size_t t = i0; ev.task_running && t < ra.ntasks; t += stride)
        ev.task_running[t] = 0;
    if (i0 == 0) {
        for (int p = 0; p < NPRIO; p++) {
            ev.st->cnt[p] = ra.init_cnt[p];
            ev.st->head[p] = 0;
            ev.st->tail[p] = (uint32_t)ra.init_cnt[p];
        }
        ev.st->done = 0;
        ev.st->abort = 0;
        ev.st->err = 0;
        ev.st->occ = 0;
    }
}

int persist_blocks_per_cu(int *per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_persistent, WG_THREADS, 0) ==
                   hipSuccess
               ? 0
               : -1;
}

int persist_launch(const ExecView &ev, const ResetArgs &ra, int grid, hipStream_t s)
{
    const size_t work = std::max<size_t>(std::max<size_t>(ev.nunits, ra.slots_total), ra.ntasks);
    const int rg = (int)std::min<size_t>(1024, (work + 255) / 256 + 1);
    hipLaunchKernelGGL(k_reset, dim3(rg), dim3(256), 0, s, ev, ra);
    if (hipGetLastError() != hipSuccess)
        return -1;
    if (ev.nitems == 0)
        return 0;
    hipLaunchKernelGGL(k_persistent, dim3(grid), dim3(WG_THREADS), 0, s, ev);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mu
