// broker_queue_check -- exercises BrokerWorkDistributor (include/
// mustard_queue.h) on the GPU, the way the reference's kernels use it
// (one lane per workgroup, mustard.h:70-129).
//
//   fill:  P producer workgroups enqueue ids into a ring that holds all of
//          them; a second launch of C consumers drains it.
//   mixed: P producers and C consumers run at once on a ring of `cap`
//          slots (far fewer than the items), so every slot is reused many
//          times while writers and readers race for it.
// Every id must come out exactly once, size() must read 0 at the end and no
// wait may give up.  Each wait is bounded (spin_limit) and every loop has an
// iteration cap, so the kernels end even if the queue were wrong; the
// program prints one line per case and exits non-zero on any failure.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/mustard_queue.h"

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                  \
            std::exit(2);                                                           \
        }                                                                           \
    } while (0)

constexpr unsigned LOOP_CAP = 1u << 20;

struct Counters {
    unsigned consumed;  // items taken so far
    unsigned capped;    // loops that hit LOOP_CAP
    int size_end;       // size() read after the run
};

__global__ void produce(BrokerWorkDistributor q, unsigned per, Counters *c)
{
    if (threadIdx.x != 0)
        return;
    for (unsigned i = 0; i < per; i++) {
        const unsigned id = blockIdx.x * per + i;
        unsigned tries = 0;
        while (!q.enqueue(id, 0)) {
            if (__hip_atomic_load(q.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                ++tries >= LOOP_CAP) {
                atomicAdd(&c->capped, 1u);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

__device__ void consume_loop(BrokerWorkDistributor &q, unsigned total, unsigned *seen,
                             Counters *c)
{
    if (threadIdx.x != 0)
        return;
    for (unsigned it = 0;; it++) {
        if (__hip_atomic_load(&c->consumed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                total ||
            __hip_atomic_load(q.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return;
        if (it >= LOOP_CAP) {
            atomicAdd(&c->capped, 1u);
            return;
        }
        bool has = false;
        unsigned v = 0;
        q.dequeue(has, v, 0);
        if (has) {
            if (v < total)
                atomicAdd(&seen[v], 1u);
            __hip_atomic_fetch_add(&c->consumed, 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

__global__ void consume(BrokerWorkDistributor q, unsigned total, unsigned *seen, Counters *c)
{
    consume_loop(q, total, seen, c);
}

// Producers and consumers in one launch: blocks [0, P) produce, the rest
// consume (all workgroups are resident: P + C is far below the CU count).
__global__ void mixed(BrokerWorkDistributor q, unsigned P, unsigned per, unsigned total,
                      unsigned *seen, Counters *c)
{
    if (blockIdx.x < P) {
        if (threadIdx.x != 0)
            return;
        for (unsigned i = 0; i < per; i++) {
            const unsigned id = blockIdx.x * per + i;
            unsigned tries = 0;
            while (!q.enqueue(id, 0)) {
                if (__hip_atomic_load(q.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    ++tries >= LOOP_CAP) {
                    atomicAdd(&c->capped, 1u);
                    return;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        return;
    }
    consume_loop(q, total, seen, c);
}

__global__ void read_size(BrokerWorkDistributor q, Counters *c) { c->size_end = q.size(0); }

static int run_case(const char *name, bool concurrent, unsigned P, unsigned C, unsigned per,
                    int cap)
{
    const unsigned total = P * per;
    BrokerWorkDistributor q(cap);
    if (!q.valid()) {
        std::printf("%s: allocation failed\n", name);
        return 1;
    }
    q.spin_limit = 1u << 20;
    unsigned *seen = nullptr;
    Counters *c = nullptr;
    CHECK(hipMalloc(&seen, total * sizeof(unsigned)));
    CHECK(hipMalloc(&c, sizeof(Counters)));
    CHECK(hipMemset(seen, 0, total * sizeof(unsigned)));
    CHECK(hipMemset(c, 0, sizeof(Counters)));
    if (concurrent) {
        hipLaunchKernelGGL(mixed, dim3(P + C), dim3(64), 0, 0, q, P, per, total, seen, c);
    } else {
        hipLaunchKernelGGL(produce, dim3(P), dim3(64), 0, 0, q, per, c);
        hipLaunchKernelGGL(consume, dim3(C), dim3(64), 0, 0, q, total, seen, c);
    }
    CHECK(hipGetLastError());
    hipLaunchKernelGGL(read_size, dim3(1), dim3(1), 0, 0, q, c);
    CHECK(hipDeviceSynchronize());
    std::vector<unsigned> h(total);
    Counters hc;
    int herr = 0;
    CHECK(hipMemcpy(h.data(), seen, total * sizeof(unsigned), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(&hc, c, sizeof hc, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(&herr, q.error, sizeof herr, hipMemcpyDeviceToHost));
    unsigned once = 0, missing = 0, dup = 0;
    for (unsigned v : h) {
        once += v == 1;
        missing += v == 0;
        dup += v > 1;
    }
    const bool ok = once == total && hc.consumed == total && hc.capped == 0 && herr == 0 &&
                    hc.size_end == 0;
    std::printf("%s: items %u once %u missing %u duplicated %u consumed %u capped %u "
                "gave_up %d size_end %d -> %s\n",
                name, total, once, missing, dup, hc.consumed, hc.capped, herr, hc.size_end,
                ok ? "ok" : "FAIL");
    CHECK(hipFree(seen));
    CHECK(hipFree(c));
    q.free_mem();
    return ok ? 0 : 1;
}

int main()
{
    int bad = 0;
    bad += run_case("fill", false, 64, 64, 100, 6400);
    bad += run_case("mixed", true, 64, 64, 500, 256);
    bad += run_case("mixed_tiny_ring", true, 32, 96, 200, 7);
    return bad ? 1 : 0;
}
