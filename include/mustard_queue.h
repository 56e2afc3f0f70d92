// mustard_queue.h -- the reference's work queue type (mustard/broker_queue.h:
// 12-150) for device code on one MI355X.
//
// A bounded multi-producer multi-consumer FIFO of 32-bit task ids in HBM: the
// broker queue (Kerbl, Kenzel, Mueller, Schmalstieg, Steinberger, "The
// Broker Queue: A Fast, Linearizable FIFO Queue for Fine-Granular Work
// Distribution on the GPU", ICS 2018).  A counter admits an operation only
// when the ring has room (enqueue) or an item (dequeue); admitted
// operations then take a position from the tail or head counter, and each
// slot's ticket orders the rounds of writers and readers of that slot
// (2r: free for the write of round r, 2r + 1: holds round r's item).
//
// Same names, members and call forms as the reference (the `pe` argument is
// accepted for source compatibility; one GPU owns the queue).  What differs:
//   * the memory is one zeroed HBM allocation (the reference's NVSHMEM
//     buffers are not initialised, broker_queue.h:26-34), and reset()
//     zeroes it again between runs;
//   * synchronisation is agent-scope acquire/release on the GPU's L2 (the
//     reference's nvshmem_fence / nvshmem_quiet order remote PEs; both
//     consistency modes mean the same here);
//   * enqueue returns whether the item went in (the reference always
//     returns false and drops an item when the ring is full);
//   * `spin_limit` (0: wait forever, as the reference does) bounds each
//     ticket wait; an operation that gives up sets *error and returns, so a
//     kernel whose partner cannot progress still ends.
// Call from one lane per wave (as the reference's kernels do from thread
// 0): a lane spinning on a ticket that another lane of the same wave must
// write would never see it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

enum class BrokerConsistency { FENCE, QUIET };

template <BrokerConsistency Consistency = BrokerConsistency::QUIET>
class BrokerWorkDistributorT {
  public:
    typedef unsigned int Ticket;

    volatile Ticket *tickets = nullptr;
    unsigned int *ring_buffer = nullptr;
    int N = 0;

    unsigned int *head = nullptr;
    unsigned int *tail = nullptr;
    int *count = nullptr;
    int *error = nullptr;         // set to 1 when a wait gave up (spin_limit)
    unsigned int spin_limit = 0;  // ticket polls before giving up; 0 = never

    BrokerWorkDistributorT() = default;

    // One allocation: counters (count, head, tail, error) in the first 256
    // bytes, then N tickets, then the N-entry ring; all zero.
    explicit BrokerWorkDistributorT(int size) : N(size)
    {
        void *p = nullptr;
        if (size <= 0 || hipMalloc(&p, bytes(size)) != hipSuccess) {
            N = 0;
            return;
        }
        base_ = p;
        unsigned char *b = static_cast<unsigned char *>(p);
        count = reinterpret_cast<int *>(b);
        head = reinterpret_cast<unsigned int *>(b + 64);
        tail = reinterpret_cast<unsigned int *>(b + 128);
        error = reinterpret_cast<int *>(b + 192);
        tickets = reinterpret_cast<volatile Ticket *>(b + 256);
        ring_buffer = reinterpret_cast<unsigned int *>(b + 256 + 4 * (size_t)size);
        if (hipMemset(p, 0, bytes(size)) != hipSuccess)
            free_mem();
    }

    // Whether the constructor's allocation succeeded.
    bool valid() const { return base_ != nullptr; }

    // Empty the queue (stream-ordered when a stream is given).
    hipError_t reset(hipStream_t s = nullptr)
    {
        return base_ ? hipMemsetAsync(base_, 0, bytes(N), s) : hipErrorInvalidValue;
    }

    void free_mem()
    {
        if (base_)
            (void)hipFree(base_);
        base_ = nullptr;
        tickets = nullptr;
        ring_buffer = nullptr;
        head = tail = nullptr;
        count = error = nullptr;
    }

    __device__ inline bool enqueue(const unsigned int &data, int pe = 0)
    {
        (void)pe;
        if (!admit(+1, N))
            return false;
        const unsigned int pos = add_u(tail, 1u);
        const unsigned int slot = pos % (unsigned)N, round = pos / (unsigned)N;
        if (!wait_ticket(slot, 2u * round))
            return false;
        ring_buffer[slot] = data;
        release_ticket(slot, 2u * round + 1u);
        return true;
    }

    __device__ inline void dequeue(bool &hasData, unsigned int &data, int pe = 0)
    {
        (void)pe;
        hasData = admit(-1, 0);
        if (!hasData)
            return;
        const unsigned int pos = add_u(head, 1u);
        const unsigned int slot = pos % (unsigned)N, round = pos / (unsigned)N;
        if (!wait_ticket(slot, 2u * round + 1u)) {
            hasData = false;
            return;
        }
        data = __hip_atomic_load(&ring_buffer[slot], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
        release_ticket(slot, 2u * round + 2u);
    }

    // Items admitted and not yet taken (may briefly overshoot by the number
    // of operations in flight).
    __device__ inline int size(int pe = 0) const
    {
        (void)pe;
        return __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

  private:
    void *base_ = nullptr;

    static size_t bytes(int n) { return 256 + 8 * (size_t)(n > 0 ? n : 0); }

    __device__ static inline unsigned int add_u(unsigned int *p, unsigned int v)
    {
        return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    __device__ inline int add_count(int v) const
    {
        return __hip_atomic_fetch_add(count, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // dir +1 (enqueue: room while count < N) or -1 (dequeue: an item while
    // count > 0).  An optimistic add is undone when it overshoots; the
    // undo's result is the fresh count for the next try.
    __device__ inline bool admit(int dir, int bound) const
    {
        int seen = size();
        while (dir > 0 ? seen < bound : seen > bound) {
            const int before = add_count(dir);
            if (dir > 0 ? before < bound : before > bound)
                return true;
            seen = add_count(-dir) - dir;
        }
        return false;
    }

    __device__ inline bool wait_ticket(unsigned int slot, Ticket want) const
    {
        Ticket *t = const_cast<Ticket *>(tickets) + slot;
        unsigned int polls = 0;
        while (__hip_atomic_load(t, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != want) {
            if (spin_limit && ++polls >= spin_limit) {
                __hip_atomic_store(error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        return true;
    }

    __device__ inline void release_ticket(unsigned int slot, Ticket v) const
    {
        __hip_atomic_store(const_cast<Ticket *>(tickets) + slot, v, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
};

using BrokerWorkDistributor = BrokerWorkDistributorT<BrokerConsistency::QUIET>;
using BrokerWorkDistributorFence = BrokerWorkDistributorT<BrokerConsistency::FENCE>;
