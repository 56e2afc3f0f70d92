// mu_sched.hpp -- the device work queue and dependency counters, written once
// for two targets: the gfx950 persistent kernel (atomics policy DevAtom in
// mu_runtime.hip) and a host emulation in which every std::thread plays one
// workgroup (HostAtom, mu_sched_host.cpp), so the protocol is tested on the
// CPU as well.
//
// Protocol (replaces mustard/broker_queue.h and kernel_scheduler /
// kernel_dep_update, mustard/mustard.h:70-129):
//  * one ring per priority; a ring holds every item of its priority once per
//    run, so it never wraps; a producer reserves positions with one fetch_add
//    on the tail and fills them, a consumer claims position head with a CAS
//    and waits for the value to appear;
//  * an item's completion decrements its unit's items_left; the last item of a
//    unit decrements the dependency count of each successor unit, and a unit
//    whose count reaches zero has all its items queued;
//  * `done` counts finished items; a workgroup leaves when done == nitems, or
//    when it has seen no progress for timeout_ticks (abort) or has polled
//    poll_limit times without progress.
#pragma once
#include "mu_types.h"

#ifdef __HIP__  // HIP translation unit: callable from the kernel and the host
#include <hip/hip_runtime.h>
#define MU_HD __host__ __device__ __forceinline__
#else
#define MU_HD inline
#endif

namespace mu {

// Queue all items of unit u on its priority ring.
template <class AT>
MU_HD void sched_push_unit(const ExecView &ev, uint32_t u)
{
    const UnitDesc *d = ev.units + u;
    const int p = d->prio;
    const uint32_t cnt = (uint32_t)d->nitems, base = d->item_base;
    const uint32_t pos = AT::add(&ev.st->tail[p], cnt);
    AT::fence_release();
    for (uint32_t r = 0; r < cnt; r++)
        AT::store(ev.slots[p] + pos + r, base + r + 1);
}

// One attempt on ring p: an item id, or -1 when the ring is empty (or the
// run was aborted while waiting for a reserved slot).
template <class AT>
MU_HD int sched_try_pop(const ExecView &ev, int p, uint64_t deadline)
{
    uint32_t *head = &ev.st->head[p];
    uint32_t h = AT::load(head);
    for (;;) {
        const uint32_t t = AT::load(&ev.st->tail[p]);
        if (h >= t)
            return -1;
        const uint32_t seen = AT::cas(head, h, h + 1);
        if (seen == h) {
            // position h was reserved by a producer that is filling it now
            uint32_t v;
            while ((v = AT::load_acquire(ev.slots[p] + h)) == 0) {
                if (AT::now() > deadline || AT::load(&ev.st->abort))
                    return -1;
                AT::pause();
            }
            return (int)(v - 1);
        }
        // another workgroup took position h: retry from the head the CAS saw,
        // but never past the deadline or an abort
        h = seen;
        if (AT::now() > deadline || AT::load(&ev.st->abort))
            return -1;
    }
}

template <class AT>
MU_HD void sched_abort(const ExecView &ev, uint32_t code)
{
    AT::store(&ev.st->err, code);
    AT::store(&ev.st->abort, 1u);
}

// Next item for this workgroup, or -1 when the run is over or aborted.
template <class AT>
MU_HD int sched_pop(const ExecView &ev)
{
    uint64_t t0 = AT::now();
    uint32_t seen = AT::load(&ev.st->done);
    uint64_t idle = 0;
    for (;;) {
        if (AT::load(&ev.st->abort))
            return -1;
        for (int p = 0; p < NPRIO; p++) {
            const int r = sched_try_pop<AT>(ev, p, t0 + ev.timeout_ticks);
            if (r >= 0)
                return r;
        }
        const uint32_t d = AT::load(&ev.st->done);
        if (d >= ev.nitems)
            return -1;
        const uint64_t now = AT::now();
        if (d != seen) {
            seen = d;
            t0 = now;
            idle = 0;
        } else if (now - t0 > ev.timeout_ticks) {
            sched_abort<AT>(ev, 1u);
            return -1;
        } else if (++idle > ev.poll_limit) {
            sched_abort<AT>(ev, 2u);
            return -1;
        }
        AT::pause();
    }
}

// Completion, step 1 (one thread): true when `item` was the last of unit u.
template <class AT>
MU_HD bool sched_item_done(const ExecView &ev, uint32_t u)
{
    return AT::add_acq_rel(ev.items_left + u, -1) == 1;
}

// Completion, step 2 (threads first, first + stride, ...): release the
// successors of the finished unit u.
template <class AT>
MU_HD void sched_release(const ExecView &ev, uint32_t u, uint32_t first, uint32_t stride)
{
    const UnitDesc *d = ev.units + u;
    for (uint32_t s = d->succ_begin + first; s < d->succ_end; s += stride) {
        const uint32_t v = ev.succ[s];
        if (AT::add_acq_rel(ev.dep_count + v, -1) == 1)
            sched_push_unit<AT>(ev, v);
    }
}

// Completion, step 3 (one thread, after step 2 of every thread).
template <class AT>
MU_HD void sched_count(const ExecView &ev)
{
    AT::add_release(&ev.st->done, 1u);
}

}  // namespace mu
