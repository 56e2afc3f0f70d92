// mu_sched.hpp -- the device work queue, the dependency counters and the
// persistent worker loop, written once for two targets: the gfx950
// persistent kernel (policy DevAtom, mu_persist.hip) and a host emulation in
// which one std::thread plays one wave and four of them one workgroup
// (HostAtom, mu_sched_host.cpp).  Both run exactly the code below.
//
// Protocol (replaces mustard/broker_queue.h and kernel_populate_queue /
// kernel_scheduler / kernel_dep_update / kernel_occupancy_update,
// mustard/mustard.h:70-129):
//  * one ring per priority; a ring holds every item of its priority once per
//    run, so it never wraps.  `cnt` counts queued, unclaimed items (the
//    broker queue's count, broker_queue.h:58-94): a producer reserves
//    positions with one fetch_add on `tail`, stores the item ids, and only
//    then adds them to `cnt` (release); a consumer first takes one unit of
//    `cnt` (undoing the decrement when it finds none), then a position with
//    one fetch_add on `head`.  A claimed position was reserved by a producer
//    that is running and storing it, so the consumer's wait for it is short
//    and never abandoned (only an abort ends it);
//  * an item's completion decrements its unit's items_left; the last item of
//    a unit decrements the dependency count of each successor unit, queues
//    the items of every unit whose count reaches zero, and adds the unit's
//    items to `done`;
//  * workers leave when done == nitems, on an abort (device watchdog: no
//    progress for timeout_ticks or poll_limit empty polls; or the host-set
//    abort word), never otherwise;
//  * optional CU budget per recorded task (the reference's --sm and
//    FLAGS_OCCUP): at most task_cap workgroups run items of one task at once,
//    and a task starts only while the budgets of the running tasks sum to
//    less than 100 (percent of the GPU).
//
// Control flow.  Every branch of the worker loop that a barrier follows is
// wave-uniform: the scheduler steps run on all 64 lanes of wave 0 with every
// value broadcast from lane 0 (readfirstlane), and the atomic itself is the
// only statement under a lane test.  An earlier form ran the scheduler under
// `if (threadIdx.x == 0)`; the optimizer merged the end-of-item test with
// the next pop's test across the loop latch, and the structurizer turned the
// first barrier into the header of a divergent inner loop in which lanes
// 1-63 re-ran the previous item while lane 0 waited, masked off, to pop the
// next one -- a workgroup that never progressed and never reached its
// watchdog.
#pragma once
#include "mu_types.h"

#ifdef __HIP__  // HIP translation unit: callable from the kernel and the host
#include <hip/hip_runtime.h>
#define MU_HD __host__ __device__ __forceinline__
#else
#define MU_HD inline
#endif

namespace mu {

// ---- wave-collective operations: all lanes of the calling wave execute them
// in uniform control flow; lane 0 performs the memory operation and every
// lane receives its result ---------------------------------------------------
template <class AT>
MU_HD uint32_t w_load(uint32_t *p)
{
    return AT::bcast(AT::load(p));
}
template <class AT>
MU_HD uint32_t w_load_acquire(uint32_t *p)
{
    return AT::bcast(AT::load_acquire(p));
}
template <class AT>
MU_HD uint32_t w_add(uint32_t *p, uint32_t v)
{
    uint32_t r = 0;
    if (AT::lane0())
        r = AT::add(p, v);
    return AT::bcast(r);
}
template <class AT>
MU_HD uint32_t w_add_release(uint32_t *p, uint32_t v)
{
    uint32_t r = 0;
    if (AT::lane0())
        r = AT::add_release(p, v);
    return AT::bcast(r);
}
template <class AT>
MU_HD int32_t w_add_acq_rel(int32_t *p, int32_t v)
{
    int32_t r = 0;
    if (AT::lane0())
        r = AT::add_acq_rel(p, v);
    return (int32_t)AT::bcast((uint32_t)r);
}
template <class AT>
MU_HD uint32_t w_cas(uint32_t *p, uint32_t expected, uint32_t desired)
{
    uint32_t r = 0;
    if (AT::lane0())
        r = AT::cas(p, expected, desired);
    return AT::bcast(r);
}
template <class AT>
MU_HD void w_store(uint32_t *p, uint32_t v)
{
    if (AT::lane0())
        AT::store(p, v);
}

template <class AT>
MU_HD void sched_abort(const ExecView &ev, uint32_t code)
{
    w_store<AT>(&ev.st->err, code);
    w_store<AT>(&ev.st->abort, 1u);
}

// The run was aborted, by the device watchdog or by the host.
template <class AT>
MU_HD bool sched_stopped(const ExecView &ev)
{
    if (w_load<AT>(&ev.st->abort))
        return true;
    return ev.host_abort && AT::bcast(AT::load_system(ev.host_abort)) != 0;
}

// Queue all items of unit u on its priority ring (one thread).
template <class AT>
MU_HD void sched_push_unit(const ExecView &ev, uint32_t u)
{
    const UnitDesc *d = ev.units + u;
    const int p = d->prio;
    const uint32_t cnt = (uint32_t)d->nitems, base = d->item_base;
    const uint32_t pos = AT::add(&ev.st->tail[p], cnt);
    AT::producer_delay();
    // the stores below publish the unit's inputs to whoever loads them
    AT::fence_release();
    for (uint32_t r = 0; r < cnt; r++)
        AT::store(ev.slots[p] + pos + r, base + r + 1);
    AT::add_release(reinterpret_cast<uint32_t *>(&ev.st->cnt[p]), cnt);
}

// Next item for this workgroup (wave-collective), or -1 when the run is
// over or aborted.  Rings are tried in priority order.  The fast path (an
// item is queued) reads only the ring counters; the abort words -- one of
// them in host memory, a PCIe round trip away -- are read on every 16th call
// (`calls` counts them for the caller) and whenever nothing is queued.
template <class AT>
MU_HD int sched_pop(const ExecView &ev, uint32_t &calls)
{
    if ((calls++ & 15u) == 0 && sched_stopped<AT>(ev))
        return -1;
    uint64_t t0 = 0, idle = 0;
    uint32_t seen = 0;
    bool first = true;
    for (;;) {
        for (int p = 0; p < NPRIO; p++) {
            uint32_t *cnt = reinterpret_cast<uint32_t *>(&ev.st->cnt[p]);
            if ((int32_t)w_load<AT>(cnt) <= 0)
                continue;
            if ((int32_t)w_add<AT>(cnt, 0xffffffffu) <= 0) {
                w_add<AT>(cnt, 1u);  // none left after all: give the unit back
                continue;
            }
            const uint32_t pos = w_add<AT>(&ev.st->head[p], 1u);
            if (pos >= ev.cap[p]) {  // more claims than items: a broken DAG
                sched_abort<AT>(ev, 3u);
                return -1;
            }
            // the producer that reserved `pos` is running and storing it
            for (;;) {
                const uint32_t v = w_load_acquire<AT>(ev.slots[p] + pos);
                if (v)
                    return (int)(v - 1);
                if (sched_stopped<AT>(ev))
                    return -1;
                AT::pause();
            }
        }
        // nothing queued: stop, finish, or wait with the watchdog armed
        if (sched_stopped<AT>(ev))
            return -1;
        const uint32_t d = w_load<AT>(&ev.st->done);
        if (d >= ev.nitems)
            return -1;
        const uint64_t now = AT::now();
        if (first || d != seen) {
            first = false;
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

// CU budget (wave-collective): wait until the task of the popped item may
// run one more workgroup.  task_running[t] > 0 exactly while task t holds
// its share of the GPU in `occ`: the first running workgroup takes the share
// before it moves the count from 0 to 1 (and gives it back when another got
// there first), and the workgroup that moves it from 1 to 0 returns it.
// The wait ends when executing items finish, so it always ends; false when
// the run was aborted.
template <class AT>
MU_HD bool sched_admit(const ExecView &ev, int task)
{
    if (ev.task_cap <= 0 || task < 0)
        return true;
    uint32_t *run = reinterpret_cast<uint32_t *>(ev.task_running + task);
    uint32_t *occ = reinterpret_cast<uint32_t *>(&ev.st->occ);
    const uint32_t share = (uint32_t)ev.task_budget;
    for (;;) {
        const uint32_t r = w_load<AT>(run);
        if (r > 0 && r < (uint32_t)ev.task_cap) {
            if (w_cas<AT>(run, r, r + 1) == r)
                return true;
            continue;
        }
        if (r == 0) {
            const uint32_t o = w_load<AT>(occ);
            if ((int32_t)o < 100 && w_cas<AT>(occ, o, o + share) == o) {
                if (w_cas<AT>(run, 0u, 1u) == 0u)
                    return true;
                w_add<AT>(occ, (uint32_t)(-(int32_t)share));
            }
        }
        if (sched_stopped<AT>(ev))
            return false;
        AT::pause();
    }
}

template <class AT>
MU_HD void sched_retire(const ExecView &ev, int task)
{
    if (ev.task_cap <= 0 || task < 0)
        return;
    uint32_t *run = reinterpret_cast<uint32_t *>(ev.task_running + task);
    if (w_add<AT>(run, 0xffffffffu) == 1u)  // the task's last running workgroup
        w_add<AT>(reinterpret_cast<uint32_t *>(&ev.st->occ),
                  (uint32_t)(-(int32_t)ev.task_budget));
}

// Completion, step 1 (wave-collective): true when the item was the last of
// unit u.
template <class AT>
MU_HD bool sched_item_done(const ExecView &ev, uint32_t u)
{
    return w_add_acq_rel<AT>(ev.items_left + u, -1) == 1;
}

// Completion, step 2 (every thread: successors first, first + stride, ...):
// release the successors of the finished unit u.
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

// The worker loop of one workgroup (replaces kernel_scheduler,
// mustard/mustard.h:108-129; there one thread device-launches a sub-graph
// per task, here every workgroup runs items inline).  W is the per-wave
// context: wave() (uniform), barrier(), the two LDS words set_item/item and
// set_last/last, fence_release(), threads(f) calling f(first, stride) for
// each thread the caller stands for, and the item body run(d, item).
template <class AT, class W>
MU_HD void sched_worker(const ExecView &ev, W &w)
{
    uint32_t calls = 0;  // sched_pop calls of this workgroup (wave 0)
    if (w.wave() == 0) {
        int it = sched_pop<AT>(ev, calls);
        if (it >= 0 && !sched_admit<AT>(ev, ev.units[ev.item_unit[it]].task))
            it = -1;
        w.set_item(it);
    }
    w.barrier();
    for (;;) {
        const int item = w.item();
        if (item < 0)
            break;
        const uint32_t u = ev.item_unit[item];
        w.run(u, item);
        // every wave's writes are complete before the barrier; wave 0's
        // release (the acq_rel decrement below) then makes them visible
        // device-wide before the item counts as done
        w.fence_release();
        w.barrier();
        if (w.wave() == 0)
            w.set_last(sched_item_done<AT>(ev, u) ? 1 : 0);
        w.barrier();
        const bool last = w.last() != 0;
        if (last)
            w.threads([&](uint32_t first, uint32_t stride) {
                sched_release<AT>(ev, u, first, stride);
            });
        w.barrier();
        if (w.wave() == 0) {
            sched_retire<AT>(ev, ev.units[u].task);
            if (last)
                w_add_release<AT>(&ev.st->done, (uint32_t)ev.units[u].nitems);
            int it = sched_pop<AT>(ev, calls);
            if (it >= 0 && !sched_admit<AT>(ev, ev.units[ev.item_unit[it]].task))
                it = -1;
            w.set_item(it);
        }
        w.barrier();
    }
}

}  // namespace mu
