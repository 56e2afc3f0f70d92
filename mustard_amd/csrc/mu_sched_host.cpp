// mu_sched_host.cpp -- host emulation of the persistent scheduler.
//
// Runs the protocol of mu_sched.hpp on the CPU: each std::thread plays one
// workgroup, items are not computed but checked (every item runs exactly
// once, and only after every item of every predecessor unit has finished).
// Used by the CPU test suite and by `mu_dag_emulate` for diagnosing DAGs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mu_dag.hpp"
#include "mu_sched.hpp"

namespace mu {

struct HostAtom {
    static uint32_t load(uint32_t *p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
    static uint32_t load_acquire(uint32_t *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
    static void store(uint32_t *p, uint32_t v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
    static uint32_t add(uint32_t *p, uint32_t v)
    {
        return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
    }
    static uint32_t add_release(uint32_t *p, uint32_t v)
    {
        return __atomic_fetch_add(p, v, __ATOMIC_RELEASE);
    }
    static int32_t add_acq_rel(int32_t *p, int32_t v)
    {
        return __atomic_fetch_add(p, v, __ATOMIC_ACQ_REL);
    }
    static uint32_t cas(uint32_t *p, uint32_t expected, uint32_t desired)
    {
        __atomic_compare_exchange_n(p, &expected, desired, false, __ATOMIC_RELAXED,
                                    __ATOMIC_RELAXED);
        return expected;
    }
    static void fence_release() { __atomic_thread_fence(__ATOMIC_RELEASE); }
    static uint64_t now()
    {
        return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                   std::chrono::steady_clock::now().time_since_epoch())
            .count();
    }
    static void pause() { std::this_thread::yield(); }
};

struct EmuResult {
    uint64_t executed = 0;
    uint64_t duplicates = 0;
    uint64_t order_violations = 0;
    uint32_t err = 0;
    uint32_t done = 0;
    uint32_t nitems = 0;
};

// `bump_unit` >= 0 adds one to that unit's initial dependency count (a
// dependency that is never satisfied), to exercise the progress watchdog.
bool emulate(const Dag &dag, const LowerOpts &lo, int threads, double timeout_s, int bump_unit,
             EmuResult &res, std::string &err)
{
    Lowered L;
    if (!lower(dag, lo, L, err))
        return false;
    const uint32_t nu = (uint32_t)L.units.size();
    std::vector<uint32_t> succ;
    uint32_t cap[NPRIO] = {0, 0, 0, 0};
    for (uint32_t u = 0; u < nu; u++) {
        L.units[u].succ_begin = (uint32_t)succ.size();
        succ.insert(succ.end(), L.succ[u].begin(), L.succ[u].end());
        L.units[u].succ_end = (uint32_t)succ.size();
        cap[L.units[u].prio] += (uint32_t)L.units[u].nitems;
    }
    std::vector<std::vector<uint32_t>> preds(nu);
    for (uint32_t u = 0; u < nu; u++)
        for (uint32_t v : L.succ[u])
            preds[v].push_back(u);
    std::vector<uint32_t> item_unit(L.nitems);
    for (uint32_t u = 0; u < nu; u++)
        for (int i = 0; i < L.units[u].nitems; i++)
            item_unit[L.units[u].item_base + i] = u;

    uint32_t total = cap[0] + cap[1] + cap[2] + cap[3];
    std::vector<uint32_t> slots(std::max<uint32_t>(total, 1), 0);
    std::vector<int32_t> dep(nu), left(nu);
    SchedState st;
    std::memset(&st, 0, sizeof st);
    uint32_t base[NPRIO];
    for (int p = 0, b = 0; p < NPRIO; p++) {
        base[p] = b;
        b += cap[p];
    }
    for (uint32_t u = 0; u < nu; u++) {
        dep[u] = L.units[u].ndeps + ((int)u == bump_unit ? 1 : 0);
        left[u] = L.units[u].nitems;
        if (dep[u] == 0)
            for (int i = 0; i < L.units[u].nitems; i++)
                slots[base[L.units[u].prio] + st.tail[L.units[u].prio]++] =
                    L.units[u].item_base + i + 1;
    }
    ExecView ev;
    std::memset(&ev, 0, sizeof ev);
    ev.units = L.units.data();
    ev.item_unit = item_unit.data();
    ev.succ = succ.data();
    ev.dep_count = dep.data();
    ev.items_left = left.data();
    for (int p = 0; p < NPRIO; p++) {
        ev.slots[p] = slots.data() + base[p];
        ev.cap[p] = cap[p];
    }
    ev.st = &st;
    ev.nunits = nu;
    ev.nitems = (uint32_t)L.nitems;
    ev.timeout_ticks = (uint64_t)(timeout_s * 1e9);
    ev.poll_limit = ~0ull;

    std::vector<std::atomic<uint32_t>> runs(L.nitems);
    for (auto &r : runs)
        r.store(0);
    std::atomic<uint64_t> violations{0};
    auto worker = [&]() {
        for (;;) {
            const int it = sched_pop<HostAtom>(ev);
            if (it < 0)
                return;
            const uint32_t u = item_unit[it];
            for (uint32_t p : preds[u])
                if (__atomic_load_n(&left[p], __ATOMIC_ACQUIRE) != 0)
                    violations++;
            runs[it]++;
            if (sched_item_done<HostAtom>(ev, u))
                sched_release<HostAtom>(ev, u, 0, 1);
            sched_count<HostAtom>(ev);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(threads, 1); t++)
        pool.emplace_back(worker);
    for (auto &t : pool)
        t.join();
    res = EmuResult();
    res.nitems = (uint32_t)L.nitems;
    res.done = st.done;
    res.err = st.err;
    res.order_violations = violations.load();
    for (auto &r : runs) {
        const uint32_t c = r.load();
        res.executed += c ? 1 : 0;
        res.duplicates += c > 1 ? c - 1 : 0;
    }
    return true;
}

}  // namespace mu
