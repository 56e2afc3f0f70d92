// mu_sched_host.cpp -- host emulation of the persistent scheduler.
//
// Runs the worker loop of mu_sched.hpp (sched_worker, the same code the
// gfx950 kernel runs) on the CPU: one std::thread plays one wave, four of
// them one workgroup, and a workgroup barrier stands for s_barrier; the
// wave-collective steps run on wave 0's thread, the per-thread release step
// on every wave thread for its 64 lanes.  Items are not computed but checked:
// every item runs exactly once, only after every item of every predecessor
// unit has finished, and no task runs on more workgroups than its CU budget.
// Used by the CPU test suite (also under ThreadSanitizer) and by
// mu_dag_emulate.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mu_dag.hpp"
#include "mu_sched.hpp"
#include "mu_sched_host.hpp"

namespace mu {

namespace {

std::atomic<int> g_producer_delay_us{0};

struct HostAtom {
    static bool lane0() { return true; }
    static uint32_t bcast(uint32_t v) { return v; }
    static uint32_t load(uint32_t *p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
    static uint32_t load_acquire(uint32_t *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
    static uint32_t load_system(uint32_t *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
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
    // a producer stalled between reserving ring positions and storing them
    static void producer_delay()
    {
        const int us = g_producer_delay_us.load(std::memory_order_relaxed);
        if (us > 0)
            std::this_thread::sleep_for(std::chrono::microseconds(us));
    }
};

// A reusable barrier for the threads of one workgroup (s_barrier).
class WgBarrier {
  public:
    explicit WgBarrier(int n) : n_(n) {}
    void wait()
    {
        std::unique_lock<std::mutex> lk(m_);
        const uint64_t gen = gen_;
        if (++count_ == n_) {
            count_ = 0;
            gen_++;
            cv_.notify_all();
        } else {
            cv_.wait(lk, [&] { return gen_ != gen; });
        }
    }

  private:
    std::mutex m_;
    std::condition_variable cv_;
    int n_, count_ = 0;
    uint64_t gen_ = 0;
};

struct HostWg {
    explicit HostWg(int waves) : bar(waves) {}
    WgBarrier bar;
    int s_item = -1, s_last = 0;  // the workgroup's two LDS words
};

void raise_max(std::atomic<int32_t> &m, int32_t v)
{
    int32_t cur = m.load();
    while (v > cur && !m.compare_exchange_weak(cur, v)) {
    }
}

struct Check {
    const std::vector<std::vector<uint32_t>> *preds;
    int32_t *left;
    std::vector<std::atomic<uint32_t>> *runs;
    std::atomic<uint64_t> violations{0};
    std::vector<std::atomic<int32_t>> *task_now;  // workgroups running a task's items now
    std::atomic<int32_t> task_max{0};
    std::atomic<int32_t> tasks_active{0};  // tasks with a workgroup running one of their items
    std::atomic<int32_t> tasks_max{0};
    // one task's count reaching 0 and the active-task count following it are
    // one step: otherwise a workgroup of the same task starting in between
    // would count the task twice
    std::mutex task_mu;
    int body_barriers = 2;
};

struct HostWave {
    const ExecView &ev;
    HostWg &wg;
    Check &chk;
    int w;
    int wave() const { return w; }
    void barrier() { wg.bar.wait(); }
    void set_item(int v) { wg.s_item = v; }
    int item() const { return wg.s_item; }
    void set_last(int v) { wg.s_last = v; }
    int last() const { return wg.s_last; }
    void fence_release() { std::atomic_thread_fence(std::memory_order_release); }
    template <class F>
    void threads(F f)
    {
        for (uint32_t lane = 0; lane < 64; lane++)
            f(64u * (uint32_t)w + lane, (uint32_t)WG_THREADS);
    }
    void run(uint32_t u, int item)
    {
        const int task = ev.units[u].task;
        if (w == 0) {
            for (uint32_t p : (*chk.preds)[u])
                if (__atomic_load_n(&chk.left[p], __ATOMIC_ACQUIRE) != 0)
                    chk.violations++;
            (*chk.runs)[item]++;
            if (task >= 0 && task < (int)chk.task_now->size()) {
                std::lock_guard<std::mutex> g(chk.task_mu);
                const int32_t now = ++(*chk.task_now)[task];
                raise_max(chk.task_max, now);
                if (now == 1)
                    raise_max(chk.tasks_max, ++chk.tasks_active);
            }
        }
        // the item body's own barriers
        for (int b = 0; b < chk.body_barriers; b++)
            barrier();
        if (w == 0 && task >= 0 && task < (int)chk.task_now->size()) {
            std::lock_guard<std::mutex> g(chk.task_mu);
            if (--(*chk.task_now)[task] == 0)
                --chk.tasks_active;
        }
    }
};

}  // namespace

bool emulate(const Dag &dag, const LowerOpts &lo, const EmuOpts &eo, EmuResult &res,
             std::string &err)
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

    const uint32_t total = cap[0] + cap[1] + cap[2] + cap[3];
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
        const int p = L.units[u].prio;
        dep[u] = L.units[u].ndeps + ((int)u == eo.bump_unit ? 1 : 0);
        left[u] = L.units[u].nitems;
        if (dep[u] == 0)
            for (int i = 0; i < L.units[u].nitems; i++)
                slots[base[p] + st.tail[p]++] = L.units[u].item_base + i + 1;
    }
    for (int p = 0; p < NPRIO; p++)
        st.cnt[p] = (int32_t)st.tail[p];
    const size_t ntasks = std::max<size_t>(dag.tasks.size(), 1);
    std::vector<int32_t> task_running(ntasks, 0);
    uint32_t host_abort = 0;

    const int wgs = std::max(eo.workgroups, 1);
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
    ev.timeout_ticks = (uint64_t)(eo.timeout_s * 1e9);
    ev.poll_limit = ~0ull;
    ev.host_abort = &host_abort;
    ev.task_running = task_running.data();
    if (eo.sm > 0) {
        ev.task_budget = std::min(eo.sm, 100);
        ev.task_cap = std::max(1, (int)(((long long)ev.task_budget * wgs + 99) / 100));
    }

    std::vector<std::atomic<uint32_t>> runs(L.nitems);
    for (auto &r : runs)
        r.store(0);
    std::vector<std::atomic<int32_t>> task_now(ntasks);
    for (auto &t : task_now)
        t.store(0);
    Check chk;
    chk.preds = &preds;
    chk.left = left.data();
    chk.runs = &runs;
    chk.task_now = &task_now;

    g_producer_delay_us.store(eo.producer_delay_us);
    const int waves = 4;
    std::vector<HostWg *> wg;
    for (int g = 0; g < wgs; g++)
        wg.push_back(new HostWg(waves));
    std::vector<std::thread> pool;
    for (int g = 0; g < wgs; g++)
        for (int w = 0; w < waves; w++)
            pool.emplace_back([&, g, w]() {
                HostWave hw{ev, *wg[g], chk, w};
                sched_worker<HostAtom>(ev, hw);
            });
    // the host side of the watchdog: past the deadline, set the abort word
    std::atomic<bool> finished{false};
    std::thread dog([&]() {
        const auto t0 = std::chrono::steady_clock::now();
        while (!finished.load()) {
            const double el =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (eo.host_deadline_s > 0 && el > eo.host_deadline_s) {
                __atomic_store_n(&host_abort, 1u, __ATOMIC_RELEASE);
                return;
            }
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
    });
    for (auto &t : pool)
        t.join();
    finished.store(true);
    dog.join();
    g_producer_delay_us.store(0);
    for (HostWg *p : wg)
        delete p;
    res = EmuResult();
    res.nitems = (uint32_t)L.nitems;
    res.done = st.done;
    res.err = st.err ? st.err : (__atomic_load_n(&host_abort, __ATOMIC_ACQUIRE) ? 4u : 0u);
    res.order_violations = chk.violations.load();
    res.max_task_workgroups = (uint32_t)chk.task_max.load();
    res.task_cap = (uint32_t)ev.task_cap;
    res.max_tasks_running = (uint32_t)chk.tasks_max.load();
    for (auto &r : runs) {
        const uint32_t c = r.load();
        res.executed += c ? 1 : 0;
        res.duplicates += c > 1 ? c - 1 : 0;
    }
    return true;
}

}  // namespace mu
