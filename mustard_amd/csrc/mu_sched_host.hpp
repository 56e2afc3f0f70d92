// mu_sched_host.hpp -- host emulation of the persistent scheduler
// (mu_sched_host.cpp).
#pragma once
#include <cstdint>
#include <string>

#include "mu_dag.hpp"

namespace mu {

struct EmuOpts {
    int workgroups = 8;          // emulated workgroups (four threads each)
    double timeout_s = 20.0;     // device watchdog: no progress for this long
    double host_deadline_s = 0;  // host watchdog: abort after this long (0 = none)
    int bump_unit = -1;          // >= 0: one never-satisfied dependency on this unit
    int producer_delay_us = 0;   // stall of every producer between reserving and storing
    int sm = 0;                  // CU budget per task, percent (0 = none)
};

struct EmuResult {
    uint64_t executed = 0;
    uint64_t duplicates = 0;
    uint64_t order_violations = 0;
    uint32_t err = 0;
    uint32_t done = 0;
    uint32_t nitems = 0;
    uint32_t max_task_workgroups = 0;
    uint32_t task_cap = 0;
    uint32_t max_tasks_running = 0;  // most tasks with a running item at once
};

bool emulate(const Dag &dag, const LowerOpts &lo, const EmuOpts &eo, EmuResult &res,
             std::string &err);

}  // namespace mu
