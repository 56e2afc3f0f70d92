// mu_kernels.hpp -- what the host runtime needs from the kernel translation
// units (mu_kern.hip, mu_kern_generic.hip, mu_persist.hip): kernel handles
// for graph nodes, launchers and the persistent scheduler's reset.
#pragma once
#include <hip/hip_runtime.h>

#include "mu_types.h"

namespace mu {

// The kernel a graph node's items run on.
enum LevelKernel : char {
    LEVEL_GENERIC = 0,
    LEVEL_GEMM = 1,
    LEVEL_SMALL = 2,
    LEVEL_TRSM = 3,
    LEVEL_TRSM_SMALL = 4,
    LEVEL_TRSM_256 = 5,
    LEVEL_TRSM_512 = 6,
    LEVEL_GEMM_NARROW = 7,
    LEVEL_GEMM_LOW = 8,  // reduced-precision GEMM/SYRK with k-contiguous operands (F_KVEC)
    LEVEL_CVT = 9,       // precision conversions (OP_CVT)
    LEVEL_KINDS = 10,
};

constexpr int TRSM_KERNEL_WAVES = 8;  // k_batch_trsm: two waves per SIMD

// Largest panel-solve triangle with inverted diagonal blocks the persistent
// kernel runs (its LDS holds the panel and the TRINV staging, 34 KB); larger
// solves are nested by the lowering.
constexpr int PERSIST_TRSM_MAX = 256;

// Kernel handle and block size of a graph node of kind `kind` (mu_kern.hip).
const void *level_kernel(int kind);
int level_threads(int kind);
// The generic kernels (mu_kern_generic.hip).
const void *generic_kernel(int kind);
// One unit's items as one launch (stream mode).
int launch_unit(const UnitDesc &d, hipStream_t s);
int launch_unit_generic(const UnitDesc &d, hipStream_t s);

// Persistent scheduler (mu_persist.hip).
struct ResetArgs {
    int32_t init_cnt[NPRIO];   // items queued before the run, per ring
    const uint32_t *slots_init;
    uint32_t *slots_flat;
    uint32_t slots_total;
    uint32_t ntasks;           // entries of task_running
};
// Workgroups of the persistent kernel one CU holds.
int persist_blocks_per_cu(int *per_cu);
// Reset the run state and launch `grid` persistent workgroups on s.
int persist_launch(const ExecView &ev, const ResetArgs &ra, int grid, hipStream_t s);

}  // namespace mu
