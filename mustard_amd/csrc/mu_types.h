// mu_types.h -- descriptors shared by the host runtime and the gfx950 kernels.
//
// A recorded tile task (the unit the reference captures as one cuBLAS /
// cuSOLVER call, mustard/lu_mustard.cu:343-535) is lowered to one or more
// *units*.  A unit is a flat set of independent work items (one workgroup
// each); units form a DAG through dependency counters and successor lists.
// Large diagonal factorizations and large triangular solves lower to a nested
// DAG of units over sub-tiles, the way the reference nests a tiled LU for
// tiles above MAX_TILE (mustard/lu_mustard.cu:133-254, mustard/mustard.h:214-248).
#pragma once
#include <stdint.h>

namespace mu {

// Unit operations.
enum Op : int32_t {
    OP_NONE = 0,
    OP_GEMM = 1,   // C <- beta C + alpha op(A) op(B); items = bm x bn output blocks
    OP_SYRK = 2,   // lower(C) <- beta C + alpha A A^T; items = lower output blocks
    OP_TRSM = 3,   // panel triangular solve T Y = Z; items = panels of <= 16 right-hand sides
    OP_GETRF = 4,  // no-pivot LU of an n <= SMALL_MAX tile in LDS; one item
    OP_POTRF = 5,  // lower Cholesky of an n <= SMALL_MAX tile in LDS; one item
    OP_TRINV = 6,  // inverses of the 16 x 16 diagonal blocks of a TRSM triangle, into the
                   // unit's workspace slot; items = groups of TRINV_BLOCKS blocks
    OP_CVT = 7,    // FP64 block (c, ldc) rounded to FP32 / BF16 storage (b, ldb); items = 64 x 64
};

// Unit flags.
enum Flag : int32_t {
    F_TA = 1 << 0,       // GEMM: op(A) = A^T
    F_TB = 1 << 1,       // GEMM: op(B) = B^T
    F_VEC = 1 << 2,      // operands 16-byte aligned with even leading dims: vector loads
    F_T_TRANS = 1 << 3,  // TRSM: triangle read transposed, T[r][c] = t[c + r*ldt]
    F_Z_ROWS = 1 << 4,   // TRSM: right-hand sides are rows, Z[r][c] = z[c + r*ldz]
    F_UNIT = 1 << 5,     // TRSM: unit diagonal
    F_LOW_F32 = 1 << 6,  // GEMM/SYRK: operands rounded to FP32, FP32 MFMA, FP32 accumulation
    F_LOW_BF16 = 1 << 7, // GEMM/SYRK: operands rounded to BF16, FP32 accumulation
    F_REV = 1 << 8,      // TRSM: row order of T and Z reversed (upper-triangular solves)
    F_INV = 1 << 9,      // TRSM: diagonal blocks solved with the inverses a TRINV unit left in b
    F_SRC_LOW = 1 << 10, // GEMM/SYRK with F_LOW_*: a and b are stored in that precision (not FP64)
    F_NARROW = 1 << 11,  // GEMM (FP64, op(B) = B) with n <= 16: items are 16-row tiles, every
                         // column in one item
    F_CVT_T = 1 << 12,   // CVT: the block is stored transposed, element (r, c) at b[c + r*ldb]
    F_KVEC = 1 << 13,    // GEMM/SYRK with F_SRC_LOW, F_TA, no F_TB: both operands k-contiguous
                         // with 16-byte aligned rows and k a multiple of 16 bytes (the
                         // k_batch_gemm_low kernel)
    F_A32 = 1 << 14,     // GEMM (op(A) = A, op(B) = B, n <= 16): a stored as FP32, b and c
                         // FP64; FP32 MFMA on b rounded to FP32, sums added into c in FP64
                         // (lowered F_NARROW: items of 64 rows x one k-slice of
                         // NARROW32_KSLICE, gn = number of slices)
    F_BLK128 = 1 << 15,  // F_KVEC GEMM/SYRK of the level-batched graph: items are 128 x 128
                         // output blocks (gm, gn count 128-blocks)
    F_CVT_BOTH = 1 << 16, // CVT with F_CVT_T: also stores the block untransposed at (a, lda)
    F_BROW = 1 << 17,    // F_A32 GEMM: B and C stored by rows (B(k, j) at b[k*ldb + j], C(i, j)
                         // at c[i*ldc + j])
};

// Kernel geometry (fixed for gfx950: 4 waves of 64 lanes per workgroup).
constexpr int WG_THREADS = 256;
constexpr int GEMM_BM = 64;     // GEMM/SYRK output block rows per item
constexpr int GEMM_BN = 64;     // ... and columns
constexpr int GEMM_BK = 8;      // k-block staged in LDS per step
constexpr int TRSM_W = 16;        // right-hand sides per TRSM item
constexpr int TRSM_NMAX = 1024;   // largest triangle a TRSM item keeps in LDS
constexpr int TRSM_NEST = 256;    // default: triangles above this order nest into blocks,
                                  // whose solves run beside GEMM workgroups on a CU
constexpr int SMALL_MAX = 128;    // largest tile a GETRF/POTRF item factors in LDS
constexpr int NPRIO = 4;          // priority levels of the device queue
constexpr int TRINV_BLOCKS = 16;  // diagonal blocks one TRINV item inverts (16 threads each)
constexpr int NARROW32_KSLICE = 512;  // k per item of a split F_A32 GEMM (a longer k is split
                                      // over items whose partial sums the last one adds)
constexpr int NARROW32_SPLIT_BELOW = 128;  // F_A32 GEMMs of fewer 64-row blocks split k

struct UnitDesc {
    int32_t op;
    int32_t flags;
    int32_t m, n, k;       // operation sizes
    int32_t lda, ldb, ldc; // leading dimensions
    int32_t gm, gn;        // item grid (GEMM: blocks; TRSM: gn panels)
    int32_t nitems;
    int32_t prio;          // 0 = most urgent
    uint32_t item_base;    // global id of item 0
    uint32_t succ_begin;   // successors: succ[succ_begin .. succ_end)
    uint32_t succ_end;
    int32_t ndeps;         // initial dependency count
    int32_t info_base;     // global column offset reported by factor items
    int32_t task;          // recorded task this unit belongs to
    double alpha, beta;
    const double *a;
    const double *b;
    double *c;
    union {
        int32_t *info;     // LAPACK-style info cell (atomic min of failing column), or null
        float *part;       // F_A32 GEMM split over gn k-slices: gm x gn partial 64 x 16 tiles,
                           // then gm arrival counters (zero between launches)
    };
    int32_t ws_slot;       // TRSM with F_INV / TRINV: workspace slot of the inverted blocks;
                           // split F_A32 GEMM: of its partial tiles; or -1
    int32_t grp;           // GEMM: an item covers grp x grp output blocks (0 or 1: one block)
};
static_assert(sizeof(UnitDesc) == 128, "UnitDesc must stay 128 bytes");

// Per-run mutable scheduler state, all in device memory.
struct SchedState {
    int32_t cnt[NPRIO];   // queued, unclaimed items per priority ring
    uint32_t head[NPRIO]; // next ring position to hand out
    uint32_t tail[NPRIO]; // next ring position to reserve
    uint32_t done;        // items of completed units
    uint32_t abort;       // set on a time-out or a host abort; every workgroup exits
    uint32_t err;         // 1 = no progress for the time-out, 2 = poll limit, 3 = broken DAG,
                          // 4 = aborted by the host
    int32_t occ;          // CU budget of the running tasks, percent of the GPU (FLAGS_OCCUP)
    uint32_t pad_[4];
};
static_assert(sizeof(SchedState) == 80, "SchedState layout");

// Everything the device scheduler needs, passed by value to the kernels.
struct ExecView {
    const UnitDesc *units;
    const uint32_t *item_unit;     // item id -> unit id
    const uint32_t *succ;          // successor unit ids (CSR)
    int32_t *dep_count;            // per unit, reset each run
    int32_t *items_left;           // per unit, reset each run
    uint32_t *slots[NPRIO];        // queue rings: item id + 1, 0 = empty
    uint32_t cap[NPRIO];
    SchedState *st;
    uint32_t nunits;
    uint32_t nitems;
    uint64_t timeout_ticks;        // wall-clock ticks without progress before abort
    uint64_t poll_limit;           // empty polls without progress before abort
    uint64_t *trace;               // optional: per item (start, end) wall-clock stamps
    uint32_t *host_abort;          // host-written abort word (pinned host memory), or null
    int32_t *task_running;         // per recorded task: workgroups running its items
    int32_t task_cap;              // CU budget: workgroups per task (0 = no budget)
    int32_t task_budget;           // ... and the task's share of the GPU in percent
};

}  // namespace mu
