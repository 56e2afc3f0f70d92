// mu_kern_generic.hip -- the kernels that run any item kind: the per-unit
// kernel of the stream mode (k_unit), the generic level kernel of the tiled
// mode (k_batch) and the 8-wave panel-solve level kernel for triangles of
// order up to 1024 (k_batch_trsm).  Their LDS is sized for the largest panel
// solve, so one workgroup fits on a CU; they live in their own translation
// unit because they instantiate every item body.
#include "mu_kernels.hpp"
#include "mu_ops.hpp"

namespace mu {

__global__ __launch_bounds__(WG_THREADS) void k_unit(UnitDesc d)
{
    __shared__ double smem[LDS_DOUBLES];
    run_item(d, blockIdx.x, smem);
}

// One launch runs every item of one graph node; block b runs items[b].
__global__ __launch_bounds__(WG_THREADS) void k_batch(const UnitDesc *units,
                                                      const uint32_t *item_unit,
                                                      const uint32_t *items)
{
    __shared__ double smem[LDS_DOUBLES];
    const uint32_t item = items[blockIdx.x];
    // a private copy of the descriptor lives in registers; a reference into
    // global memory would be re-read after every barrier of the item body
    const UnitDesc d = units[item_unit[item]];
    run_item(d, (int)(item - d.item_base), smem);
}

__global__ __launch_bounds__(64 * TRSM_KERNEL_WAVES) void k_batch_trsm(const UnitDesc *units,
                                                                      const uint32_t *item_unit,
                                                                      const uint32_t *items)
{
    __shared__ double smem[TRSM_NMAX * 16];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_trsm_inv_item<TRSM_KERNEL_WAVES>(d, (int)(item - d.item_base), smem);
}

const void *generic_kernel(int kind)
{
    if (kind == LEVEL_TRSM)
        return reinterpret_cast<const void *>(k_batch_trsm);
    return reinterpret_cast<const void *>(k_batch);
}

int launch_unit_generic(const UnitDesc &d, hipStream_t s)
{
    hipLaunchKernelGGL(k_unit, dim3(d.nitems), dim3(WG_THREADS), 0, s, d);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mu
