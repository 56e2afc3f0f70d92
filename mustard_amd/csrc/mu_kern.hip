// mu_kern.hip -- the specialised level kernels of the tiled mode (one launch
// runs every item of one graph node; block b runs items[b]) and the per-unit
// GEMM kernel of the stream mode.  Each needs only its own item kind's LDS
// and registers, so several workgroups share a CU: the bulk GEMM/SYRK blocks
// at four per CU, and the panel work of the next step beside them.
#include "mu_kernels.hpp"
#include "mu_ops.hpp"

namespace mu {

#define MU_GEMM_KERNEL __launch_bounds__(WG_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4)))

__global__ MU_GEMM_KERNEL void k_unit_gemm(UnitDesc d)
{
    __shared__ double smem[LDS_GEMM];
    run_gemm_item(d, blockIdx.x, smem);
}

__global__ MU_GEMM_KERNEL void k_batch_gemm(const UnitDesc *units, const uint32_t *item_unit,
                                            const uint32_t *items)
{
    __shared__ double smem[LDS_GEMM];
    const uint32_t item = items[blockIdx.x];
    // a private copy of the descriptor lives in registers; a reference into
    // global memory would be re-read after every barrier of the item body
    const UnitDesc d = units[item_unit[item]];
    run_gemm_item(d, (int)(item - d.item_base), smem);
}

// Diagonal GETRF/POTRF (and precision conversions), 4 waves held to 168
// registers per lane: a CU running three GEMM workgroups (104 registers per
// wave) still has room for it, so the diagonal factorization of the next step
// starts as soon as one bulk GEMM workgroup leaves a CU instead of waiting for
// the bulk update to drain.
__global__ __launch_bounds__(WG_THREADS) __attribute__((amdgpu_waves_per_eu(3, 3))) void
k_batch_small(const UnitDesc *units, const uint32_t *item_unit, const uint32_t *items)
{
    __shared__ double smem[LDS_SMALL];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_small_item<4>(d, (int)(item - d.item_base), smem);
}

// Reduced-precision GEMM/SYRK units whose operands are k-contiguous in
// their storage precision (the trailing updates of an FP32 / BF16
// factorization reading the converted panels): long k-slices; 128 x 128
// blocks in the level-batched graph (72 KB of LDS, two workgroups per CU).
__global__ __launch_bounds__(WG_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void
k_batch_gemm_low(const UnitDesc *units, const uint32_t *item_unit, const uint32_t *items)
{
    __shared__ double smem[LDS_GEMM_LOWK128];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_gemm_lowk_item(d, (int)(item - d.item_base), smem);
}

// Precision conversions (the panels' reduced-precision shadows, the FP32
// copies of the factors): 16 KB of LDS and few registers, so eight
// workgroups share a CU -- a conversion on the lookahead chain waits for no
// CU to drain, and the bulk ones stream at full occupancy.
__global__ __launch_bounds__(WG_THREADS) void k_batch_cvt(const UnitDesc *units,
                                                          const uint32_t *item_unit,
                                                          const uint32_t *items)
{
    __shared__ double smem[4 * 16 * 65 / 2 + 4];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    cvt_block(d, (int)(item - d.item_base), smem);
}

// GEMMs with at most 16 output columns (substitution steps of a solve).
__global__ __launch_bounds__(WG_THREADS) void k_batch_gemm_narrow(const UnitDesc *units,
                                                                  const uint32_t *item_unit,
                                                                  const uint32_t *items)
{
    __shared__ double smem[LDS_GEMM];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_gemm_narrow_item(d, (int)(item - d.item_base), smem);
}

// Panel solves of order <= SMALL_MAX (inside nested diagonal factorizations).
__global__ __launch_bounds__(WG_THREADS) void k_batch_trsm_small(const UnitDesc *units,
                                                                 const uint32_t *item_unit,
                                                                 const uint32_t *items)
{
    __shared__ double smem[LDS_TRSM_SMALL];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    run_trsm_small_item(d, (int)(item - d.item_base), smem);
}

// Panel solves with inverted diagonal blocks of order <= NMAX, and the TRINV
// items before them, on 4 waves and at most NMAX * 128 bytes of LDS
// (k_batch_trsm_256: 34 KB, k_batch_trsm_512: 64 KB), so GEMM workgroups keep
// running beside them on the same CU.
template <int NMAX>
__global__ __launch_bounds__(WG_THREADS) void k_batch_trsm_mid(const UnitDesc *units,
                                                               const uint32_t *item_unit,
                                                               const uint32_t *items)
{
    constexpr int LDSD = NMAX * 16 > TRINV_BLOCKS * 16 * 17 ? NMAX * 16 : TRINV_BLOCKS * 16 * 17;
    __shared__ double smem[LDSD];
    const uint32_t item = items[blockIdx.x];
    const UnitDesc d = units[item_unit[item]];
    if (d.op == OP_TRINV)
        run_trinv_item(d, (int)(item - d.item_base), smem);
    else
        run_trsm_inv_item<4, NMAX>(d, (int)(item - d.item_base), smem);
}

const void *level_kernel(int kind)
{
    switch (kind) {
    case LEVEL_GEMM: return reinterpret_cast<const void *>(k_batch_gemm);
    case LEVEL_SMALL: return reinterpret_cast<const void *>(k_batch_small);
    case LEVEL_TRSM_SMALL: return reinterpret_cast<const void *>(k_batch_trsm_small);
    case LEVEL_GEMM_NARROW: return reinterpret_cast<const void *>(k_batch_gemm_narrow);
    case LEVEL_GEMM_LOW: return reinterpret_cast<const void *>(k_batch_gemm_low);
    case LEVEL_CVT: return reinterpret_cast<const void *>(k_batch_cvt);
    case LEVEL_TRSM_256: return reinterpret_cast<const void *>(k_batch_trsm_mid<256>);
    case LEVEL_TRSM_512: return reinterpret_cast<const void *>(k_batch_trsm_mid<512>);
    default: return generic_kernel(kind);
    }
}

int level_threads(int kind)
{
    return kind == LEVEL_TRSM ? 64 * TRSM_KERNEL_WAVES : WG_THREADS;
}

int launch_unit(const UnitDesc &d, hipStream_t s)
{
    if ((d.op == OP_GEMM || d.op == OP_SYRK) && !(d.flags & F_NARROW)) {
        hipLaunchKernelGGL(k_unit_gemm, dim3(d.nitems), dim3(WG_THREADS), 0, s, d);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    return launch_unit_generic(d, s);
}

}  // namespace mu
