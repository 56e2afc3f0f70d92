// mu_noalloc.hip -- mu_dtrsm_noalloc: mu_dtrsm for callers that cannot
// allocate device memory.
//
// A triangle of order above SMALL_MAX is solved with its 16 x 16 diagonal
// blocks inverted first (a TRINV unit, then the panel solve reading the
// inverses); mu_dtrsm takes the inverses' workspace from a stream-ordered
// allocation per call.  This variant keeps them in `scratch`, a fixed device
// array of MU_NOALLOC_SCRATCH_BYTES that each GPU holds until the process
// ends, and otherwise runs the same lowering and the same kernels, so its
// result is bit-identical to mu_dtrsm's.  Every 16 x 16 inverse the solve
// reads is written in full by the TRINV items before it, so nothing in the
// array needs clearing between calls.  Calls that overlap on one GPU (two
// streams, or two host threads) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/mustard_hip.h"
#include "mu_capi.hpp"
#include "mu_dag.hpp"
#include "mu_runtime.hpp"

namespace {

constexpr size_t SCRATCH_DOUBLES = MU_NOALLOC_SCRATCH_BYTES / sizeof(double);
static_assert(MU_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch is a whole number of 256 B lines");

// Host code never uses this name as an address: the device address of each
// GPU's copy comes from hipGetSymbolAddress (scratch_on_current_device).
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_scratch[MAX_DEVICES];  // per device, once looked up

// The current device's copy of `scratch`, looked up once per device.
int scratch_on_current_device(double **ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return mu::set_error(MU_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES)
        return mu::set_error(MU_ENOTSUP, "mu_dtrsm_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return mu::set_error(MU_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                              hipGetErrorString(e));
        g_scratch[dev] = (double *)p;
    }
    *ptr = g_scratch[dev];
    return MU_OK;
}

}  // namespace

extern "C" int mu_dtrsm_noalloc(void *stream, char side, char uplo, char transa, char diag,
                                int m, int n, double alpha, const double *A, int lda, double *B,
                                int ldb)
{
    mu::OpRec op;
    int rc = mu::dtrsm_op(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, op);
    if (rc != MU_OK)
        return rc;
    if (mu::stream_recording(stream))
        return mu::set_error(MU_ESTATE, "mu_dtrsm_noalloc: the stream is being recorded into a "
                                        "DAG (record with mu_dtrsm; the executable owns its "
                                        "workspace)");
    if (op.m > MU_NOALLOC_MAX_ORDER)
        return mu::set_error(MU_EINVAL, "mu_dtrsm_noalloc: a triangle of order " +
                                            std::to_string(op.m) + " is above the " +
                                            std::to_string(MU_NOALLOC_MAX_ORDER) +
                                            " the scratch array is sized for; use mu_dtrsm");
    mu::Dag d;
    d.begin_task({0, 0}, {});
    d.add_op(op);
    d.end_task();
    mu::Lowered low;
    std::string err;
    if (!mu::lower(d, mu::LowerOpts(), low, err))
        return mu::set_error(MU_EINVAL, err);
    const size_t need = mu::ws_total_doubles(low);
    if (need > SCRATCH_DOUBLES)
        return mu::set_error(
            MU_EINVAL, "mu_dtrsm_noalloc: the inverted diagonal blocks of a triangle of order " +
                           std::to_string(op.m) + " need " + std::to_string(need * 8) +
                           " bytes of scratch, above the " +
                           std::to_string(MU_NOALLOC_SCRATCH_BYTES) +
                           " it holds (order <= 8192 fits); use mu_dtrsm");
    if (need) {
        double *ws = nullptr;
        if ((rc = scratch_on_current_device(&ws)) != MU_OK)
            return rc;
        mu::patch_workspace(low.units, low, ws);
    }
    if (mu::run_units(low, (hipStream_t)stream, err) != 0)
        return mu::set_error(MU_EHIP, err);
    return MU_OK;
}
