// mu_solve.cpp -- A X = B by a tiled factorization with reduced-precision
// trailing updates and FP64 iterative refinement.
//
// No counterpart in the reference (it factors only, lu_mustard.cu:547-660).
// The refinement follows LAPACK's dsgesv / dsposv: factor once in the lower
// precision, solve, then repeat r = b - A x (FP64), solve A d = r with the
// same factors, x = x + d, until ||r_j||_inf <= ||x_j||_inf ||A||_inf eps
// sqrt(n) for every column j (eps = 2^-53, LAPACK's dlamch('E')), or ITERMAX
// iterations; on failure LAPACK refactors in full precision, and so does this
// driver when asked to.  Here the factors are kept in FP64 storage: only the
// trailing GEMM/SYRK of the factorization run in FP32 or BF16 (the diagonal
// factorizations, the panel solves, the triangular solves of the refinement,
// the residual and the correction are FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/mustard_hip.h"
#include "mu_runtime.hpp"

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p)
            (void)hipFree(p);
    }
};

struct Dag {
    mu_dag_t h = nullptr;
    ~Dag()
    {
        if (h)
            mu_dag_destroy(h);
    }
};

struct Ex {
    mu_exec_t h = nullptr;
    ~Ex()
    {
        if (h)
            mu_exec_destroy(h);
    }
};

struct Ev {
    hipEvent_t e = nullptr;
    ~Ev()
    {
        if (e)
            (void)hipEventDestroy(e);
    }
};

}  // namespace

namespace mu {
int set_error(int code, const std::string &msg);  // mu_capi.cpp: the mu_last_error() text
}
static void mu_solve_ir_impl_set_error(const std::string &msg) { mu::set_error(MU_EINVAL, msg); }

#define CK(x)                                                                               \
    do {                                                                                    \
        int rc_ = (x);                                                                      \
        if (rc_ != 0)                                                                       \
            return rc_;                                                                     \
    } while (0)
#define HK(x)                                                                               \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            mu_solve_ir_impl_set_error(std::string(#x) + ": " + hipGetErrorString(e_));     \
            return MU_EHIP;                                                                 \
        }                                                                                   \
    } while (0)
#define KK(x)                                                                               \
    do {                                                                                    \
        if ((x) != 0) {                                                                     \
            mu_solve_ir_impl_set_error(#x " failed");                                       \
            return MU_EHIP;                                                                 \
        }                                                                                   \
    } while (0)

extern "C" int mu_solve_ir(const double *A, int n, int lda, const double *B, int ldb, double *X,
                           int ldx, int nrhs, const mu_ir_opts_t *opts, mu_ir_result_t *res,
                           void *stream)
{
    if (!A || !B || !X || !opts || !res || n <= 0 || nrhs <= 0 || lda < n || ldb < n || ldx < n ||
        opts->tiles <= 0 || n % opts->tiles != 0 ||
        (opts->kind != MU_SOLVE_LU && opts->kind != MU_SOLVE_CHOL) || opts->prec < MU_PREC_FP64 ||
        opts->prec > MU_PREC_BF16) {
        mu_solve_ir_impl_set_error("mu_solve_ir: bad arguments (n % tiles == 0, leading "
                                   "dimensions >= n, nrhs > 0, kind, prec)");
        return MU_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const int t = opts->tiles, itermax = opts->itermax > 0 ? opts->itermax : 30;
    *res = mu_ir_result_t{};

    // workspace: factors F (n x n), right-hand side / correction R (n x nrhs),
    // the info cell and the norms
    DevBuf dF, dR, dN, dI;
    HK(hipMalloc(&dF.p, (size_t)n * n * sizeof(double)));
    HK(hipMalloc(&dR.p, (size_t)n * nrhs * sizeof(double)));
    HK(hipMalloc(&dN.p, (size_t)(2 * nrhs + 1) * sizeof(double)));
    HK(hipMalloc(&dI.p, sizeof(int)));
    double *F = (double *)dF.p, *R = (double *)dR.p;
    double *rn = (double *)dN.p, *xn = rn + nrhs, *an = xn + nrhs;
    int *info = (int *)dI.p;


    // factorization and solve graphs over F and R
    auto build = [&](int prec, Dag &fd, Ex &fx, Dag &sd, Ex &sx) -> int {
        CK(mu_dag_create(&fd.h));
        if (opts->kind == MU_SOLVE_LU)
            CK(mu_record_tiled_lu(fd.h, F, n, t, n, info, prec));
        else
            CK(mu_record_tiled_cholesky(fd.h, F, n, t, n, info, prec));
        CK(mu_exec_create(fd.h, MU_MODE_TILED, nullptr, &fx.h));
        CK(mu_dag_create(&sd.h));
        CK(mu_record_tiled_solve(sd.h, F, n, t, n, opts->kind, R, n, nrhs));
        CK(mu_exec_create(sd.h, MU_MODE_TILED, nullptr, &sx.h));
        return 0;
    };
    auto factor = [&](Ex &fx) -> int {
        KK(mu::copy_matrix(A, lda, F, n, n, n, s));
        CK(mu_info_reset(info, s));
        CK(mu_exec_launch(fx.h, s));
        HK(hipMemcpyAsync(&res->info, info, sizeof(int), hipMemcpyDeviceToHost, s));
        HK(hipStreamSynchronize(s));
        if (res->info == INT32_MAX)
            res->info = 0;
        return 0;
    };
    // r = b - A x; returns per-column ||r||_inf, ||x||_inf on the host
    std::vector<double> hr(nrhs), hx(nrhs), hb(nrhs);
    auto residual = [&]() -> int {
        KK(mu::copy_matrix(B, ldb, R, n, n, nrhs, s));
        CK(mu_dgemm(stream, 'N', 'N', n, nrhs, n, -1.0, A, lda, X, ldx, 1.0, R, n));
        KK(mu::colmax_abs(R, n, n, nrhs, rn, s));
        KK(mu::colmax_abs(X, ldx, n, nrhs, xn, s));
        HK(hipMemcpyAsync(hr.data(), rn, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        HK(hipMemcpyAsync(hx.data(), xn, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        HK(hipStreamSynchronize(s));
        return 0;
    };

    // graphs are built (lowered, uploaded, instantiated) before the clock starts
    Dag fd, sd;
    Ex fx, sx;
    CK(build(opts->prec, fd, fx, sd, sx));
    Ev e0, e1, e2;
    HK(hipEventCreate(&e0.e));
    HK(hipEventCreate(&e1.e));
    HK(hipEventCreate(&e2.e));
    HK(hipEventRecord(e0.e, s));
    KK(mu::inf_norm(A, lda, n, n, an, s));
    KK(mu::colmax_abs(B, ldb, n, nrhs, rn, s));
    HK(hipMemcpyAsync(&res->anorm, an, sizeof(double), hipMemcpyDeviceToHost, s));
    HK(hipMemcpyAsync(hb.data(), rn, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
    HK(hipStreamSynchronize(s));
    const double eps = std::ldexp(1.0, -53);
    const double cte = res->anorm * eps * std::sqrt((double)n);
    auto berr = [&]() {
        double w = 0.0;
        for (int j = 0; j < nrhs; j++) {
            const double d = res->anorm * hx[j] + hb[j];
            w = std::max(w, d > 0.0 ? hr[j] / d : 0.0);
        }
        return w;
    };
    auto small = [&]() {
        for (int j = 0; j < nrhs; j++)
            if (!(hr[j] <= hx[j] * cte))
                return false;
        return true;
    };

    CK(factor(fx));
    HK(hipEventRecord(e1.e, s));
    if (res->info != 0) {
        mu_solve_ir_impl_set_error("mu_solve_ir: the factorization failed (info > 0)");
        return MU_ERUN;
    }
    // x0 = solve(b)
    KK(mu::copy_matrix(B, ldb, R, n, n, nrhs, s));
    CK(mu_exec_launch(sx.h, s));
    KK(mu::copy_matrix(R, n, X, ldx, n, nrhs, s));
    int it = 0;
    for (;; it++) {
        CK(residual());
        const double w = berr();
        if (it == 0)
            res->berr0 = w;
        res->berr = w;
        if (it < (int)(sizeof res->berr_hist / sizeof res->berr_hist[0]))
            res->berr_hist[it] = w;
        if (small()) {
            res->converged = 1;
            break;
        }
        if (it >= itermax || !std::isfinite(w))
            break;
        // d = solve(r) in place, x += d
        CK(mu_exec_launch(sx.h, s));
        KK(mu::add_matrix(X, ldx, R, n, n, nrhs, s));
    }
    res->iters = it;
    if (!res->converged && opts->fallback && opts->prec != MU_PREC_FP64) {
        // LAPACK's fallback: refactor in full precision and solve
        Dag fd2, sd2;
        Ex fx2, sx2;
        CK(build(MU_PREC_FP64, fd2, fx2, sd2, sx2));
        CK(factor(fx2));
        if (res->info != 0) {
            mu_solve_ir_impl_set_error("mu_solve_ir: the FP64 factorization failed (info > 0)");
            return MU_ERUN;
        }
        KK(mu::copy_matrix(B, ldb, R, n, n, nrhs, s));
        CK(mu_exec_launch(sx2.h, s));
        KK(mu::copy_matrix(R, n, X, ldx, n, nrhs, s));
        CK(residual());
        res->berr = berr();
        res->converged = small() ? 1 : 0;
        res->fell_back = 1;
    }
    HK(hipEventRecord(e2.e, s));
    HK(hipEventSynchronize(e2.e));
    float ms = 0.f;
    HK(hipEventElapsedTime(&ms, e0.e, e1.e));
    res->fact_ms = ms;
    HK(hipEventElapsedTime(&ms, e1.e, e2.e));
    res->solve_ms = ms;
    HK(hipEventElapsedTime(&ms, e0.e, e2.e));
    res->total_ms = ms;
    return MU_OK;
}
