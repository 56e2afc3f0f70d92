// mu_solve.cpp -- A X = B by a tiled factorization with reduced-precision
// trailing updates and FP64 iterative refinement.
//
// No counterpart in the reference (it factors only, lu_mustard.cu:547-660).
// The refinement follows LAPACK's dsgesv / dsposv: factor once in the lower
// precision, solve, then repeat r = b - A x (FP64), solve A d = r with the
// same factors, x = x + d, until ||r_j||_inf <= ||x_j||_inf ||A||_inf eps
// sqrt(n) for every column j (eps = 2^-53, LAPACK's dlamch('E')), or ITERMAX
// iterations; on failure LAPACK refactors in full precision, and so does this
// driver when asked to.  The factors are kept in FP64 storage: the trailing
// GEMM/SYRK of the factorization run in FP32 or BF16, the diagonal
// factorizations and the panel solves in FP64.  As in dsgesv / dsposv the
// refinement's triangular solves then run in FP32, on FP32 copies of the
// factors made inside the factorization's graph; the residual and the
// correction are FP64.  The triangular solves use the inverses of the
// diagonal blocks of two tiles (computed once per factorization), so each
// substitution step is a GEMM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "../../include/mustard_hip.h"
#include "mu_capi.hpp"
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

    // workspace: factors F (n x n), right-hand side / correction R and the
    // forward-substitution result Y (n x nrhs each), the inverses of the
    // diagonal blocks of two tiles (DL, and DU with LU; 2b x 2b each, see
    // below) with their b x b temporaries, the info cell and the norms
    const int b = n / t, npair = (t + 1) / 2, b2 = 2 * b;
    const int ninv = opts->kind == MU_SOLVE_LU ? 2 : 1;
    DevBuf dF, dR, dY, dD, dT, dN, dI, dP;
    HK(hipMalloc(&dF.p, (size_t)n * n * sizeof(double)));
    HK(hipMalloc(&dR.p, (size_t)n * nrhs * sizeof(double)));
    HK(hipMalloc(&dY.p, (size_t)n * nrhs * sizeof(double)));
    HK(hipMalloc(&dD.p, (size_t)ninv * npair * b2 * b2 * sizeof(double)));
    HK(hipMalloc(&dT.p, (size_t)ninv * npair * b * b * sizeof(double)));
    HK(hipMalloc(&dN.p, (size_t)(2 * nrhs + 1) * sizeof(double)));
    HK(hipMalloc(&dI.p, sizeof(int)));
    HK(hipMalloc(&dP.p, std::max(mu::residual_partial_doubles(n, n),
                                 mu::inf_norm_work_doubles(n, n)) * sizeof(double)));
    double *F = (double *)dF.p, *R = (double *)dR.p, *Y = (double *)dY.p;
    double *DL = (double *)dD.p, *DU = DL + (size_t)npair * b2 * b2;
    double *TL = (double *)dT.p, *TU = TL + (size_t)npair * b * b;
    double *rn = (double *)dN.p, *xn = rn + nrhs, *an = xn + nrhs;
    int *info = (int *)dI.p;
    // reduced precision: FP32 copies, for the refinement's substitutions, of
    // the factors (G, n x n: L below the diagonal blocks of two tiles, U --
    // or L^T with Cholesky -- above them) and of the two-tile inverses (DL32,
    // and DU32 = DU, or DL^T with Cholesky), the way dsgesv / dsposv solve
    // with their single-precision factors
    const size_t g_elems = ((size_t)n * n + 63) / 64 * 64, d_elems = (size_t)npair * b2 * b2;
    // Their solves keep the right-hand sides stored by rows (Rr, and Y):
    // a 16-column step then reads each row of B as one 128-byte line.
    DevBuf dG, dRr;
    float *G = nullptr, *DL32 = nullptr, *DU32 = nullptr;
    double *Rr = nullptr;
    if (opts->prec != MU_PREC_FP64) {
        HK(hipMalloc(&dG.p, (g_elems + 2 * d_elems) * sizeof(float)));
        G = (float *)dG.p;
        DL32 = G + g_elems;
        DU32 = DL32 + d_elems;
        HK(hipMalloc(&dRr.p, (size_t)n * nrhs * sizeof(double)));
        Rr = (double *)dRr.p;
    }

    // Graphs over F, R, Y: the factorization (with the inverses of its
    // diagonal blocks of two tiles in the same graph) and the solve.  With
    // the inverses every step of the substitutions is a GEMM over all of its
    // output blocks instead of a single-workgroup panel solve, and a step
    // covers two tiles, so each substitution is a chain of t/2 steps of two
    // GEMM levels.  Pair q holds tiles a = 2q, c = 2q + 1 (the last pair of
    // an odd t only tile a):
    //   DL_q = [La 0; Lca Lc]^-1 = [Da 0; X Dc],  Da = La^-1, Dc = Lc^-1,
    //          X = -Dc (Lca Da)
    //   DU_q = [Ua Uac; 0 Uc]^-1 = [Ea Z; 0 Ec],  Z = -Ea (Uac Ec)   (LU)
    //   forward   Y_q = DL_q R_q;        R_i -= L_{i,q} Y_q      (tiles i after q)
    //   backward  R_q = DU_q Y_q (LU)  or  DL_q^T Y_q (Cholesky);
    //             Y_i -= U_{i,q} R_q  or  L_{q,i}^T R_q          (tiles i before q)
    // leaving the solution in R.  DL_q and DU_q are 2b x 2b, leading
    // dimension 2b, reset to the identity before every factorization; the
    // diagonal-block inverses are TRSMs against that identity in place.
    auto fb = [&](int i, int j) { return F + (size_t)i * b + (size_t)j * b * n; };
    auto dpair = [&](double *D, int q, int r, int c) {  // block (r, c) of pair q
        return D + (size_t)q * b2 * b2 + (size_t)r * b + (size_t)c * b * b2;
    };
    auto tmp = [&](double *T, int q) { return T + (size_t)q * b * b; };
    auto gt = [&](int i, int j) { return G + (size_t)i * b + (size_t)j * b * n; };
    auto d32 = [&](float *D, int q) { return D + (size_t)q * b2 * b2; };
    auto width = [&](int q) { return 2 * q + 1 < t ? b2 : b; };  // rows of pair q
    auto build = [&](int prec, Dag &fd, Ex &fx, Dag &sd, Ex &sx) -> int {
        CK(mu_dag_create(&fd.h));
        if (opts->kind == MU_SOLVE_LU)
            CK(mu_record_tiled_lu(fd.h, F, n, t, n, info, prec));
        else
            CK(mu_record_tiled_cholesky(fd.h, F, n, t, n, info, prec));
        // the inverses join the factorization's DAG: each task reads only
        // the factored tiles it needs, so it runs beside the rest of the
        // factorization.  Workspace keys: (tile, t) / (tile, t + 1) for the
        // diagonal inverses of L / U, (q, t + 2) / (q, t + 3) for the
        // temporaries, (q, t + 4) / (q, t + 5) for X / Z.
        void *key = fd.h;
        const bool lu = opts->kind == MU_SOLVE_LU;
        CK(mu_dag_capture_begin(fd.h, key));
        for (int k = 0; k < t; k++) {
            const int q = k / 2, o = k % 2;
            const int rd[2] = {k, k};
            CK(mu_dag_begin_task(fd.h, k, t, rd, 1, nullptr));
            CK(mu_dtrsm(key, 'L', 'L', 'N', lu ? 'U' : 'N', b, b, 1.0, fb(k, k), n,
                        dpair(DL, q, o, o), b2));
            CK(mu_dag_end_task(fd.h));
            if (lu) {
                CK(mu_dag_begin_task(fd.h, k, t + 1, rd, 1, nullptr));
                CK(mu_dtrsm(key, 'L', 'U', 'N', 'N', b, b, 1.0, fb(k, k), n,
                            dpair(DU, q, o, o), b2));
                CK(mu_dag_end_task(fd.h));
            }
        }
        for (int q = 0; 2 * q + 1 < t; q++) {
            const int a = 2 * q, c = a + 1;
            // X = -Dc (Lca Da)
            const int r1[4] = {c, a, a, t};
            CK(mu_dag_begin_task(fd.h, q, t + 2, r1, 2, nullptr));
            CK(mu_dgemm(key, 'N', 'N', b, b, b, 1.0, fb(c, a), n, dpair(DL, q, 0, 0), b2, 0.0,
                        tmp(TL, q), b));
            CK(mu_dag_end_task(fd.h));
            const int r2[4] = {q, t + 2, c, t};
            CK(mu_dag_begin_task(fd.h, q, t + 4, r2, 2, nullptr));
            CK(mu_dgemm(key, 'N', 'N', b, b, b, -1.0, dpair(DL, q, 1, 1), b2, tmp(TL, q), b,
                        0.0, dpair(DL, q, 1, 0), b2));
            CK(mu_dag_end_task(fd.h));
            if (lu) {
                // Z = -Ea (Uac Ec)
                const int r3[4] = {a, c, c, t + 1};
                CK(mu_dag_begin_task(fd.h, q, t + 3, r3, 2, nullptr));
                CK(mu_dgemm(key, 'N', 'N', b, b, b, 1.0, fb(a, c), n, dpair(DU, q, 1, 1), b2,
                            0.0, tmp(TU, q), b));
                CK(mu_dag_end_task(fd.h));
                const int r4[4] = {q, t + 3, a, t + 1};
                CK(mu_dag_begin_task(fd.h, q, t + 5, r4, 2, nullptr));
                CK(mu_dgemm(key, 'N', 'N', b, b, b, -1.0, dpair(DU, q, 0, 0), b2, tmp(TU, q), b,
                            0.0, dpair(DU, q, 0, 1), b2));
                CK(mu_dag_end_task(fd.h));
            }
        }
        if (prec != MU_PREC_FP64) {
            // the FP32 copies, each as soon as its source is final.  Keys:
            // (i, 2t + 6 + j) for tile (i, j) of G, (q, 3t + 6) / (q, 3t + 7)
            // for DL32 / DU32 of pair q
            for (int j = 0; j < t; j++)
                for (int i = 0; i < t; i++) {
                    if (i / 2 == j / 2 || (!lu && i < j))
                        continue;
                    const int rd[2] = {i, j};
                    CK(mu_dag_begin_task(fd.h, i, 2 * t + 6 + j, rd, 1, nullptr));
                    if (lu)
                        CK(mu::cvt_f32(key, fb(i, j), n, gt(i, j), n, b, b, false));
                    else  // L, and L^T above the diagonal, from one read
                        CK(mu::cvt_f32(key, fb(i, j), n, gt(j, i), n, b, b, true, gt(i, j), n));
                    CK(mu_dag_end_task(fd.h));
                }
            for (int q = 0; q < npair; q++) {
                const int a = 2 * q, c = a + 1, nr = c < t ? 3 : 1;
                const int rl[6] = {a, t, q, t + 4, c, t};
                CK(mu_dag_begin_task(fd.h, q, 3 * t + 6, rl, nr, nullptr));
                if (lu)
                    CK(mu::cvt_f32(key, dpair(DL, q, 0, 0), b2, d32(DL32, q), b2, b2, b2, false));
                else
                    CK(mu::cvt_f32(key, dpair(DL, q, 0, 0), b2, d32(DU32, q), b2, b2, b2, true,
                                   d32(DL32, q), b2));
                CK(mu_dag_end_task(fd.h));
                if (lu) {
                    const int ru[6] = {a, t + 1, q, t + 5, c, t + 1};
                    CK(mu_dag_begin_task(fd.h, q, 3 * t + 7, ru, nr, nullptr));
                    CK(mu::cvt_f32(key, dpair(DU, q, 0, 0), b2, d32(DU32, q), b2, b2, b2,
                                   false));
                    CK(mu_dag_end_task(fd.h));
                }
            }
        }
        CK(mu_dag_capture_end(fd.h, key));
        CK(mu_exec_create(fd.h, MU_MODE_TILED, nullptr, &fx.h));
        // the solve: tile k of R is (k, 0), of Y is (k, 1); a pair's step
        // writes (a, 0) / (a, 1) for both its tiles
        CK(mu_dag_create(&sd.h));
        key = sd.h;
        CK(mu_dag_capture_begin(sd.h, key));
        auto rb = [&](double *M, int k) { return M + (size_t)k * b; };
        if (prec != MU_PREC_FP64) {
            // on the FP32 copies, in groups of <= 16 right-hand sides.  The
            // update of a step is one GEMM over every tile after (forward) or
            // before (backward) the pair: written as key (a', 0) for the
            // first tile a' after the pair, and (0, 1) for the rows above it
            // (every backward step reads (0, 1) as well)
            // right-hand sides by rows: row i of Rr / Y at i * nrhs
            auto rr = [&](double *M, int k) { return M + (size_t)k * b * nrhs; };
            auto rhs = [&](auto &&op) -> int {
                for (int j0 = 0; j0 < nrhs; j0 += 16)
                    CK(op((size_t)j0, std::min(16, nrhs - j0)));
                return 0;
            };
            for (int q = 0; q < npair; q++) {
                const int a = 2 * q, w = width(q), last = a + w / b - 1;
                const int rd0[4] = {a, 0, last, 0};
                CK(mu_dag_begin_task(sd.h, a, 1, rd0, 2, nullptr));
                CK(rhs([&](size_t o, int nj) {
                    return mu::gemm_a32(key, w, nj, w, 1.0, d32(DL32, q), b2, rr(Rr, a) + o, nrhs,
                                        0.0, rr(Y, a) + o, nrhs, true);
                }));
                CK(mu_dag_end_task(sd.h));
                if (last + 1 < t) {
                    const int rd[4] = {a, 1, last + 1, 0};
                    CK(mu_dag_begin_task(sd.h, last + 1, 0, rd, 2, nullptr));
                    CK(rhs([&](size_t o, int nj) {
                        return mu::gemm_a32(key, (t - last - 1) * b, nj, w, -1.0,
                                            gt(last + 1, a), n, rr(Y, a) + o, nrhs, 1.0,
                                            rr(Rr, last + 1) + o, nrhs, true);
                    }));
                    CK(mu_dag_end_task(sd.h));
                }
            }
            for (int q = npair - 1; q >= 0; q--) {
                const int a = 2 * q, w = width(q), last = a + w / b - 1;
                const int rd0[6] = {a, 1, last, 1, 0, 1};
                CK(mu_dag_begin_task(sd.h, a, 0, rd0, 3, nullptr));
                CK(rhs([&](size_t o, int nj) {
                    return mu::gemm_a32(key, w, nj, w, 1.0, d32(DU32, q), b2, rr(Y, a) + o, nrhs,
                                        0.0, rr(Rr, a) + o, nrhs, true);
                }));
                CK(mu_dag_end_task(sd.h));
                if (a > 0) {
                    const int rd[4] = {a, 0, 0, 1};
                    CK(mu_dag_begin_task(sd.h, 0, 1, rd, 2, nullptr));
                    CK(rhs([&](size_t o, int nj) {
                        return mu::gemm_a32(key, a * b, nj, w, -1.0, gt(0, a), n, rr(Rr, a) + o,
                                            nrhs, 1.0, rr(Y, 0) + o, nrhs, true);
                    }));
                    CK(mu_dag_end_task(sd.h));
                }
            }
            CK(mu_dag_capture_end(sd.h, key));
            CK(mu_exec_create(sd.h, MU_MODE_TILED, nullptr, &sx.h));
            return 0;
        }
        for (int q = 0; q < npair; q++) {
            const int a = 2 * q, w = width(q), last = a + w / b - 1;
            const int rd0[4] = {a, 0, last, 0};
            CK(mu_dag_begin_task(sd.h, a, 1, rd0, 2, nullptr));
            CK(mu_dgemm(key, 'N', 'N', w, nrhs, w, 1.0, dpair(DL, q, 0, 0), b2, rb(R, a), n, 0.0,
                        rb(Y, a), n));
            CK(mu_dag_end_task(sd.h));
            for (int i = last + 1; i < t; i++) {
                int rd[4] = {a, 1, i, 0};
                CK(mu_dag_begin_task(sd.h, i, 0, rd, 2, nullptr));
                CK(mu_dgemm(key, 'N', 'N', b, nrhs, w, -1.0, fb(i, a), n, rb(Y, a), n, 1.0,
                            rb(R, i), n));
                CK(mu_dag_end_task(sd.h));
            }
        }
        const bool lu_ = opts->kind == MU_SOLVE_LU;
        for (int q = npair - 1; q >= 0; q--) {
            const int a = 2 * q, w = width(q), last = a + w / b - 1;
            const int rd0[4] = {a, 1, last, 1};
            CK(mu_dag_begin_task(sd.h, a, 0, rd0, 2, nullptr));
            CK(mu_dgemm(key, lu_ ? 'N' : 'T', 'N', w, nrhs, w, 1.0,
                        lu_ ? dpair(DU, q, 0, 0) : dpair(DL, q, 0, 0), b2, rb(Y, a), n, 0.0,
                        rb(R, a), n));
            CK(mu_dag_end_task(sd.h));
            for (int i = 0; i < a; i++) {
                int rd[4] = {a, 0, i, 1};
                CK(mu_dag_begin_task(sd.h, i, 1, rd, 2, nullptr));
                CK(mu_dgemm(key, lu_ ? 'N' : 'T', 'N', b, nrhs, w, -1.0,
                            lu_ ? fb(i, a) : fb(a, i), n, rb(R, a), n, 1.0, rb(Y, i), n));
                CK(mu_dag_end_task(sd.h));
            }
        }
        CK(mu_dag_capture_end(sd.h, key));
        CK(mu_exec_create(sd.h, MU_MODE_TILED, nullptr, &sx.h));
        return 0;
    };
    // F = A (unless the norm's pass already copied it), then factor
    auto factor = [&](Ex &fx, bool copied) -> int {
        if (!copied)
            KK(mu::copy_matrix(A, lda, F, n, n, n, s));
        CK(mu_info_reset(info, s));
        KK(mu::identity_tiles(DL, b2, npair * ninv, s));
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
        KK(mu::residual(A, lda, X, ldx, B, ldb, R, n, n, n, nrhs, (double *)dP.p, s));
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
    KK(mu::copy_inf_norm(A, lda, F, n, n, n, an, (double *)dP.p, s));
    KK(mu::colmax_abs(B, ldb, n, nrhs, rn, s));
    HK(hipMemcpyAsync(&res->anorm, an, sizeof(double), hipMemcpyDeviceToHost, s));
    HK(hipMemcpyAsync(hb.data(), rn, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
    HK(hipStreamSynchronize(s));
    const double eps = std::ldexp(1.0, -53);
    const double cte = res->anorm * eps * std::sqrt((double)n);
    // a non-finite norm makes the backward error NaN (std::max would drop it)
    auto berr = [&]() {
        double w = 0.0;
        for (int j = 0; j < nrhs; j++) {
            const double d = res->anorm * hx[j] + hb[j];
            if (!std::isfinite(hr[j]) || !std::isfinite(d))
                return std::numeric_limits<double>::quiet_NaN();
            w = std::max(w, d > 0.0 ? hr[j] / d : 0.0);
        }
        return w;
    };
    // the stopping test fails on any non-finite norm (the column maxima carry
    // NaN through), so a NaN or overflow never counts as convergence
    auto small = [&]() {
        if (!std::isfinite(cte))
            return false;
        for (int j = 0; j < nrhs; j++)
            if (!std::isfinite(hr[j]) || !std::isfinite(hx[j]) || !(hr[j] <= hx[j] * cte))
                return false;
        return true;
    };

    // LAPACK dsgesv/dsposv: a matrix with an entry beyond the single-
    // precision range (here ||A||_inf > FLT_MAX bounds every entry) is not
    // rounded to FP32 / BF16 at all; it goes straight to the FP64 path
    const bool out_of_range = opts->prec != MU_PREC_FP64 && !(res->anorm <= FLT_MAX);
    if (out_of_range && !opts->fallback) {
        mu_solve_ir_impl_set_error(
            "mu_solve_ir: ||A||_inf exceeds the single-precision range; use FP64 or fallback");
        return MU_EINVAL;
    }
    bool low_failed = false;
    if (!out_of_range) {
        CK(factor(fx, true));
        // dsgesv/dsposv (ITER = -3): a reduced-precision factorization that
        // breaks down (a zero pivot, or a Cholesky pivot the rounded updates
        // drove non-positive) goes to the FP64 path when a fallback is allowed
        low_failed = res->info != 0 && opts->prec != MU_PREC_FP64 && opts->fallback;
        if (res->info != 0 && !low_failed) {
            mu_solve_ir_impl_set_error("mu_solve_ir: the factorization failed (info > 0)");
            return MU_ERUN;
        }
    }
    HK(hipEventRecord(e1.e, s));
    if (out_of_range || low_failed) {
        res->iters = 0;
    } else {
        // the solve of R in place (through Rr for the reduced-precision graph)
        const bool rows = opts->prec != MU_PREC_FP64;
        auto solve = [&]() -> int {
            if (rows)
                KK(mu::to_rows(R, n, Rr, n, nrhs, s));
            CK(mu_exec_launch(sx.h, s));
            if (rows)
                KK(mu::from_rows(Rr, R, n, n, nrhs, s));
            return 0;
        };
        // x0 = solve(b)
        KK(mu::copy_matrix(B, ldb, R, n, n, nrhs, s));
        CK(solve());
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
            CK(solve());
            KK(mu::add_matrix(X, ldx, R, n, n, nrhs, s));
        }
        res->iters = it;
    }
    if (!res->converged && opts->fallback && opts->prec != MU_PREC_FP64) {
        // LAPACK's fallback: refactor in full precision and solve
        Dag fd2, sd2;
        Ex fx2, sx2;
        CK(build(MU_PREC_FP64, fd2, fx2, sd2, sx2));
        CK(factor(fx2, false));
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
