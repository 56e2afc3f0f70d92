// mu_adjoint.cpp -- reverse mode of the tiled Cholesky factorization.
//
// No counterpart in the reference (it has no derivatives).  Given the factor
// L of A = L L^T and the gradient Lbar of a scalar loss with respect to (the
// lower triangle of) L, the gradient with respect to the symmetric A is
//     G = (Abar + Abar^T) / 2,   Abar = L^-T Phi(L^T Lbar) L^-1,
// Phi(X) = lower triangle of X with its diagonal halved (I. Murray,
// "Differentiation of the Cholesky decomposition", arXiv:1602.07527, eq. 10).
// It runs as three tiled task DAGs of tile operations over T x T tiles of B:
//   1. P(i, j) = L(i:T, i)^T Lbar(i:T, j) for i >= j    (one GEMM per tile)
//   2. P <- Phi(P); W = P L^-1 by block columns j = T-1 .. 0:
//        W(j:T, j) -= W(j:T, j+1:T) L(j+1:T, j);  W(j:T, j) <- W(j:T, j) L(j, j)^-1
//   3. Abar = L^-T W by block rows i = T-1 .. 0:
//        Abar(i, :) -= L(i+1:T, i)^T Abar(i+1:T, :); Abar(i, :) <- L(i, i)^-T Abar(i, :)
// then G = (Abar + Abar^T) / 2.  W is lower block triangular and Abar is
// formed in the same buffer (its upper blocks start at zero).  Only L and Lbar
// are read: nothing of the forward pass is recomputed.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mustard_hip.h"
#include "mu_runtime.hpp"

namespace mu {
int set_error(int code, const std::string &msg);
}

namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p)
            (void)hipFree(p);
    }
};
struct DagH {
    mu_dag_t h = nullptr;
    ~DagH()
    {
        if (h)
            mu_dag_destroy(h);
    }
};
struct ExH {
    mu_exec_t h = nullptr;
    ~ExH()
    {
        if (h)
            mu_exec_destroy(h);
    }
};
}  // namespace

#define CK(x)                                                                                 \
    do {                                                                                      \
        int rc_ = (x);                                                                        \
        if (rc_ != 0)                                                                         \
            return rc_;                                                                       \
    } while (0)

extern "C" int mu_chol_backward(const double *L, int ldl, const double *Lbar, int ldlb, double *G,
                                int ldg, int n, int t, int prec, void *stream)
{
    if (!L || !Lbar || !G || n <= 0 || t <= 0 || n % t != 0 || ldl < n || ldlb < n || ldg < n ||
        prec < MU_PREC_FP64 || prec > MU_PREC_BF16)
        return mu::set_error(MU_EINVAL, "mu_chol_backward: need n % t == 0, leading dims >= n");
    hipStream_t s = (hipStream_t)stream;
    const int b = n / t;
    const size_t nn = (size_t)n * n;
    DevBuf buf;
    hipError_t he = hipMalloc(&buf.p, 3 * nn * sizeof(double));
    if (he != hipSuccess)
        return mu::set_error(MU_EHIP, std::string("mu_chol_backward: hipMalloc: ") +
                                          hipGetErrorString(he));
    double *Lz = (double *)buf.p, *Lb = Lz + nn, *P = Lb + nn;
    // lower triangles of L and Lbar (the strict upper parts of the inputs are
    // ignored: a tiled Cholesky leaves A's upper triangle in place), P = 0
    if (mu::tril_scaled(L, ldl, Lz, n, n, 1.0, s) != 0 ||
        mu::tril_scaled(Lbar, ldlb, Lb, n, n, 1.0, s) != 0 ||
        hipMemsetAsync(P, 0, nn * sizeof(double), s) != hipSuccess)
        return mu::set_error(MU_EHIP, "mu_chol_backward: kernel launch failed");
    auto lz = [&](int i, int j) { return Lz + (size_t)i * b + (size_t)j * b * n; };
    auto lb = [&](int i, int j) { return Lb + (size_t)i * b + (size_t)j * b * n; };
    auto pp = [&](int i, int j) { return P + (size_t)i * b + (size_t)j * b * n; };
    // records a phase as one DAG and runs it as one graph
    auto phase = [&](auto body) -> int {
        DagH d;
        ExH x;
        CK(mu_dag_create(&d.h));
        void *key = d.h;
        CK(mu_dag_capture_begin(d.h, key));
        int rc = body(d.h, key);
        int rc2 = mu_dag_capture_end(d.h, key);
        CK(rc);
        CK(rc2);
        CK(mu_exec_create(d.h, MU_MODE_TILED, nullptr, &x.h));
        CK(mu_exec_launch(x.h, stream));
        return hipStreamSynchronize(s) == hipSuccess ? 0
                                                      : mu::set_error(MU_EHIP, "phase failed");
    };
    auto gemm = [&](void *key, char ta, char tb, int m, int nn_, int k, double al, const double *A,
                    const double *Bm, double be, double *C) -> int {
        CK(mu_set_update_precision(key, prec));
        int rc = mu_dgemm(key, ta, tb, m, nn_, k, al, A, n, Bm, n, be, C, n);
        mu_set_update_precision(key, MU_PREC_FP64);
        return rc;
    };
    // 1. P(i, j) = L(i:T, i)^T Lbar(i:T, j), i >= j
    CK(phase([&](mu_dag_t d, void *key) -> int {
        for (int i = 0; i < t; i++)
            for (int j = 0; j <= i; j++) {
                CK(mu_dag_begin_task(d, i, j, nullptr, 0, nullptr));
                CK(gemm(key, 'T', 'N', b, b, (t - i) * b, 1.0, lz(i, i), lb(i, j), 0.0, pp(i, j)));
                CK(mu_dag_end_task(d));
            }
        return 0;
    }));
    // 2. Phi, then W = P L^-1 (right, lower, not transposed) by block columns
    if (mu::tril_scaled(P, n, P, n, n, 0.5, s) != 0)
        return mu::set_error(MU_EHIP, "mu_chol_backward: kernel launch failed");
    CK(phase([&](mu_dag_t d, void *key) -> int {
        for (int j = t - 1; j >= 0; j--) {
            std::vector<int> rd;
            for (int q = j + 1; q < t; q++) {
                rd.push_back(q);
                rd.push_back(0);
            }
            CK(mu_dag_begin_task(d, j, 0, rd.data(), (int)rd.size() / 2, nullptr));
            const int m = (t - j) * b;
            if (j < t - 1)
                CK(gemm(key, 'N', 'N', m, b, (t - 1 - j) * b, -1.0, pp(j, j + 1), lz(j + 1, j), 1.0,
                        pp(j, j)));
            CK(mu_dtrsm(key, 'R', 'L', 'N', 'N', m, b, 1.0, lz(j, j), n, pp(j, j), n));
            CK(mu_dag_end_task(d));
        }
        return 0;
    }));
    // 3. Abar = L^-T W (left, lower, transposed) by block rows
    CK(phase([&](mu_dag_t d, void *key) -> int {
        for (int i = t - 1; i >= 0; i--) {
            std::vector<int> rd;
            for (int q = i + 1; q < t; q++) {
                rd.push_back(q);
                rd.push_back(1);
            }
            CK(mu_dag_begin_task(d, i, 1, rd.data(), (int)rd.size() / 2, nullptr));
            if (i < t - 1)
                CK(gemm(key, 'T', 'N', b, n, (t - 1 - i) * b, -1.0, lz(i + 1, i), pp(i + 1, 0), 1.0,
                        pp(i, 0)));
            CK(mu_dtrsm(key, 'L', 'L', 'T', 'N', b, n, 1.0, lz(i, i), n, pp(i, 0), n));
            CK(mu_dag_end_task(d));
        }
        return 0;
    }));
    if (mu::sym_avg(P, n, G, ldg, n, s) != 0)
        return mu::set_error(MU_EHIP, "mu_chol_backward: kernel launch failed");
    return hipStreamSynchronize(s) == hipSuccess ? MU_OK
                                                  : mu::set_error(MU_EHIP, "mu_chol_backward failed");
}
