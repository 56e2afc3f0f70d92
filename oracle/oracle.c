/*
 * oracle.c -- CPU oracle for the single-GPU tiled LU / Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Every function states the
 * reference lines it restates.  The reference's arithmetic lives in cuBLAS /
 * cuSOLVER (closed, CUDA >= 12.3, not under the reference tree); the LAPACK
 * semantics of those calls are restated here, either through a dlopen'ed host
 * LAPACK (the SciPy-bundled OpenBLAS, symbol prefix "scipy_") or through the
 * plain loops below when no library is loaded.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* generators                                                               */
/* ======================================================================== */

void or_gen_ref(double *h_A, size_t n)
{
    /* include/gen.h:23-38, same loop order and the same int loop indices. */
    srand(420);
    for (int i = 0; i < (int)n; i++)
        for (int j = i; j < (int)n; j++)
            h_A[(size_t)i * n + j] = (double)rand() / (double)RAND_MAX;
    for (int i = 0; i < (int)n; i++)
        for (int j = i; j >= 0; j--)
            h_A[(size_t)i * n + j] = h_A[(size_t)j * n + i];
    for (int i = 0; i < (int)n; i++)
        h_A[(size_t)i * n + i] = h_A[(size_t)i * n + i] + (double)n;
}

/* SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014).  Integer-only, so the device generator
 * in the product library produces the same bits. */
static inline uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

double or_u01(uint64_t seed, uint64_t k)
{
    uint64_t z = sm64(sm64(seed) + k);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0); /* 2^-53 */
}

void or_gen_uniform(double *M, size_t count, uint64_t seed, double lo, double hi)
{
    for (size_t k = 0; k < count; k++)
        M[k] = lo + (hi - lo) * or_u01(seed, k);
}

/* ======================================================================== */
/* host BLAS / LAPACK through dlopen                                        */
/* ======================================================================== */

typedef void (*dgemm_t)(const char *, const char *, const int *, const int *, const int *,
                        const double *, const double *, const int *, const double *, const int *,
                        const double *, double *, const int *, size_t, size_t);
typedef void (*dtrsm_t)(const char *, const char *, const char *, const char *, const int *,
                        const int *, const double *, const double *, const int *, double *,
                        const int *, size_t, size_t, size_t, size_t);
typedef void (*dsyrk_t)(const char *, const char *, const int *, const int *, const double *,
                        const double *, const int *, const double *, double *, const int *,
                        size_t, size_t);
typedef void (*dpotrf_t)(const char *, const int *, double *, const int *, int *, size_t);
typedef void (*setthr_t)(int);

static struct {
    void *h;
    dgemm_t gemm;
    dtrsm_t trsm;
    dsyrk_t syrk;
    dpotrf_t potrf;
    setthr_t setthr;
} g_blas;

int or_blas_load(const char *path, const char *prefix)
{
    char name[128];
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h)
        return -1;
#define SYM(field, base, type)                                   \
    snprintf(name, sizeof name, "%s%s", prefix, base);           \
    g_blas.field = (type)dlsym(h, name);                         \
    if (!g_blas.field) { dlclose(h); memset(&g_blas, 0, sizeof g_blas); return -2; }
    SYM(gemm, "dgemm_", dgemm_t)
    SYM(trsm, "dtrsm_", dtrsm_t)
    SYM(syrk, "dsyrk_", dsyrk_t)
    SYM(potrf, "dpotrf_", dpotrf_t)
#undef SYM
    snprintf(name, sizeof name, "%sopenblas_set_num_threads", prefix);
    g_blas.setthr = (setthr_t)dlsym(h, name);
    g_blas.h = h;
    return 0;
}

int or_blas_loaded(void) { return g_blas.h != NULL; }

void or_blas_set_threads(int n)
{
    if (g_blas.setthr)
        g_blas.setthr(n);
}

#define AT(P, ld, r, c) (P)[(size_t)(r) + (size_t)(c) * (size_t)(ld)]

static void gemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc)
{
    if (m <= 0 || n <= 0)
        return;
    if (g_blas.gemm) {
        g_blas.gemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc, 1, 1);
        return;
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int p = 0; p < k; p++) {
                double a = ta == 'N' ? AT(A, lda, i, p) : AT(A, lda, p, i);
                double b = tb == 'N' ? AT(B, ldb, p, j) : AT(B, ldb, j, p);
                s += a * b;
            }
            AT(C, ldc, i, j) = (beta == 0.0 ? 0.0 : beta * AT(C, ldc, i, j)) + alpha * s;
        }
}

/* ======================================================================== */
/* tile operations                                                          */
/* ======================================================================== */

int or_getrf_nopiv_unblocked(int n, double *A, int lda)
{
    /* Right-looking Doolittle elimination without row exchanges: the
     * factorisation cusolverDnDgetrf computes when devIpiv == NULL. */
    int info = 0;
    for (int c = 0; c < n; c++) {
        double piv = AT(A, lda, c, c);
        if (piv == 0.0) {
            if (!info)
                info = c + 1;
            continue;
        }
        for (int r = c + 1; r < n; r++)
            AT(A, lda, r, c) /= piv;
        for (int cc = c + 1; cc < n; cc++) {
            double u = AT(A, lda, c, cc);
            for (int r = c + 1; r < n; r++)
                AT(A, lda, r, cc) -= AT(A, lda, r, c) * u;
        }
    }
    return info;
}

int or_getrf_nopiv(int n, double *A, int lda)
{
    /* Blocked right-looking version (panel of 64 columns, then TRSM + GEMM on
     * the host BLAS).  LAPACK's dgetrf always pivots, so the no-pivot form is
     * written out here; on a strictly diagonally dominant matrix dgetrf picks
     * identity pivots and both agree. */
    const int nb = 64;
    int info = 0;
    if (!g_blas.gemm)
        return or_getrf_nopiv_unblocked(n, A, lda);
    for (int j = 0; j < n; j += nb) {
        int jb = n - j < nb ? n - j : nb;
        for (int c = j; c < j + jb; c++) {
            double piv = AT(A, lda, c, c);
            if (piv == 0.0) {
                if (!info)
                    info = c + 1;
                continue;
            }
            for (int r = c + 1; r < n; r++)
                AT(A, lda, r, c) /= piv;
            for (int cc = c + 1; cc < j + jb; cc++) {
                double u = AT(A, lda, c, cc);
                for (int r = c + 1; r < n; r++)
                    AT(A, lda, r, cc) -= AT(A, lda, r, c) * u;
            }
        }
        int rest = n - j - jb;
        if (rest > 0) {
            or_trsm_llnu(jb, rest, &AT(A, lda, j, j), lda, &AT(A, lda, j, j + jb), lda);
            gemm('N', 'N', rest, rest, jb, -1.0, &AT(A, lda, j + jb, j), lda,
                 &AT(A, lda, j, j + jb), lda, 1.0, &AT(A, lda, j + jb, j + jb), lda);
        }
    }
    return info;
}

int or_potrf_lower_unblocked(int n, double *A, int lda)
{
    /* Right-looking lower Cholesky (LAPACK dpotf2 semantics, upper triangle
     * left untouched, info = first non-positive pivot). */
    for (int c = 0; c < n; c++) {
        double d = AT(A, lda, c, c);
        if (!(d > 0.0))
            return c + 1;
        d = sqrt(d);
        AT(A, lda, c, c) = d;
        for (int r = c + 1; r < n; r++)
            AT(A, lda, r, c) /= d;
        for (int cc = c + 1; cc < n; cc++) {
            double l = AT(A, lda, cc, c);
            for (int r = cc; r < n; r++)
                AT(A, lda, r, cc) -= AT(A, lda, r, c) * l;
        }
    }
    return 0;
}

int or_potrf_lower(int n, double *A, int lda)
{
    if (g_blas.potrf) {
        int info = 0;
        g_blas.potrf("L", &n, A, &lda, &info, 1);
        return info;
    }
    return or_potrf_lower_unblocked(n, A, lda);
}

void or_trsm_llnu(int m, int n, const double *L, int ldl, double *B, int ldb)
{
    if (g_blas.trsm) {
        const double one = 1.0;
        g_blas.trsm("L", "L", "N", "U", &m, &n, &one, L, &ldl, B, &ldb, 1, 1, 1, 1);
        return;
    }
    for (int j = 0; j < n; j++)
        for (int r = 0; r < m; r++) {
            double x = AT(B, ldb, r, j);
            for (int q = r + 1; q < m; q++)
                AT(B, ldb, q, j) -= AT(L, ldl, q, r) * x;
        }
}

void or_trsm_runn(int m, int n, const double *U, int ldu, double *B, int ldb)
{
    if (g_blas.trsm) {
        const double one = 1.0;
        g_blas.trsm("R", "U", "N", "N", &m, &n, &one, U, &ldu, B, &ldb, 1, 1, 1, 1);
        return;
    }
    /* X U = B: column c of X = (B[:,c] - sum_{q<c} X[:,q] U[q,c]) / U[c,c] */
    for (int c = 0; c < n; c++) {
        for (int q = 0; q < c; q++) {
            double u = AT(U, ldu, q, c);
            for (int i = 0; i < m; i++)
                AT(B, ldb, i, c) -= AT(B, ldb, i, q) * u;
        }
        double d = AT(U, ldu, c, c);
        for (int i = 0; i < m; i++)
            AT(B, ldb, i, c) /= d;
    }
}

void or_trsm_rltn(int m, int n, const double *L, int ldl, double *B, int ldb)
{
    if (g_blas.trsm) {
        const double one = 1.0;
        g_blas.trsm("R", "L", "T", "N", &m, &n, &one, L, &ldl, B, &ldb, 1, 1, 1, 1);
        return;
    }
    /* X L^T = B: column c of X = (B[:,c] - sum_{q<c} X[:,q] L[c,q]) / L[c,c] */
    for (int c = 0; c < n; c++) {
        for (int q = 0; q < c; q++) {
            double l = AT(L, ldl, c, q);
            for (int i = 0; i < m; i++)
                AT(B, ldb, i, c) -= AT(B, ldb, i, q) * l;
        }
        double d = AT(L, ldl, c, c);
        for (int i = 0; i < m; i++)
            AT(B, ldb, i, c) /= d;
    }
}

void or_gemm_nn_sub(int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                    double *C, int ldc)
{
    gemm('N', 'N', m, n, k, -1.0, A, lda, B, ldb, 1.0, C, ldc);
}

void or_gemm_nt_sub(int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                    double *C, int ldc)
{
    gemm('N', 'T', m, n, k, -1.0, A, lda, B, ldb, 1.0, C, ldc);
}

void or_syrk_ln_sub(int n, int k, const double *A, int lda, double *C, int ldc)
{
    if (g_blas.syrk) {
        const double m1 = -1.0, one = 1.0;
        g_blas.syrk("L", "N", &n, &k, &m1, A, &lda, &one, C, &ldc, 1, 1);
        return;
    }
    for (int j = 0; j < n; j++)
        for (int i = j; i < n; i++) {
            double s = 0.0;
            for (int p = 0; p < k; p++)
                s += AT(A, lda, i, p) * AT(A, lda, j, p);
            AT(C, ldc, i, j) -= s;
        }
}

/* ======================================================================== */
/* tiled factorizations                                                     */
/* ======================================================================== */

#define TILE(A, n, b, i, j) ((A) + (size_t)(i) * (b) + (size_t)(j) * (b) * (size_t)(n))

int or_tiled_lu(int n, int t, double *A)
{
    /* mustard/lu_mustard.cu:343-535 */
    int b = n / t, info = 0;
    for (int k = 0; k < t; k++) {
        int r = or_getrf_nopiv(b, TILE(A, n, b, k, k), n);
        if (r && !info)
            info = k * b + r;
        for (int i = k + 1; i < t; i++)
            or_trsm_llnu(b, b, TILE(A, n, b, k, k), n, TILE(A, n, b, k, i), n);
        for (int i = k + 1; i < t; i++) {
            or_trsm_runn(b, b, TILE(A, n, b, k, k), n, TILE(A, n, b, i, k), n);
            for (int j = k + 1; j < t; j++)
                or_gemm_nn_sub(b, b, b, TILE(A, n, b, i, k), n, TILE(A, n, b, k, j), n,
                               TILE(A, n, b, i, j), n);
        }
    }
    return info;
}

int or_tiled_cholesky(int n, int t, double *A)
{
    /* mustard/cholesky_mustard.cu:221-400 */
    int b = n / t, info = 0;
    for (int k = 0; k < t; k++) {
        int r = or_potrf_lower(b, TILE(A, n, b, k, k), n);
        if (r && !info)
            info = k * b + r;
        for (int i = k + 1; i < t; i++)
            or_trsm_rltn(b, b, TILE(A, n, b, k, k), n, TILE(A, n, b, i, k), n);
        for (int i = k + 1; i < t; i++) {
            or_syrk_ln_sub(b, b, TILE(A, n, b, i, k), n, TILE(A, n, b, i, i), n);
            for (int j = i + 1; j < t; j++)
                or_gemm_nt_sub(b, b, b, TILE(A, n, b, j, k), n, TILE(A, n, b, i, k), n,
                               TILE(A, n, b, j, i), n);
        }
    }
    return info;
}

int or_tiled_lu_omp(int n, int t, double *A, int threads)
{
    /* The task DAG of or_tiled_lu, submitted in the same order as OpenMP tasks
     * whose depend clauses name the tiles each task reads and writes. */
    int b = n / t, info = 0;
    char *dep = calloc((size_t)t * t, 1);
    or_blas_set_threads(1);
#pragma omp parallel num_threads(threads)
#pragma omp single
    for (int k = 0; k < t; k++) {
#pragma omp task depend(inout : dep[k * t + k]) firstprivate(k) shared(info)
        {
            int r = or_getrf_nopiv(b, TILE(A, n, b, k, k), n);
            if (r) {
#pragma omp critical
                if (!info)
                    info = k * b + r;
            }
        }
        for (int i = k + 1; i < t; i++) {
#pragma omp task depend(in : dep[k * t + k]) depend(inout : dep[k * t + i]) firstprivate(k, i)
            or_trsm_llnu(b, b, TILE(A, n, b, k, k), n, TILE(A, n, b, k, i), n);
        }
        for (int i = k + 1; i < t; i++) {
#pragma omp task depend(in : dep[k * t + k]) depend(inout : dep[i * t + k]) firstprivate(k, i)
            or_trsm_runn(b, b, TILE(A, n, b, k, k), n, TILE(A, n, b, i, k), n);
            for (int j = k + 1; j < t; j++) {
#pragma omp task depend(in : dep[i * t + k], dep[k * t + j]) depend(inout : dep[i * t + j]) \
    firstprivate(k, i, j)
                or_gemm_nn_sub(b, b, b, TILE(A, n, b, i, k), n, TILE(A, n, b, k, j), n,
                               TILE(A, n, b, i, j), n);
            }
        }
    }
    free(dep);
    return info;
}

int or_tiled_cholesky_omp(int n, int t, double *A, int threads)
{
    int b = n / t, info = 0;
    char *dep = calloc((size_t)t * t, 1);
    or_blas_set_threads(1);
#pragma omp parallel num_threads(threads)
#pragma omp single
    for (int k = 0; k < t; k++) {
#pragma omp task depend(inout : dep[k * t + k]) firstprivate(k) shared(info)
        {
            int r = or_potrf_lower(b, TILE(A, n, b, k, k), n);
            if (r) {
#pragma omp critical
                if (!info)
                    info = k * b + r;
            }
        }
        for (int i = k + 1; i < t; i++) {
#pragma omp task depend(in : dep[k * t + k]) depend(inout : dep[i * t + k]) firstprivate(k, i)
            or_trsm_rltn(b, b, TILE(A, n, b, k, k), n, TILE(A, n, b, i, k), n);
        }
        for (int i = k + 1; i < t; i++) {
#pragma omp task depend(in : dep[i * t + k]) depend(inout : dep[i * t + i]) firstprivate(k, i)
            or_syrk_ln_sub(b, b, TILE(A, n, b, i, k), n, TILE(A, n, b, i, i), n);
            for (int j = i + 1; j < t; j++) {
#pragma omp task depend(in : dep[j * t + k], dep[i * t + k]) depend(inout : dep[j * t + i]) \
    firstprivate(k, i, j)
                or_gemm_nt_sub(b, b, b, TILE(A, n, b, j, k), n, TILE(A, n, b, i, k), n,
                               TILE(A, n, b, j, i), n);
            }
        }
    }
    free(dep);
    return info;
}

/* ======================================================================== */
/* verification                                                             */
/* ======================================================================== */

double or_verify_lu_l1(const double *A, const double *F, int n)
{
    /* include/verify.h:17-71: row-major unit-lower L and upper U from the
     * column-major packed factor, newA = L*U by triple loop (i, j, k order),
     * error = sum |A - newA|.  A is read in the reference's row-major view. */
    size_t nn = (size_t)n * n;
    double *L = malloc(nn * sizeof(double)), *U = calloc(nn, sizeof(double));
    double *P = calloc(nn, sizeof(double));
    memcpy(L, F, nn * sizeof(double));
    for (int i = 0; i < n; i++) {
        for (int j = i; j < n; j++) {
            double t = L[i + (size_t)j * n];
            L[i + (size_t)j * n] = L[(size_t)i * n + j];
            L[(size_t)i * n + j] = t;
            U[(size_t)i * n + j] = L[(size_t)i * n + j];
            L[(size_t)i * n + j] = 0;
        }
        L[(size_t)i * n + i] = 1;
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++)
                P[(size_t)i * n + j] += L[(size_t)i * n + k] * U[(size_t)k * n + j];
    double err = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            err += fabs(A[(size_t)i * n + j] - P[(size_t)i * n + j]);
    free(L);
    free(U);
    free(P);
    return err;
}

double or_verify_chol_l1(const double *A, const double *F, int n)
{
    /* include/verify.h:79-129 */
    size_t nn = (size_t)n * n;
    double *L = malloc(nn * sizeof(double)), *P = calloc(nn, sizeof(double));
    memcpy(L, F, nn * sizeof(double));
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            L[i + (size_t)j * n] = 0;
            double t = L[i + (size_t)j * n];
            L[i + (size_t)j * n] = L[(size_t)i * n + j];
            L[(size_t)i * n + j] = t;
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++)
                P[(size_t)i * n + j] += L[(size_t)i * n + k] * L[k + (size_t)j * n];
    double err = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            err += fabs(A[(size_t)i * n + j] - P[(size_t)i * n + j]);
    free(L);
    free(P);
    return err;
}

static double frob(const double *X, size_t count)
{
    double s = 0;
    for (size_t k = 0; k < count; k++)
        s += X[k] * X[k];
    return sqrt(s);
}

double or_scaled_residual_lu(const double *A, const double *F, int n)
{
    size_t nn = (size_t)n * n;
    double *L = calloc(nn, sizeof(double)), *U = calloc(nn, sizeof(double));
    double *R = malloc(nn * sizeof(double));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            if (i > j)
                AT(L, n, i, j) = AT(F, n, i, j);
            else
                AT(U, n, i, j) = AT(F, n, i, j);
            if (i == j)
                AT(L, n, i, j) = 1.0;
        }
    memcpy(R, A, nn * sizeof(double));
    gemm('N', 'N', n, n, n, -1.0, L, n, U, n, 1.0, R, n);
    double r = frob(R, nn) / (frob(A, nn) * n * DBL_EPSILON);
    free(L);
    free(U);
    free(R);
    return r;
}

double or_scaled_residual_chol(const double *A, const double *F, int n)
{
    size_t nn = (size_t)n * n;
    double *L = calloc(nn, sizeof(double)), *R = malloc(nn * sizeof(double));
    for (int j = 0; j < n; j++)
        for (int i = j; i < n; i++)
            AT(L, n, i, j) = AT(F, n, i, j);
    memcpy(R, A, nn * sizeof(double));
    gemm('N', 'T', n, n, n, -1.0, L, n, L, n, 1.0, R, n);
    double r = frob(R, nn) / (frob(A, nn) * n * DBL_EPSILON);
    free(L);
    free(R);
    return r;
}

void or_gen_spd(double *A, size_t n, uint64_t seed)
{
    size_t nn = n * n;
    double *M = malloc(nn * sizeof(double));
    or_gen_uniform(M, nn, seed, -1.0, 1.0);
    gemm('N', 'T', (int)n, (int)n, (int)n, 1.0, M, (int)n, M, (int)n, 0.0, A, (int)n);
    for (size_t j = 0; j < n; j++)
        for (size_t i = 0; i < n; i++)
            A[i + j * n] = A[i + j * n] / (double)n + (i == j ? (double)n : 0.0);
    /* make the matrix exactly symmetric (the product is symmetric up to the
     * summation order of the two triangles) */
    for (size_t j = 0; j < n; j++)
        for (size_t i = j + 1; i < n; i++)
            A[j + i * n] = A[i + j * n];
    free(M);
}

/* ======================================================================== */
/* dense solves                                                             */
/* ======================================================================== */

void or_lu_solve(int n, int nrhs, const double *F, double *X)
{
    or_trsm_llnu(n, nrhs, F, n, X, n);
    /* X <- U^-1 X : back substitution column by column */
    if (g_blas.trsm) {
        const double one = 1.0;
        g_blas.trsm("L", "U", "N", "N", &n, &nrhs, &one, F, &n, X, &n, 1, 1, 1, 1);
        return;
    }
    for (int j = 0; j < nrhs; j++)
        for (int r = n - 1; r >= 0; r--) {
            double x = AT(X, n, r, j) / AT(F, n, r, r);
            AT(X, n, r, j) = x;
            for (int q = 0; q < r; q++)
                AT(X, n, q, j) -= AT(F, n, q, r) * x;
        }
}

void or_chol_solve(int n, int nrhs, const double *F, double *X)
{
    if (g_blas.trsm) {
        const double one = 1.0;
        g_blas.trsm("L", "L", "N", "N", &n, &nrhs, &one, F, &n, X, &n, 1, 1, 1, 1);
        g_blas.trsm("L", "L", "T", "N", &n, &nrhs, &one, F, &n, X, &n, 1, 1, 1, 1);
        return;
    }
    for (int j = 0; j < nrhs; j++) {
        for (int r = 0; r < n; r++) {
            double x = AT(X, n, r, j) / AT(F, n, r, r);
            AT(X, n, r, j) = x;
            for (int q = r + 1; q < n; q++)
                AT(X, n, q, j) -= AT(F, n, q, r) * x;
        }
        for (int r = n - 1; r >= 0; r--) {
            double s = AT(X, n, r, j);
            for (int q = r + 1; q < n; q++)
                s -= AT(F, n, q, r) * AT(X, n, q, j);
            AT(X, n, r, j) = s / AT(F, n, r, r);
        }
    }
}
