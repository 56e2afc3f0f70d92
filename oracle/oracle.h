/*
 * oracle.h -- CPU oracle for the single-GPU tiled LU / Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py are the only callers, always as the checker
 * (or as the reported CPU baseline), never as the thing that is measured or
 * shipped.  The product library (mustard_amd/libmustard_hip.so) never links
 * or loads this file.
 *
 * All matrices are column-major doubles with a leading dimension, exactly as
 * the reference keeps them (mustard/lu_mustard.cu:278-281: tile (i,j) of an
 * N x N matrix cut in T x T tiles of B = N/T starts at A + i*B + j*B*N).
 */
#ifndef MUSTARD_ORACLE_H
#define MUSTARD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generators -------------------------------------------------------- */
/* include/gen.h:23-38: glibc srand(420), upper triangle row by row from
 * rand()/RAND_MAX, mirrored, N added to the diagonal. */
void or_gen_ref(double *A, size_t n);
/* Counter-based uniform stream: value k of stream `seed` in [0,1). */
double or_u01(uint64_t seed, uint64_t k);
/* M[k] = lo + (hi-lo) * or_u01(seed, k) for k < count. */
void or_gen_uniform(double *M, size_t count, uint64_t seed, double lo, double hi);
/* Config-1 Cholesky matrix: A = M M^T / n + n I, M uniform [-1,1) (col-major
 * M[i + j*n] = stream element i + j*n). Uses the host BLAS when loaded. */
void or_gen_spd(double *A, size_t n, uint64_t seed);

/* ---- host BLAS/LAPACK ---------------------------------------------------- */
/* dlopen a LAPACK/BLAS library.  `prefix` is the symbol prefix ("scipy_" for
 * the SciPy-bundled OpenBLAS, "" for a system liblapack).  Returns 0 on
 * success.  Without a library every routine below uses plain C loops. */
int or_blas_load(const char *path, const char *prefix);
int or_blas_loaded(void);
void or_blas_set_threads(int n);

/* ---- tile operations (the reference's library calls, restated) ---------- */
/* cusolverDnDgetrf with devIpiv = NULL (lu_mustard.cu:360-369): in-place
 * Doolittle LU without pivoting, unit-lower L below the diagonal, U on and
 * above it.  Returns 0 or the 1-based column of the first zero pivot. */
int or_getrf_nopiv(int n, double *A, int lda);
/* Unblocked triple-loop version of the same (brute force for small n). */
int or_getrf_nopiv_unblocked(int n, double *A, int lda);
/* cusolverDnDpotrf(LOWER) (cholesky_mustard.cu:238-246): lower Cholesky in
 * place, upper triangle untouched.  Returns 0 or LAPACK-style info. */
int or_potrf_lower(int n, double *A, int lda);
int or_potrf_lower_unblocked(int n, double *A, int lda);
/* B <- L^-1 B, L unit lower m x m, B m x n (lu_mustard.cu:413-422). */
void or_trsm_llnu(int m, int n, const double *L, int ldl, double *B, int ldb);
/* B <- B U^-1, U upper non-unit n x n, B m x n (lu_mustard.cu:459-468). */
void or_trsm_runn(int m, int n, const double *U, int ldu, double *B, int ldb);
/* B <- B L^-T, L lower non-unit n x n, B m x n (cholesky_mustard.cu:281-290). */
void or_trsm_rltn(int m, int n, const double *L, int ldl, double *B, int ldb);
/* C <- C - A B (lu_mustard.cu:511-522). A m x k, B k x n. */
void or_gemm_nn_sub(int m, int n, int k, const double *A, int lda,
                    const double *B, int ldb, double *C, int ldc);
/* C <- C - A B^T (cholesky_mustard.cu:376-387). A m x k, B n x k. */
void or_gemm_nt_sub(int m, int n, int k, const double *A, int lda,
                    const double *B, int ldb, double *C, int ldc);
/* lower(C) <- lower(C) - A A^T (cholesky_mustard.cu:327-333). A n x k. */
void or_syrk_ln_sub(int n, int k, const double *A, int lda, double *C, int ldc);

/* ---- tiled factorizations in the reference's task order ----------------- */
/* lu_mustard.cu:343-535 (per k: GETRF; TRSM_L(k,i) for i>k; then for i>k:
 * TRSM_R(i,k) followed by GEMM(i,j) for j>k).  Returns the first nonzero
 * GETRF info (global 1-based column) or 0. */
int or_tiled_lu(int n, int t, double *A);
/* cholesky_mustard.cu:221-400 (per k: POTRF; TRSM(i,k) for i>k; for i>k:
 * SYRK(i,i) then GEMM(j,i) for j>i). */
int or_tiled_cholesky(int n, int t, double *A);
/* The same task DAG executed as OpenMP tasks with depend clauses on tiles and
 * single-threaded BLAS per tile, on `threads` host threads (the CPU baseline). */
int or_tiled_lu_omp(int n, int t, double *A, int threads);
int or_tiled_cholesky_omp(int n, int t, double *A, int threads);

/* ---- verification -------------------------------------------------------- */
/* include/verify.h:17-71: the packed LU factor is split into unit-lower L and
 * upper U, L*U is formed by triple loop and sum |A - LU| is returned (the
 * reference passes when it is <= 1e-6).  O(n^3): small n only. */
double or_verify_lu_l1(const double *A, const double *F, int n);
/* include/verify.h:79-129: the same with L L^T. */
double or_verify_chol_l1(const double *A, const double *F, int n);
/* ||A - L U||_F / (||A||_F * n * eps) and ||A - L L^T||_F / (...), formed
 * with the host BLAS (or loops). */
double or_scaled_residual_lu(const double *A, const double *F, int n);
double or_scaled_residual_chol(const double *A, const double *F, int n);

/* ---- dense solves (the FP64 LAPACK oracle of the refinement forms) ------ */
/* X <- A^-1 X through the packed factors (n x n, nrhs columns, col-major). */
void or_lu_solve(int n, int nrhs, const double *F, double *X);
void or_chol_solve(int n, int nrhs, const double *F, double *X);

#ifdef __cplusplus
}
#endif
#endif
