// mu_capi.hpp -- helpers of the extern "C" layer shared by its source files.
#pragma once
#include <string>

#include "mu_dag.hpp"

namespace mu {

// Set the text mu_last_error() returns; returns `code`.
int set_error(int code, const std::string &msg);

// mu_dtrsm's arguments as one recorded operation (a TRSM, or for alpha = 0
// the zero fill of B); MU_OK, or an error code with the message set.
int dtrsm_op(char side, char uplo, char transa, char diag, int m, int n, double alpha,
             const double *A, int lda, double *B, int ldb, OpRec &op);

// Whether `stream` is being recorded into a DAG (mu_dag_capture_begin).
bool stream_recording(void *stream);

// C = beta C + alpha A B with A an m x k FP32 matrix (column-major, lda in
// elements), B k x n and C m x n FP64, n <= 16: FP32 products on B rounded
// to FP32, added into C in FP64 (the substitution steps of a refinement
// solve on FP32 copies of the factors).  B and C are column-major, or with
// `rows` stored by rows (B(k, j) at B[k*ldb + j], C(i, j) at C[i*ldc + j]).
// Recorded into the DAG capturing `stream`, or run on it now.
int gemm_a32(void *stream, int m, int n, int k, double alpha, const float *A, int lda,
             const double *B, int ldb, double beta, double *C, int ldc, bool rows = false);

// The FP64 m x n block (src, lds) rounded to FP32 into (dst, ldd); with
// `transposed`, element (r, c) goes to dst[c + r*ldd], and with `plain` as
// well the untransposed copy to (plain, ldp), from the same read.  Recorded
// or run like gemm_a32.
int cvt_f32(void *stream, const double *src, int lds, float *dst, int ldd, int m, int n,
            bool transposed, float *plain = nullptr, int ldp = 0);

}  // namespace mu
