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

}  // namespace mu
