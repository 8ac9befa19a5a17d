// Internal GEMM entry points shared by the kernel files.
#pragma once
#include "afq_common.h"

namespace afq {

// Batched, strided C = alpha * op(A) * op(B) (+ beta * C) with complex output.
// Element A(m,k) = A[b*sA + m*lda_m + k*lda_k] (same for B(k,n)); A and B are
// real (double) or complex (double2) as flagged; conjA conjugates A.
// If R != nullptr the epilogue is the Taylor update: C = alpha*acc (beta
// ignored) and R += C.
struct GemmArgs {
  int m = 0, n = 0, k = 0, batch = 1;
  const void* A = nullptr;
  long lda_m = 0, lda_k = 0, sA = 0;
  bool a_cplx = false, conjA = false;
  const void* B = nullptr;
  long ldb_k = 0, ldb_n = 0, sB = 0;
  bool b_cplx = false, conjB = false;
  cplx* C = nullptr;
  long ldc = 0, sC = 0;
  cplx* R = nullptr;
  long ldr = 0, sR = 0;
  cplx alpha = {1.0, 0.0};
  double beta = 0.0;
};

int gemm_fp64(hipStream_t stream, const GemmArgs& g);

}  // namespace afq
