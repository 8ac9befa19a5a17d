#pragma once
#include "afq_common.h"

namespace afq {
// mode 0: overlap only; 1: Ghalf (W,N,M) + overlap.  Complex interleaved.
// inv (optional, greens_inv_bytes): S^-1 scratch of the two-kernel Ghalf path.
int greens_launch(hipStream_t s, int mode, int W, int M, int Na, int Nb, const double* phia,
                  const double* phib, const double* tu, const double* td, double* Gha,
                  double* Ghb, double* ovlp, const GhalfBf16* gbf = nullptr,
                  double* inv = nullptr);
// S^-1 buffer of greens_launch (identity trials, N <= 40: Ghalf by a second kernel)
size_t greens_inv_bytes(int W, int Na, int Nb);
int full_g_launch(hipStream_t s, int W, int M, int N, const double* trial, const double* Ghalf,
                  double* G);
}  // namespace afq
