// Walker re-orthonormalisation (OmegaQMC/afqmc_gto.py:126-181).
//
// The reference takes a Householder QR and multiplies the columns of Q by
// sign(diag R), i.e. it returns the unique thin-QR factor whose R has a
// positive real diagonal, and log_detR = sum_j log|R_jj|.  That factor is
// produced here directly by classical Gram-Schmidt with one
// re-orthogonalisation pass (CGS2, r_jj = ||v_j|| > 0), which agrees with
// Householder to rounding for numerically full-rank walkers.  sign_detR is
// returned as 1: the reference's value depends on LAPACK's Householder
// sign convention and the driver discards it (afqmc_gto.py:1392).
//
// One workgroup per walker; each thread owns rows p = t, t + 256, ...;
// dot products are reduced across the workgroup through LDS.
#include "afq_kernels.h"

namespace afq {

namespace {

constexpr int NT = 256;
constexpr int NMAXC = 64;

__device__ void cgs2_spin(int M, int N, cplx* __restrict__ phi, double* logdet, double* red) {
  // red layout: [4 waves][NMAXC] complex partials + [NMAXC] complex coefficients
  cplx* part = (cplx*)red;
  cplx* coef = part + 4 * NMAXC;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  double ld = 0.0;
  for (int j = 0; j < N; ++j) {
    for (int pass = 0; pass < 2; ++pass) {
      if (j > 0) {
        for (int k = 0; k < j; ++k) {
          cplx a = mkc(0, 0);
          for (int p = t; p < M; p += NT) cfma(a, cconj(phi[(long)p * N + k]), phi[(long)p * N + j]);
          a.x = wave_sum(a.x);
          a.y = wave_sum(a.y);
          if (lane == 0) part[wid * NMAXC + k] = a;
        }
        __syncthreads();
        if (t < j) {
          cplx c = part[t];
          for (int q = 1; q < (NT >> 6); ++q) c = cadd(c, part[q * NMAXC + t]);
          coef[t] = c;
        }
        __syncthreads();
        for (int p = t; p < M; p += NT) {
          cplx v = phi[(long)p * N + j];
          for (int k = 0; k < j; ++k) cfms(v, phi[(long)p * N + k], coef[k]);
          phi[(long)p * N + j] = v;
        }
        __syncthreads();
      }
    }
    double s = 0.0;
    for (int p = t; p < M; p += NT) {
      cplx v = phi[(long)p * N + j];
      s = fma(v.x, v.x, fma(v.y, v.y, s));
    }
    s = block_sum(s, (double*)(coef + NMAXC));
    const double r = sqrt(s);
    for (int p = t; p < M; p += NT) {
      cplx v = phi[(long)p * N + j];
      phi[(long)p * N + j] = mkc(v.x / r, v.y / r);
    }
    ld += log(r);
    __syncthreads();
  }
  if (t == 0) *logdet = ld;
}

__global__ __launch_bounds__(NT) void k_ortho(int M, int Na, int Nb, cplx* __restrict__ phia,
                                              cplx* __restrict__ phib, double* __restrict__ ldr,
                                              cplx* __restrict__ sgn) {
  __shared__ double red[(4 * NMAXC + NMAXC) * 2 + 32];
  __shared__ double lds[2];
  const long w = blockIdx.x;
  if (threadIdx.x == 0) { lds[0] = 0.0; lds[1] = 0.0; }
  __syncthreads();
  cgs2_spin(M, Na, phia + w * M * Na, &lds[0], red);
  __syncthreads();
  cgs2_spin(M, Nb, phib + w * M * Nb, &lds[1], red);
  __syncthreads();
  if (threadIdx.x == 0) {
    ldr[w] = lds[0] + lds[1];
    if (sgn) sgn[w] = mkc(1.0, 0.0);
  }
}

}  // namespace

int ortho_launch(hipStream_t s, int W, int M, int Na, int Nb, double* phia, double* phib,
                 double* log_detR, double* sign_detR) {
  if (Na > NMAXC || Nb > NMAXC) return fail(AFQ_EUNSUPPORTED, "nocc > 64 is not supported");
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_ortho, dim3(W), dim3(NT), 0, s, M, Na, Nb, (cplx*)phia, (cplx*)phib,
                     log_detR, (cplx*)sign_detR);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace afq

extern "C" int afq_orthogonalize_walkers(afq_ctx* ctx, int W, int M, int Na, int Nb, double* phia,
                                         double* phib, double* log_detR, double* sign_detR) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M, "bad shape");
  AFQ_REQUIRE(log_detR != nullptr, "log_detR is required");
  return afq::ortho_launch(ctx->stream, W, M, Na, Nb, phia, phib, log_detR, sign_detR);
}
