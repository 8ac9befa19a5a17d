// Walker re-orthonormalisation (OmegaQMC/afqmc_gto.py:126-181).
//
// The reference takes a Householder QR and multiplies the columns of Q by
// sign(diag R): it returns the unique thin-QR factor whose R has a positive
// real diagonal, and log_detR = sum_j log|R_jj|.  That factor is computed
// here by Cholesky QR applied twice (CholeskyQR2): R1 = chol(phi^H phi),
// Q1 = phi R1^-1, R2 = chol(Q1^H Q1), Q = Q1 R2^-1, so R = R2 R1 and
// log_detR = sum_j (log R1_jj + log R2_jj).  When the first Gram matrix is not
// numerically positive definite, the first pass is repeated with the diagonal
// shift s = 11 (M N + N (N+1)) u ||phi||_F^2 and a third pass is added
// (shifted CholeskyQR3, Fukaya, Kannan, Nakatsukasa, Yamamoto and Yanagisawa,
// SIAM J. Sci. Comput. 42 (2020)).  sign_detR is returned as 1: the
// reference's value depends on LAPACK's Householder sign convention and the
// driver discards it (afqmc_gto.py:1392).
//
// One workgroup (4 waves) per (walker, spin).  The Gram matrix and the product
// phi R^-1 run on FP64 MFMA (A[i][k] lane i+16k, B[k][j] lane j+16k, D[i][j]
// lane j+16(i%4) register i/4); the N x N Cholesky factor and its inverse are
// computed by one wavefront with one lane per column, in LDS.
#include "afq_kernels.h"

namespace afq {

namespace {

constexpr int NT = 256;
typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma64(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// G (column-major, G[j*LD + i] = (phi^H phi)[i][j]) on MFMA, all waves.
__device__ void gram(const cplx* __restrict__ phi, int M, int N, int LD, cplx* Gm) {
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lq = l >> 4;
  const int RT = (N + 15) >> 4;
  const int kM = (M + 3) & ~3;
  for (int pi = wv; pi < RT * RT; pi += NT / 64) {
    const int ti = pi / RT, tj = pi % RT;
    const int ci = 16 * ti + li, cj = 16 * tj + li;
    d4v cr = {0, 0, 0, 0}, cim = {0, 0, 0, 0};
    for (int k0 = 0; k0 < kM; k0 += 4) {
      const int p = k0 + lq;
      const cplx a = (p < M && ci < N) ? phi[(long)p * N + ci] : mkc(0.0, 0.0);
      const cplx b = (p < M && cj < N) ? phi[(long)p * N + cj] : mkc(0.0, 0.0);
      // conj(a) b
      cr = mfma64(a.x, b.x, cr);
      cim = mfma64(a.x, b.y, cim);
      cr = mfma64(a.y, b.y, cr);
      cim = mfma64(-a.y, b.x, cim);
    }
    const int j = 16 * tj + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * ti + lq + 4 * r;
      if (i < N && j < N) Gm[j * LD + i] = mkc(cr[r], cim[r]);
    }
  }
}

// In place: G -> R (upper, G = R^H R), one wavefront, lane = column.  Returns
// false (uniform) if a pivot is not positive and finite.  *logd += sum log R_jj.
__device__ bool wave_chol(cplx* Gm, int LD, int N, double shift, double* logd, bool* ok_lds) {
  const int j = threadIdx.x & 63;
  bool ok = true;
  double ld = 0.0;
  if (shift != 0.0 && j < N) Gm[j * LD + j].x += shift;
  wave_sync();
  for (int k = 0; k < N; ++k) {
    const double dkk = Gm[k * LD + k].x;
    const double d = sqrt(dkk);
    if (!(dkk > 0.0) || !(d <= 1.7976931348623157e308)) ok = false;
    if (!ok) break;
    ld += log(d);
    wave_sync();
    if (j == k) Gm[k * LD + k] = mkc(d, 0.0);
    if (j > k && j < N) Gm[j * LD + k] = mkc(Gm[j * LD + k].x / d, Gm[j * LD + k].y / d);
    wave_sync();
    if (j > k && j < N) {
      const cplx rkj = Gm[j * LD + k];
      for (int i = k + 1; i <= j; ++i) cfms(Gm[j * LD + i], cconj(Gm[i * LD + k]), rkj);
    }
    wave_sync();
  }
  if (j == 0) {
    *ok_lds = ok;
    if (ok) *logd += ld;
  }
  return ok;
}

// Rinv (column-major) = R^-1, one wavefront, lane = column.
__device__ void wave_rinv(const cplx* R, int LD, int N, cplx* Ri) {
  const int j = threadIdx.x & 63;
  if (j >= N) return;
  for (int i = 0; i < N; ++i) Ri[j * LD + i] = mkc(0.0, 0.0);
  Ri[j * LD + j] = cdiv(mkc(1.0, 0.0), R[j * LD + j]);
  for (int i = j - 1; i >= 0; --i) {
    cplx x = mkc(0.0, 0.0);
    for (int k = i + 1; k <= j; ++k) cfma(x, R[k * LD + i], Ri[j * LD + k]);
    Ri[j * LD + i] = cdiv(mkc(-x.x, -x.y), R[i * LD + i]);
  }
}

// phi <- phi Rinv in place; a wave owns whole row tiles (reads them fully
// before writing them).
template <int CT>
__device__ void apply_rinv(cplx* phi, int M, int N, int LD, const cplx* Ri) {
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lq = l >> 4;
  const int RTm = (M + 15) >> 4, RTn = (N + 15) >> 4;
  const int kN = (N + 3) & ~3;
  for (int tm = wv; tm < RTm; tm += NT / 64) {
    d4v cr[CT], ci[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) { cr[c] = d4v{0, 0, 0, 0}; ci[c] = d4v{0, 0, 0, 0}; }
    const int row = 16 * tm + li;
    for (int k0 = 0; k0 < kN; k0 += 4) {
      const int k = k0 + lq;
      const cplx a = (row < M && k < N) ? phi[(long)row * N + k] : mkc(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        if (c >= RTn) continue;
        const int j = 16 * c + li;
        const cplx b = (j < N && k < N) ? Ri[j * LD + k] : mkc(0.0, 0.0);
        cr[c] = mfma64(a.x, b.x, cr[c]);
        ci[c] = mfma64(a.x, b.y, ci[c]);
        cr[c] = mfma64(-a.y, b.y, cr[c]);
        ci[c] = mfma64(a.y, b.x, ci[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      if (c >= RTn) continue;
      const int j = 16 * c + li;
      if (j >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * tm + lq + 4 * r;
        if (i < M) phi[(long)i * N + j] = mkc(cr[c][r], ci[c][r]);
      }
    }
  }
}

template <int CT>
__global__ __launch_bounds__(NT) void k_cholqr(int M, int Na, int Nb, cplx* phia, cplx* phib,
                                               double* __restrict__ ldr) {
  extern __shared__ __align__(16) char smem[];
  __shared__ double logd;
  __shared__ bool ok;
  __shared__ double fro;
  const long w = blockIdx.x;
  const int spin = blockIdx.y;
  const int N = spin ? Nb : Na;
  cplx* phi = spin ? phib + w * M * Nb : phia + w * M * Na;
  if (threadIdx.x == 0) logd = 0.0;
  if (N == 0) {
    __syncthreads();
    if (threadIdx.x == 0) ldr[2 * w + spin] = 0.0;
    return;
  }
  const int LD = N | 1;
  cplx* Gm = (cplx*)smem;
  cplx* Ri = Gm + N * LD;
  int passes = 2;  // uniform across the workgroup
  for (int pass = 0; pass < passes; ++pass) {
    __syncthreads();
    gram(phi, M, N, LD, Gm);
    __syncthreads();
    if (threadIdx.x < 64) wave_chol(Gm, LD, N, 0.0, &logd, &ok);
    __syncthreads();
    if (!ok) {
      if (pass > 0) break;  // not numerically full rank: leave NaN in log_detR
      // shifted first pass: s = 11 (M N + N (N+1)) u ||phi||_F^2
      gram(phi, M, N, LD, Gm);
      __syncthreads();
      if (threadIdx.x == 0) {
        double f = 0.0;
        for (int i = 0; i < N; ++i) f += Gm[i * LD + i].x;
        fro = f;
      }
      __syncthreads();
      const double sft = 11.0 * ((double)M * N + (double)N * (N + 1)) * 0x1.0p-53 * fro;
      if (threadIdx.x < 64) wave_chol(Gm, LD, N, sft, &logd, &ok);
      __syncthreads();
      if (!ok) break;
      passes = 3;
    }
    if (threadIdx.x < 64) wave_rinv(Gm, LD, N, Ri);
    __syncthreads();
    apply_rinv<CT>(phi, M, N, LD, Ri);
  }
  __syncthreads();
  if (threadIdx.x == 0) ldr[2 * w + spin] = ok ? logd : __builtin_nan("");
}

__global__ void k_sum_spins(int W, const double* __restrict__ l2, double* __restrict__ ldr,
                            cplx* __restrict__ sgn) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  ldr[w] = l2[2 * w] + l2[2 * w + 1];
  if (sgn) sgn[w] = mkc(1.0, 0.0);
}

}  // namespace

size_t ortho_scratch_bytes(int W) { return sizeof(double) * 2 * (size_t)W; }

int ortho_launch(hipStream_t s, int W, int M, int Na, int Nb, double* phia, double* phib,
                 double* log_detR, double* sign_detR, double* scratch2w) {
  if (Na > 64 || Nb > 64) return fail(AFQ_EUNSUPPORTED, "nocc > 64 is not supported");
  if (W <= 0) return AFQ_OK;
  const int N = std::max(Na, Nb);
  const size_t sm = (size_t)2 * N * (N | 1) * sizeof(cplx) + 64;
  const dim3 grid(W, 2);
  const int RTn = (N + 15) / 16;
  if (RTn <= 1)
    hipLaunchKernelGGL((k_cholqr<1>), grid, dim3(NT), sm, s, M, Na, Nb, (cplx*)phia, (cplx*)phib, scratch2w);
  else if (RTn == 2)
    hipLaunchKernelGGL((k_cholqr<2>), grid, dim3(NT), sm, s, M, Na, Nb, (cplx*)phia, (cplx*)phib, scratch2w);
  else if (RTn == 3)
    hipLaunchKernelGGL((k_cholqr<3>), grid, dim3(NT), sm, s, M, Na, Nb, (cplx*)phia, (cplx*)phib, scratch2w);
  else
    hipLaunchKernelGGL((k_cholqr<4>), grid, dim3(NT), sm, s, M, Na, Nb, (cplx*)phia, (cplx*)phib, scratch2w);
  AFQ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sum_spins, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, scratch2w, log_detR,
                     (cplx*)sign_detR);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace afq

extern "C" int afq_orthogonalize_walkers(afq_ctx* ctx, int W, int M, int Na, int Nb, double* phia,
                                         double* phib, double* log_detR, double* sign_detR) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M, "bad shape");
  AFQ_REQUIRE(log_detR != nullptr, "log_detR is required");
  double* sc = (double*)afq::scratch(ctx, afq::ortho_scratch_bytes(W));
  if (!sc && W) return AFQ_ENOMEM;
  return afq::ortho_launch(ctx->stream, W, M, Na, Nb, phia, phib, log_detR, sign_detR, sc);
}
