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
// lane j+16(i%4) register i/4); R^-1 comes from one wavefront: for N <= 40 by
// an LDL^H elimination held in registers (one lane per row), else by the
// column-by-column Cholesky factor and its inverse in LDS.
#include "afq_gj.h"
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

// G (column-major, G[j*LD + i] = (phi^H phi)[i][j]) on MFMA, all waves: the
// upper tiles are computed and mirrored (G is Hermitian).
__device__ void gram(const cplx* __restrict__ phi, int M, int N, int LD, cplx* Gm) {
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lq = l >> 4;
  const int RT = (N + 15) >> 4;
  const int kM = (M + 3) & ~3;
  for (int pi = wv; pi < RT * (RT + 1) / 2; pi += NT / 64) {
    int ti = 0, t = pi;
    while (t >= RT - ti) {
      t -= RT - ti;
      ++ti;
    }
    const int tj = ti + t;
    const int ci = 16 * ti + li, cj = 16 * tj + li;
    d4v cr = {0, 0, 0, 0}, cim = {0, 0, 0, 0};
#pragma unroll 5
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
      if (i < N && j < N) {
        Gm[j * LD + i] = mkc(cr[r], cim[r]);
        if (ti != tj) Gm[i * LD + j] = mkc(cr[r], -cim[r]);
      }
    }
  }
}

// G = L D L^H by elimination without pivoting, one wavefront with lane i
// holding row i of [G | I] in registers (G read from LDS, plus `shift` on the
// diagonal).  As in the Gauss-Jordan kernels the row is kept rotated so that
// the active column is register 0 and the identity part grows into the last
// registers: step k subtracts f = G_ik / d_k times the pivot row from the rows
// below k and appends the new column (1 for the pivot row, -f below, 0 above),
// so after N steps row i holds row i of L^-1 (column c in register
// c - N + NB).  d_k = Re of the k-th pivot is the k-th Cholesky pivot, so
// R = D^1/2 L^H and R^-1 = L^-H D^-1/2 is written to Ri (column-major).
// lg = sum_k log sqrt(d_k) (every lane).  Returns false (uniform) if a pivot
// is not positive and finite, as the column-by-column factorisation would.
// The pivot row reaches the other lanes by readlane (scalar operands of the
// update), not through LDS.
template <int NB>
__device__ bool wave_ldl_rinv(const cplx* Gm, int LD, int N, double shift, cplx* Ri, double& lg) {
  const int l = threadIdx.x & 63;
  cplx a[NB];
  // G[l][j] = conj(G[j][l]), read along the contiguous column l of Gm
  const cplx* gl = Gm + (l < N ? l : 0) * LD;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const cplx g = gl[j < N ? j : 0];
    a[j] = (l < N && j < N) ? mkc(g.x, -g.y) : mkc(0.0, 0.0);
    if (j == l) a[j].x += shift;
  }
  bool ok = true;
  double myd = 1.0;
  for (int k = 0; k < N; ++k) {
    const double d = gj::rl64(a[0].x, k);
    ok = ok && d > 0.0 && d <= 1.7976931348623157e308;
    const bool below = l > k && l < N;
    const double id = 1.0 / d;
    const cplx f = below ? mkc(a[0].x * id, a[0].y * id) : mkc(0.0, 0.0);
#pragma unroll
    for (int j = 0; j + 1 < NB; ++j) {
      cplx x = a[j + 1];
      cfms(x, f, gj::rlc(a[j + 1], k));
      a[j] = x;
    }
    a[NB - 1] = l == k ? mkc(1.0, 0.0) : (below ? mkc(-f.x, -f.y) : mkc(0.0, 0.0));
    if (l == k) myd = d;
  }
  if (l < N) {
    const double sc = 1.0 / sqrt(myd);
    cplx* rl = Ri + l * LD + N - NB;  // column c of row l at rl[c + NB - N]
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int c = q + N - NB;
      if (c < 0) continue;
      rl[q] = c <= l ? mkc(a[q].x * sc, -a[q].y * sc) : mkc(0.0, 0.0);
    }
  }
  lg = l < N ? 0.5 * log(myd) : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lg += __shfl_xor(lg, off);
  wave_sync();
  return ok;
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
#pragma unroll 2
    for (int k0 = 0; k0 < kN; k0 += 4) {
      const int k = k0 + lq;
      const cplx a = (row < M && k < N) ? phi[(long)row * N + k] : mkc(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        // R^-1 is upper triangular: column tile c has no k >= 16 (c + 1)
        // (the skipped products are exact zeros)
        if (c >= RTn || k0 >= 16 * (c + 1)) continue;  // uniform
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

// N > 40: the whole QR of one (walker, spin) in one workgroup, R^-1 from the
// column-by-column Cholesky factor and its inverse in LDS.
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
  // Passes of CholeskyQR2; if the first Gram matrix is not numerically
  // positive definite the first pass is redone with the shift and a third
  // pass is added.  Every branch below is uniform across the workgroup.
  int pass = 0, passes = 2;
  bool shifted = false;
  while (pass < passes) {
    __syncthreads();
    gram(phi, M, N, LD, Gm);
    __syncthreads();
    double sft = 0.0;
    if (shifted && pass == 0) {
      // s = 11 (M N + N (N+1)) u ||phi||_F^2
      if (threadIdx.x == 0) {
        double f = 0.0;
        for (int i = 0; i < N; ++i) f += Gm[i * LD + i].x;
        fro = f;
      }
      __syncthreads();
      sft = 11.0 * ((double)M * N + (double)N * (N + 1)) * 0x1.0p-53 * fro;
    }
    if (threadIdx.x < 64 && wave_chol(Gm, LD, N, sft, &logd, &ok)) wave_rinv(Gm, LD, N, Ri);
    __syncthreads();
    if (!ok) {
      if (pass > 0 || shifted) break;  // not numerically full rank: NaN in log_detR
      shifted = true;
      passes = 3;
      continue;
    }
    apply_rinv<CT>(phi, M, N, LD, Ri);
    ++pass;
  }
  __syncthreads();
  if (threadIdx.x == 0) ldr[2 * w + spin] = ok ? logd : __builtin_nan("");
}

// N <= 40: one launch of each of the three kernels below per pass.  The
// state of a (walker, spin) decides which passes it runs:
//   0 two passes; 1 the first Gram matrix was not numerically positive
//   definite, so the first pass used the shift and a third pass follows;
//   2 not numerically full rank (stops, log_detR NaN).
__device__ __forceinline__ bool qr_runs(int pass, int st) {
  return st != 2 && (pass < 2 || st == 1);
}

// G_s = phi_s^H phi_s into Gbuf (N x N column-major per (walker, spin), stride NN)
template <int CT>
__global__ __launch_bounds__(NT) void k_qr_gram(int pass, int M, int Na, int Nb, int NN,
                                                const cplx* phia, const cplx* phib,
                                                const int* __restrict__ state,
                                                cplx* __restrict__ Gbuf) {
  const long w = blockIdx.x;
  const int spin = blockIdx.y;
  const int N = spin ? Nb : Na;
  const long id = 2 * w + spin;
  if (N == 0 || (pass > 0 && !qr_runs(pass, state[id]))) return;
  gram(spin ? phib + w * M * Nb : phia + w * M * Na, M, N, N, Gbuf + id * NN);
}

// R^-1 of the pass (one wavefront per (walker, spin)), the state and log_detR.
template <int NB>
__global__ __launch_bounds__(64) void k_qr_factor(int pass, int M, int Na, int Nb, int NN,
                                                  const cplx* __restrict__ Gbuf,
                                                  cplx* __restrict__ Ribuf, int* __restrict__ state,
                                                  double* __restrict__ ldr2) {
  const long w = blockIdx.x;
  const int spin = blockIdx.y;
  const int N = spin ? Nb : Na;
  const long id = 2 * w + spin;
  if (N == 0) {
    if (pass == 0 && threadIdx.x == 0) {
      ldr2[id] = 0.0;
      state[id] = 0;
    }
    return;
  }
  int st = pass == 0 ? 0 : state[id];
  if (!qr_runs(pass, st)) return;
  const cplx* G = Gbuf + id * NN;
  cplx* Ri = Ribuf + id * NN;
  double lg = 0.0;
  bool ok = wave_ldl_rinv<NB>(G, N, N, 0.0, Ri, lg);
  if (!ok && pass == 0) {
    // s = 11 (M N + N (N+1)) u ||phi||_F^2
    double f = 0.0;
    for (int i = 0; i < N; ++i) f += G[i * N + i].x;
    const double sft = 11.0 * ((double)M * N + (double)N * (N + 1)) * 0x1.0p-53 * f;
    ok = wave_ldl_rinv<NB>(G, N, N, sft, Ri, lg);
    st = ok ? 1 : 2;
  } else if (!ok) {
    st = 2;
  }
  if (threadIdx.x == 0) {
    const double prev = pass == 0 ? 0.0 : ldr2[id];
    ldr2[id] = ok ? prev + lg : __builtin_nan("");
    state[id] = st;
  }
}

// phi <- phi R^-1 for the (walker, spin) pairs that run this pass
template <int CT>
__global__ __launch_bounds__(NT) void k_qr_apply(int pass, int M, int Na, int Nb, int NN,
                                                 cplx* phia, cplx* phib,
                                                 const int* __restrict__ state,
                                                 const cplx* __restrict__ Ribuf) {
  extern __shared__ __align__(16) char smem[];
  const long w = blockIdx.x;
  const int spin = blockIdx.y;
  const int N = spin ? Nb : Na;
  const long id = 2 * w + spin;
  if (N == 0 || !qr_runs(pass, state[id])) return;
  const int LD = N | 1;
  cplx* Ri = (cplx*)smem;
  const cplx* src = Ribuf + id * NN;
  for (int e = threadIdx.x; e < N * N; e += NT) Ri[(e / N) * LD + e % N] = src[e];
  __syncthreads();
  apply_rinv<CT>(spin ? phib + w * M * Nb : phia + w * M * Na, M, N, LD, Ri);
}

__global__ void k_sum_spins(int W, const double* __restrict__ l2, double* __restrict__ ldr,
                            cplx* __restrict__ sgn) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  ldr[w] = l2[2 * w] + l2[2 * w + 1];
  if (sgn) sgn[w] = mkc(1.0, 0.0);
}

}  // namespace

size_t ortho_scratch_bytes(int W, int Na, int Nb) {
  const size_t N = std::max(Na, Nb);
  size_t n = align256(sizeof(double) * 2 * (size_t)W) + align256(sizeof(int) * 2 * (size_t)W);
  if (N <= 40) n += 2 * align256(sizeof(cplx) * 2 * (size_t)W * N * N);
  return n;
}

int ortho_launch(hipStream_t s, int W, int M, int Na, int Nb, double* phia, double* phib,
                 double* log_detR, double* sign_detR, void* scratch) {
  if (Na > 64 || Nb > 64) return fail(AFQ_EUNSUPPORTED, "nocc > 64 is not supported");
  if (W <= 0) return AFQ_OK;
  const int N = std::max(Na, Nb);
  char* p = (char*)scratch;
  double* ldr2 = (double*)p;
  p += align256(sizeof(double) * 2 * (size_t)W);
  int* state = (int*)p;
  p += align256(sizeof(int) * 2 * (size_t)W);
  const dim3 grid(W, 2);
  if (N <= 40) {
    const int NN = N * N;
    cplx* Gbuf = (cplx*)p;
    cplx* Ribuf = (cplx*)(p + align256(sizeof(cplx) * 2 * (size_t)W * NN));
    const size_t sm = (size_t)N * (N | 1) * sizeof(cplx) + 16;
    for (int pass = 0; pass < 3; ++pass) {
#define AFQ_QR_PASS(CT, NB)                                                                       \
  do {                                                                                            \
    hipLaunchKernelGGL((k_qr_gram<CT>), grid, dim3(NT), 0, s, pass, M, Na, Nb, NN,                \
                       (const cplx*)phia, (const cplx*)phib, state, Gbuf);                       \
    hipLaunchKernelGGL((k_qr_factor<NB>), grid, dim3(64), 0, s, pass, M, Na, Nb, NN, Gbuf, Ribuf, \
                       state, ldr2);                                                              \
    hipLaunchKernelGGL((k_qr_apply<CT>), grid, dim3(NT), sm, s, pass, M, Na, Nb, NN, (cplx*)phia, \
                       (cplx*)phib, state, Ribuf);                                                \
  } while (0)
      if (N <= 8) AFQ_QR_PASS(1, 8);
      else if (N <= 16) AFQ_QR_PASS(1, 16);
      else if (N <= 24) AFQ_QR_PASS(2, 24);
      else if (N <= 32) AFQ_QR_PASS(2, 32);
      else AFQ_QR_PASS(3, 40);
#undef AFQ_QR_PASS
      AFQ_CHECK_LAUNCH();
    }
  } else {
    const size_t sm = (size_t)2 * N * (N | 1) * sizeof(cplx) + 64;
    if (N <= 48)
      hipLaunchKernelGGL((k_cholqr<3>), grid, dim3(NT), sm, s, M, Na, Nb, (cplx*)phia,
                         (cplx*)phib, ldr2);
    else
      hipLaunchKernelGGL((k_cholqr<4>), grid, dim3(NT), sm, s, M, Na, Nb, (cplx*)phia,
                         (cplx*)phib, ldr2);
    AFQ_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_sum_spins, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, ldr2, log_detR,
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
  void* sc = afq::scratch(ctx, afq::ortho_scratch_bytes(W, Na, Nb));
  if (!sc && W) return AFQ_ENOMEM;
  return afq::ortho_launch(ctx->stream, W, M, Na, Nb, phia, phib, log_detR, sign_detR, sc);
}
