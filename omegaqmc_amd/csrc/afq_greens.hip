// Overlap matrices, LU (partial pivoting), slogdet, inverse and the
// half-rotated Green's function, one workgroup per walker.
//
// Reference: OmegaQMC/observables/greens.py:12-198
//   S = phi^T Psi*      (einsum 'wpi,pj->wij')
//   Ghalf = S^-1 phi^T  (einsum 'wij,wqj->wiq')
//   G = Psi* Ghalf      (einsum 'pi,wiq->wpq')
//   ovlp = sign * exp(logdet) from slogdet(S), product over the two spins.
//
// LU follows LAPACK getrf semantics: the pivot of column k is the first row
// of maximal |re| + |im|; a zero pivot leaves its column unscaled.  A
// singular S gives ovlp = 0 (slogdet's sign 0, logdet -inf), and its
// inverse carries inf/NaN into Ghalf as in the reference.
//
// LDS: LU (N x N) + S^-1 (N x N) + a 16/32-row slab of phi, complex128; N <= 64.
#include "afq_gemm.h"
#include "afq_greens.h"

namespace afq {

namespace {

constexpr int NT = 256;

__device__ __forceinline__ double cabs1(cplx a) { return fabs(a.x) + fabs(a.y); }

// Factor S (in LDS, N x N row-major) in place; returns overlap via *ov (thread 0).
__device__ void lu_slogdet(cplx* LU, int* perm, int N, cplx* ov_out, double* red_v, int* red_i) {
  const int t = threadIdx.x;
  for (int k = 0; k < N; ++k) {
    // pivot search in column k, rows k..N-1 (wave 0)
    if (t < 64) {
      double best = -1.0;
      int bi = k;
      for (int r = k + t; r < N; r += 64) {
        double v = cabs1(LU[r * N + k]);
        if (v > best || (v != v && best == -1.0)) { best = v; bi = r; }
      }
      // reduce (max value, then smallest row)
      for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(best, off);
        int oi = __shfl_xor(bi, off);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (t == 0) { red_i[0] = bi; perm[k] = bi; }
    }
    __syncthreads();
    const int piv = red_i[0];
    if (piv != k) {
      for (int j = t; j < N; j += NT) {
        cplx a = LU[k * N + j];
        LU[k * N + j] = LU[piv * N + j];
        LU[piv * N + j] = a;
      }
    }
    __syncthreads();
    const cplx p = LU[k * N + k];
    const bool nz = (p.x != 0.0 || p.y != 0.0);
    if (nz) {
      const cplx rp = cdiv(mkc(1.0, 0.0), p);
      for (int i = k + 1 + t; i < N; i += NT) LU[i * N + k] = cmul(LU[i * N + k], rp);
    }
    __syncthreads();
    const int rem = N - k - 1;
    for (int e = t; e < rem * rem; e += NT) {
      int i = k + 1 + e / rem, j = k + 1 + e % rem;
      cfms(LU[i * N + j], LU[i * N + k], LU[k * N + j]);
    }
    __syncthreads();
  }
  if (t == 0) {
    double logabs = 0.0;
    cplx ph = mkc(1.0, 0.0);
    bool sing = false;
    for (int k = 0; k < N; ++k) {
      cplx d = LU[k * N + k];
      double a = hypot(d.x, d.y);
      if (a == 0.0) sing = true;
      else {
        logabs += log(a);
        ph = cmul(ph, mkc(d.x / a, d.y / a));
      }
      if (perm[k] != k) ph = mkc(-ph.x, -ph.y);
    }
    *ov_out = sing ? mkc(0.0, 0.0) : cscale(ph, exp(logabs));
  }
  (void)red_v;
}

// Column j of S^-1 from the packed LU (threads j < N), into Inv (row-major).
__device__ void lu_inverse(const cplx* LU, const int* perm, int N, cplx* Inv) {
  const int j = threadIdx.x;
  if (j < N) {
    int pos = j;
    for (int k = 0; k < N; ++k) {
      if (pos == k) pos = perm[k];
      else if (pos == perm[k]) pos = k;
    }
    for (int i = 0; i < N; ++i) Inv[i * N + j] = mkc(i == pos ? 1.0 : 0.0, 0.0);
    for (int i = pos + 1; i < N; ++i) {
      cplx x = Inv[i * N + j];
      for (int k = pos; k < i; ++k) cfms(x, LU[i * N + k], Inv[k * N + j]);
      Inv[i * N + j] = x;
    }
    for (int i = N - 1; i >= 0; --i) {
      cplx x = Inv[i * N + j];
      for (int k = i + 1; k < N; ++k) cfms(x, LU[i * N + k], Inv[k * N + j]);
      Inv[i * N + j] = cdiv(x, LU[i * N + i]);
    }
  }
}

// One spin of one walker.  mode 0: overlap only; 1: + Ghalf.  eye != 0
// asserts trial == eye(M, N) exactly, so S[i][j] = phi[j][i] (the products
// with the zero entries of the trial contribute exact zeros).
__device__ void greens_spin(int mode, int eye, int M, int N, int QCH,
                            const cplx* __restrict__ phi, const cplx* __restrict__ trial,
                            cplx* __restrict__ Ghalf, cplx* ov_out, char* smem) {
  const int t = threadIdx.x;
  if (N == 0) {
    if (t == 0) *ov_out = mkc(1.0, 0.0);
    __syncthreads();
    return;
  }
  cplx* LU = (cplx*)smem;
  cplx* Inv = LU + N * N;
  cplx* slab = Inv + N * N;
  cplx* slab_t = slab + QCH * N;
  int* perm = (int*)(slab_t + QCH * N);
  int* red_i = perm + 64;
  double* red_v = (double*)(red_i + 4);

  if (eye) {
    for (int e = t; e < N * N; e += NT) LU[e] = phi[(long)(e % N) * N + e / N];
  } else {
    // S[i][j] = sum_p phi[p][i] conj(trial[p][j]), p staged through LDS
    constexpr int RMAX = (64 * 64 + NT - 1) / NT;
    cplx acc[RMAX];
#pragma unroll
    for (int r = 0; r < RMAX; ++r) acc[r] = mkc(0.0, 0.0);
    for (int q0 = 0; q0 < M; q0 += QCH) {
      const int nq = min(QCH, M - q0);
      for (int e = t; e < nq * N; e += NT) {
        slab[e] = phi[(long)q0 * N + e];
        slab_t[e] = trial[(long)q0 * N + e];
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const int e = t + NT * r;
        if (e < N * N) {
          const int i = e / N, j = e % N;
          for (int q = 0; q < nq; ++q) cfma(acc[r], slab[q * N + i], cconj(slab_t[q * N + j]));
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int e = t + NT * r;
      if (e < N * N) LU[e] = acc[r];
    }
  }
  __syncthreads();
  lu_slogdet(LU, perm, N, ov_out, red_v, red_i);
  __syncthreads();
  if (mode == 0) return;
  lu_inverse(LU, perm, N, Inv);
  __syncthreads();
  // Ghalf[i][q] = sum_j Inv[i][j] phi[q][j], q in slabs of QCH rows of phi
  for (int q0 = 0; q0 < M; q0 += QCH) {
    int nq = min(QCH, M - q0);
    for (int e = t; e < nq * N; e += NT) slab[e] = phi[(long)q0 * N + e];
    __syncthreads();
    for (int e = t; e < nq * N; e += NT) {
      int i = e / nq, ql = e % nq;
      cplx acc = mkc(0.0, 0.0);
      for (int j = 0; j < N; ++j) cfma(acc, Inv[i * N + j], slab[ql * N + j]);
      Ghalf[(long)i * M + q0 + ql] = acc;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(NT) void k_greens(int mode, int M, int Na, int Nb, int qch,
                                               const cplx* __restrict__ phia,
                                               const cplx* __restrict__ phib,
                                               const cplx* __restrict__ tu,
                                               const cplx* __restrict__ td,
                                               cplx* __restrict__ Gha, cplx* __restrict__ Ghb,
                                               cplx* __restrict__ ovlp) {
  extern __shared__ __align__(16) char smem[];
  __shared__ cplx ov[2];
  const long w = blockIdx.x;
  greens_spin(mode & 1, (mode >> 1) & 1, M, Na, qch, phia + w * M * Na, tu,
              (mode & 1) ? Gha + w * Na * M : nullptr, &ov[0], smem);
  __syncthreads();
  greens_spin(mode & 1, (mode >> 2) & 1, M, Nb, qch, phib + w * M * Nb, td,
              (mode & 1) ? Ghb + w * Nb * M : nullptr, &ov[1], smem);
  __syncthreads();
  if (threadIdx.x == 0) ovlp[w] = cmul(ov[0], ov[1]);
}

}  // namespace

static int greens_qch(int Na, int Nb) { return (Na > 48 || Nb > 48) ? 8 : 16; }

size_t greens_smem_bytes(int Na, int Nb) {
  int N = Na > Nb ? Na : Nb;
  return (size_t)(2 * N * N + 2 * greens_qch(Na, Nb) * N) * sizeof(cplx) + 64 * sizeof(int) + 4 * sizeof(int) +
         64 * sizeof(double) + 64;
}

int greens_launch(hipStream_t s, int mode, int W, int M, int Na, int Nb, const double* phia,
                  const double* phib, const double* tu, const double* td, double* Gha,
                  double* Ghb, double* ovlp) {
  if (Na > 64 || Nb > 64) return fail(AFQ_EUNSUPPORTED, "nocc > 64 is not supported (got %d, %d)", Na, Nb);
  if (W <= 0) return AFQ_OK;
  size_t sm = greens_smem_bytes(Na, Nb);
  if (sm + 64 > max_lds_bytes())
    return fail(AFQ_EUNSUPPORTED, "greens: %zu bytes of LDS needed, device allows %zu", sm,
                max_lds_bytes());
  hipLaunchKernelGGL(k_greens, dim3(W), dim3(NT), sm, s, mode, M, Na, Nb, greens_qch(Na, Nb),
                     (const cplx*)phia,
                     (const cplx*)phib, (const cplx*)tu, (const cplx*)td, (cplx*)Gha, (cplx*)Ghb,
                     (cplx*)ovlp);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

// G_s[w] = conj(trial_s) (M x N) @ Ghalf_s[w] (N x M)
int full_g_launch(hipStream_t s, int W, int M, int N, const double* trial, const double* Ghalf,
                  double* G) {
  GemmArgs g;
  g.m = M; g.n = M; g.k = N; g.batch = W;
  g.A = trial; g.lda_m = N; g.lda_k = 1; g.sA = 0; g.a_cplx = true; g.conjA = true;
  g.B = Ghalf; g.ldb_k = M; g.ldb_n = 1; g.sB = (long)N * M; g.b_cplx = true;
  g.C = (cplx*)G; g.ldc = M; g.sC = (long)M * M;
  return gemm_fp64(s, g);
}

}  // namespace afq

extern "C" {

static int check_greens_args(afq_ctx* ctx, int W, int M, int Na, int Nb) {
  AFQ_REQUIRE(ctx != nullptr, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M,
              "bad shape W=%d M=%d Na=%d Nb=%d", W, M, Na, Nb);
  return AFQ_OK;
}

int afq_greens_function_force_bias(afq_ctx* ctx, int W, int M, int Na, int Nb, const double* phia,
                                   const double* phib, const double* tu, const double* td,
                                   double* Gha, double* Ghb, double* ovlp) {
  int rc = check_greens_args(ctx, W, M, Na, Nb);
  if (rc) return rc;
  return afq::greens_launch(ctx->stream, 1, W, M, Na, Nb, phia, phib, tu, td, Gha, Ghb, ovlp);
}

int afq_greens_function_overlap(afq_ctx* ctx, int W, int M, int Na, int Nb, const double* phia,
                                const double* phib, const double* tu, const double* td,
                                double* ovlp) {
  int rc = check_greens_args(ctx, W, M, Na, Nb);
  if (rc) return rc;
  return afq::greens_launch(ctx->stream, 0, W, M, Na, Nb, phia, phib, tu, td, nullptr, nullptr,
                            ovlp);
}

int afq_greens_function(afq_ctx* ctx, int W, int M, int Na, int Nb, const double* phia,
                        const double* phib, const double* tu, const double* td, double* Ga,
                        double* Gb, double* Gha, double* Ghb, double* ovlp) {
  int rc = check_greens_args(ctx, W, M, Na, Nb);
  if (rc) return rc;
  rc = afq::greens_launch(ctx->stream, 1, W, M, Na, Nb, phia, phib, tu, td, Gha, Ghb, ovlp);
  if (rc) return rc;
  rc = afq::full_g_launch(ctx->stream, W, M, Na, tu, Gha, Ga);
  if (rc) return rc;
  return afq::full_g_launch(ctx->stream, W, M, Nb, td, Ghb, Gb);
}

}  // extern "C"
