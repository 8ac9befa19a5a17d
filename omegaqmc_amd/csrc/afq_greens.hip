// Overlap matrices, LU (partial pivoting), slogdet, inverse and the
// half-rotated Green's function; one workgroup (4 waves) per walker, spins
// in sequence.
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
// The N x N work (N <= 64) is done by one wavefront with one lane per row
// (LU, column-major in LDS) or per column (inverse): no workgroup barriers
// inside the factorisation.  Ghalf = S^-1 phi^T runs on FP64 MFMA
// (v_mfma_f64_16x16x4f64: A[i][k] lane i+16k, B[k][j] lane j+16k, D[i][j]
// lane j+16(i%4) register i/4) over all four waves.
#include "afq_gemm.h"
#include "afq_gj.h"
#include "afq_greens.h"

namespace afq {

namespace {

constexpr int NT = 256;
typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma64(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double cabs1(cplx a) { return fabs(a.x) + fabs(a.y); }

// Orders the LDS accesses of one wavefront (lanes exchange data through LDS).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LU of A (column-major, A[j*LD + i] = S[i][j]) by one wavefront, lane = row.
// Returns the overlap sign*exp(logdet) in *ov (lane 0).
__device__ void wave_lu(cplx* A, int LD, int N, int* perm, cplx* ov) {
  const int l = threadIdx.x & 63;
  for (int k = 0; k < N; ++k) {
    double v = -1.0;
    if (l >= k && l < N) {
      v = cabs1(A[k * LD + l]);
      if (v != v) v = -0.5;  // a NaN row is taken only if nothing else is left
    }
    int bi = l;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ov2 = __shfl_xor(v, off);
      const int oi = __shfl_xor(bi, off);
      if (ov2 > v || (ov2 == v && oi < bi)) { v = ov2; bi = oi; }
    }
    const int piv = bi;
    if (l == 0) perm[k] = piv;
    if (piv != k && l < N) {
      const cplx a = A[l * LD + k];
      A[l * LD + k] = A[l * LD + piv];
      A[l * LD + piv] = a;
    }
    wave_sync();
    const cplx p = A[k * LD + k];
    if (l > k && l < N && (p.x != 0.0 || p.y != 0.0))
      A[k * LD + l] = cmul(A[k * LD + l], cdiv(mkc(1.0, 0.0), p));
    wave_sync();
    if (l > k && l < N) {
      const cplx li = A[k * LD + l];
      for (int j = k + 1; j < N; ++j) cfms(A[j * LD + l], li, A[j * LD + k]);
    }
    wave_sync();
  }
  if (l == 0) {
    double logabs = 0.0;
    cplx ph = mkc(1.0, 0.0);
    bool sing = false;
    for (int k = 0; k < N; ++k) {
      const cplx d = A[k * LD + k];
      const double a = hypot(d.x, d.y);
      if (a == 0.0) sing = true;
      else {
        logabs += log(a);
        ph = cmul(ph, mkc(d.x / a, d.y / a));
      }
      if (perm[k] != k) ph = mkc(-ph.x, -ph.y);
    }
    *ov = sing ? mkc(0.0, 0.0) : cscale(ph, exp(logabs));
  }
}

// Inv (row-major, Inv[i*LD + j]) = S^-1 from the LU, one wavefront, lane = column.
__device__ void wave_inverse(const cplx* A, int LD, const int* perm, int N, cplx* Inv) {
  const int j = threadIdx.x & 63;
  if (j >= N) return;
  int pos = j;
  for (int k = 0; k < N; ++k) {
    if (pos == k) pos = perm[k];
    else if (pos == perm[k]) pos = k;
  }
  for (int i = 0; i < N; ++i) Inv[i * LD + j] = mkc(i == pos ? 1.0 : 0.0, 0.0);
  for (int i = pos + 1; i < N; ++i) {
    cplx x = Inv[i * LD + j];
    for (int k = pos; k < i; ++k) cfms(x, A[k * LD + i], Inv[k * LD + j]);
    Inv[i * LD + j] = x;
  }
  for (int i = N - 1; i >= 0; --i) {
    cplx x = Inv[i * LD + j];
    for (int k = i + 1; k < N; ++k) cfms(x, A[k * LD + i], Inv[k * LD + j]);
    Inv[i * LD + j] = cdiv(x, A[i * LD + i]);
  }
}

// Ghalf (N x M) = Inv (N x N) . phi^T on MFMA; optional FP64 and split-bf16 outputs.
__device__ void ghalf_mfma(const cplx* Inv, int LD, int N, int M, const cplx* __restrict__ phi,
                           cplx* __restrict__ Ghalf, bf16_t* __restrict__ gb, long Ks,
                           long gplane, int ldq) {
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 15, lq = l >> 4;
  const int RTn = (N + 15) >> 4, RTm = (M + 15) >> 4;
  const int kN = (N + 3) & ~3;
  for (int pi = wv; pi < RTn * RTm; pi += NT / 64) {
    const int ti = pi / RTm, tj = pi % RTm;
    d4v cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    const int ai = 16 * ti + li;  // row of Inv
    const int bj = 16 * tj + li;  // row of phi (column of phi^T)
    for (int k0 = 0; k0 < kN; k0 += 4) {
      const int k = k0 + lq;
      const cplx a = (ai < N && k < N) ? Inv[ai * LD + k] : mkc(0.0, 0.0);
      const cplx b = (bj < M && k < N) ? phi[(long)bj * N + k] : mkc(0.0, 0.0);
      cr = mfma64(a.x, b.x, cr);
      ci = mfma64(a.x, b.y, ci);
      cr = mfma64(-a.y, b.y, cr);
      ci = mfma64(a.y, b.x, ci);
    }
    const int q = 16 * tj + li;
    if (q >= M) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * ti + lq + 4 * r;
      if (i >= N) continue;
      const long k = (long)i * M + q;
      if (Ghalf) Ghalf[k] = mkc(cr[r], ci[r]);
      if (gb) {
        const long k = (long)i * ldq + q;
        bf16_t h, lo;
        split2(cr[r], h, lo);
        gb[k] = h;
        gb[gplane + k] = lo;
        split2(ci[r], h, lo);
        gb[Ks + k] = h;
        gb[gplane + Ks + k] = lo;
      }
    }
  }
}

// One spin of one walker.  mode 0: overlap only; 1: + Ghalf.  eye != 0
// asserts trial == eye(M, N) exactly, so S[i][j] = phi[j][i] (the products
// with the zero entries of the trial contribute exact zeros).
// gb (optional): split-bf16 planes of Ghalf at this walker's real-part row;
// the imaginary row follows at +Ks and the lo plane at +gplane.
template <bool GENERAL>
__device__ __forceinline__ void greens_spin(int mode, int eye, int M, int N, int QCH,
                            const cplx* __restrict__ phi, const cplx* __restrict__ trial,
                            cplx* __restrict__ Ghalf, cplx* ov_out, char* smem, bf16_t* gb,
                            long Ks, long gplane, int ldq) {
  const int t = threadIdx.x;
  if (N == 0) {
    if (t == 0) *ov_out = mkc(1.0, 0.0);
    __syncthreads();
    return;
  }
  const int LD = N | 1;
  cplx* LU = (cplx*)smem;          // column-major [N][LD]
  cplx* Inv = LU + N * LD;         // row-major [N][LD]
  cplx* slab = Inv + N * LD;       // [QCH][N]
  cplx* slab_t = slab + QCH * N;   // [QCH][N]
  int* perm = (int*)(slab_t + QCH * N);

  if (!GENERAL || eye) {
    for (int e = t; e < N * N; e += NT) {
      const int j = e / N, i = e % N;
      LU[j * LD + i] = phi[(long)j * N + i];
    }
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
      if (e < N * N) LU[(e % N) * LD + e / N] = acc[r];
    }
  }
  __syncthreads();
  if (t < 64) wave_lu(LU, LD, N, perm, ov_out);
  __syncthreads();
  if (mode == 0) return;
  if (t < 64) wave_inverse(LU, LD, perm, N, Inv);
  __syncthreads();
  ghalf_mfma(Inv, LD, N, M, phi, Ghalf, gb, Ks, gplane, ldq);
  __syncthreads();
}

// GENERAL = false: both trials are identity matrices (S is a gather).
template <bool GENERAL>
__global__ __launch_bounds__(NT) void k_greens(int mode, int M, int Na, int Nb, int qch,
                                               const cplx* __restrict__ phia,
                                               const cplx* __restrict__ phib,
                                               const cplx* __restrict__ tu,
                                               const cplx* __restrict__ td,
                                               cplx* __restrict__ Gha, cplx* __restrict__ Ghb,
                                               cplx* __restrict__ ovlp, GhalfBf16 gbf) {
  extern __shared__ __align__(16) char smem[];
  __shared__ cplx ov[2];
  const long w = blockIdx.x;
  const long W = gridDim.x;
  const bool gh = (mode & 1) != 0;
  bf16_t* ga = (gh && gbf.a) ? gbf.a + 2 * w * gbf.Ka : nullptr;
  bf16_t* gbb = (gh && gbf.b) ? gbf.b + 2 * w * gbf.Kb : nullptr;
  greens_spin<GENERAL>(mode & 1, (mode >> 1) & 1, M, Na, qch, phia + w * M * Na, tu,
              (gh && Gha) ? Gha + w * Na * M : nullptr, &ov[0], smem, ga, gbf.Ka,
              2 * W * gbf.Ka, gbf.ld);
  __syncthreads();
  greens_spin<GENERAL>(mode & 1, (mode >> 2) & 1, M, Nb, qch, phib + w * M * Nb, td,
              (gh && Ghb) ? Ghb + w * Nb * M : nullptr, &ov[1], smem, gbb, gbf.Kb,
              2 * W * gbf.Kb, gbf.ld);
  __syncthreads();
  if (threadIdx.x == 0) ovlp[w] = cmul(ov[0], ov[1]);
}

using gj::GJ_LD;
using gj::GJ_NMAX;

// LDS of one wavefront of k_greens_gj: the pivot order, plus for Ghalf S^-1
// (N x GJ_LD complex) and the 2 KB staging of the split planes.
size_t gj_wave_lds(int mode, int N, bool planes) {
  return 256 + ((mode & 1) ? 16UL * N * GJ_LD : 0) + (((mode & 1) && planes) ? 2048 : 0);
}

// One walker per workgroup of two wavefronts, wave s = spin s; identity trials.
// GH: the Ghalf modes (a separate instantiation, so that the overlap-only
// kernel keeps the small register footprint of the elimination alone); DET:
// the overlap alone (no inverse), by the determinant-only elimination.
template <int NB, bool GH, bool DET = false>
__global__ __launch_bounds__(128) void k_greens_gj(int mode, int M, int Na, int Nb, int wlds,
                                                   const cplx* __restrict__ phia,
                                                   const cplx* __restrict__ phib,
                                                   cplx* __restrict__ Gha, cplx* __restrict__ Ghb,
                                                   cplx* __restrict__ ovlp, GhalfBf16 gbf,
                                                   cplx* __restrict__ inv_out) {
  extern __shared__ __align__(16) char smem[];
  __shared__ cplx ov[2];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const long w = blockIdx.x;
  const long W = gridDim.x;
  const int N = wv ? Nb : Na;
  char* base = smem + (long)wv * wlds;
  int* order = (int*)base;
  cplx* Inv = (cplx*)(order + 64);
  bf16_t* sb = (bf16_t*)(Inv + ((mode & 1) ? N * GJ_LD : 0));
  const cplx* phi = (wv ? phib + w * M * Nb : phia + w * M * Na);
  if (N == 0) {
    if (l == 0) ov[wv] = mkc(1.0, 0.0);
  } else {
    // S[i][j] = phi[j][i] for the identity trial: lane i loads row i
    cplx a[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) a[j] = (l < N && j < N) ? phi[(long)j * N + l] : mkc(0.0, 0.0);
    int mystep = 0;
    gj::wave_gj<NB, DET>(a, N, order, mystep, &ov[wv]);
    if (!GH && !DET && inv_out)
      gj::gj_store_inverse<NB>(a, N, mystep, order,
                               inv_out + w * ((long)Na * Na + (long)Nb * Nb) + (wv ? (long)Na * Na : 0),
                               N);
    if (GH && (mode & 1)) {
      gj::gj_store_inverse<NB>(a, N, mystep, order, Inv);
      const bool gh = true;
      cplx* G = wv ? Ghb : Gha;
      bf16_t* gb = wv ? gbf.b : gbf.a;
      const long Ks = wv ? gbf.Kb : gbf.Ka;
      gj::ghalf_mfma_wave<(NB + 3) / 4>(Inv, GJ_LD, N, M, phi, (gh && G) ? G + w * N * M : nullptr,
                                    gb ? gb + 2 * w * Ks : nullptr, Ks, 2 * W * Ks, gbf.ld, sb);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) ovlp[w] = cmul(ov[0], ov[1]);
}

// Ghalf = S^-1 phi^T for identity trials from S^-1 in global memory (row
// major, N x N per (walker, spin), spin b after spin a; written by
// k_greens_gj).  GHW wavefronts per (walker, spin), wavefront v taking the
// 16-orbital column tiles v, v + GHW, ...; each holds the S^-1 fragments in
// registers, and two wavefronts share a SIMD, so that one's phi loads and
// epilogue (split, LDS staging, stores) run under the other's FP64 MFMAs.
constexpr int GHW = 4;

template <int KQ>
__global__ __launch_bounds__(64 * GHW, 2) void k_ghalf_inv(int M, int Na, int Nb,
                                                           const cplx* __restrict__ inv,
                                                           const cplx* __restrict__ phia,
                                                           const cplx* __restrict__ phib,
                                                           cplx* __restrict__ Gha,
                                                           cplx* __restrict__ Ghb, GhalfBf16 gbf) {
  __shared__ __align__(16) bf16_t sbuf[GHW][1024];  // split-plane staging, 2 KB per wavefront
  const int wv = threadIdx.x >> 6;
  const long w = blockIdx.x, W = gridDim.x;
  const int spin = blockIdx.y;
  const int N = spin ? Nb : Na;
  if (N == 0) return;  // whole workgroup; no barriers below
  const cplx* Inv = inv + w * ((long)Na * Na + (long)Nb * Nb) + (spin ? (long)Na * Na : 0);
  const cplx* phi = spin ? phib + w * M * Nb : phia + w * M * Na;
  cplx* G = spin ? Ghb : Gha;
  bf16_t* gb = spin ? gbf.b : gbf.a;
  const long Ks = spin ? gbf.Kb : gbf.Ka;
  gj::ghalf_mfma_wave<KQ, false>(Inv, N, N, M, phi, G ? G + w * N * M : nullptr,
                          gb ? gb + 2 * w * Ks : nullptr, Ks, 2 * W * Ks, gbf.ld, sbuf[wv], wv,
                          GHW);
}

}  // namespace

size_t greens_inv_bytes(int W, int Na, int Nb) {
  return (size_t)W * ((size_t)Na * Na + (size_t)Nb * Nb) * sizeof(cplx);
}

static int greens_qch(int Na, int Nb) { return (Na > 48 || Nb > 48) ? 8 : 16; }

size_t greens_smem_bytes(int Na, int Nb) {
  const int N = Na > Nb ? Na : Nb;
  const int LD = N | 1;
  return (size_t)(2 * N * LD + 2 * greens_qch(Na, Nb) * N) * sizeof(cplx) + 64 * sizeof(int) + 64;
}

int greens_launch(hipStream_t s, int mode, int W, int M, int Na, int Nb, const double* phia,
                  const double* phib, const double* tu, const double* td, double* Gha,
                  double* Ghb, double* ovlp, const GhalfBf16* gbf, double* inv) {
  if (Na > 64 || Nb > 64) return fail(AFQ_EUNSUPPORTED, "nocc > 64 is not supported (got %d, %d)", Na, Nb);
  if (W <= 0) return AFQ_OK;
  const size_t sm = greens_smem_bytes(Na, Nb);
  if (sm + 64 > max_lds_bytes())
    return fail(AFQ_EUNSUPPORTED, "greens: %zu bytes of LDS needed, device allows %zu", sm,
                max_lds_bytes());
  GhalfBf16 g{};
  if (gbf) {
    g = *gbf;
    if (g.ld < M) return fail(AFQ_EINVAL, "greens: bf16 row stride %d < M=%d", g.ld, M);
    if ((g.a && g.Ka < (long)Na * g.ld) || (g.b && g.Kb < (long)Nb * g.ld))
      return fail(AFQ_EINVAL, "greens: bf16 plane row too short");
  }
  const bool eye_both = ((mode >> 1) & 1 || Na == 0) && ((mode >> 2) & 1 || Nb == 0);
  const int Nmax = Na > Nb ? Na : Nb;
  const bool planes = g.a || g.b;
  // with an S^-1 buffer, Ghalf is a second kernel (k_ghalf_inv) after the
  // elimination, which then runs in its overlap-only form
  const bool two = (mode & 1) && inv != nullptr;
  const int gmode = two ? (mode & ~1) : mode;
  const size_t wl = (gj_wave_lds(gmode, Nmax, planes) + 15) / 16 * 16;
  if (eye_both && Nmax <= GJ_NMAX && 2 * wl + 64 <= max_lds_bytes()) {
    const size_t sm2 = 2 * wl;
#define AFQ_GJ(NB)                                                                            \
  do {                                                                                          \
  if (gmode & 1)                                                                                \
    hipLaunchKernelGGL((k_greens_gj<NB, true>), dim3(W), dim3(128), sm2, s, gmode, M, Na, Nb,   \
                       (int)wl, (const cplx*)phia, (const cplx*)phib, (cplx*)Gha, (cplx*)Ghb,   \
                       (cplx*)ovlp, g, nullptr);                                                \
  else if (two)                                                                                 \
    hipLaunchKernelGGL((k_greens_gj<NB, false>), dim3(W), dim3(128), sm2, s, gmode, M, Na, Nb,  \
                       (int)wl, (const cplx*)phia, (const cplx*)phib, (cplx*)Gha, (cplx*)Ghb,   \
                       (cplx*)ovlp, g, (cplx*)inv);                                             \
  else                                                                                          \
    hipLaunchKernelGGL((k_greens_gj<NB, false, true>), dim3(W), dim3(128), sm2, s, gmode, M, Na, \
                       Nb, (int)wl, (const cplx*)phia, (const cplx*)phib, (cplx*)Gha,           \
                       (cplx*)Ghb, (cplx*)ovlp, g, nullptr);                                    \
  if (two) {                                                                                    \
    AFQ_CHECK_LAUNCH();                                                                         \
    hipLaunchKernelGGL((k_ghalf_inv<(NB + 3) / 4>), dim3(W, 2), dim3(64 * GHW), 0, s, M, Na, Nb, \
                       (const cplx*)inv, (const cplx*)phia, (const cplx*)phib, (cplx*)Gha,      \
                       (cplx*)Ghb, g);                                                          \
  }                                                                                             \
  } while (0)
    if (Nmax <= 8) AFQ_GJ(8);
    else if (Nmax <= 16) AFQ_GJ(16);
    else if (Nmax <= 24) AFQ_GJ(24);
    else if (Nmax <= 32) AFQ_GJ(32);
    else AFQ_GJ(40);
#undef AFQ_GJ
  } else if (eye_both)
    hipLaunchKernelGGL((k_greens<false>), dim3(W), dim3(NT), sm, s, mode, M, Na, Nb,
                       greens_qch(Na, Nb), (const cplx*)phia, (const cplx*)phib, (const cplx*)tu,
                       (const cplx*)td, (cplx*)Gha, (cplx*)Ghb, (cplx*)ovlp, g);
  else
    hipLaunchKernelGGL((k_greens<true>), dim3(W), dim3(NT), sm, s, mode, M, Na, Nb,
                       greens_qch(Na, Nb), (const cplx*)phia, (const cplx*)phib, (const cplx*)tu,
                       (const cplx*)td, (cplx*)Gha, (cplx*)Ghb, (cplx*)ovlp, g);
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
