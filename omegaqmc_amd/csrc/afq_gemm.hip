// Generic batched FP64 GEMM with complex output (tiled through LDS, FMA
// micro-tiles).  It serves every contraction that has no dedicated MFMA
// kernel: small per-walker products (S = phi^T Psi*, G = Psi* Ghalf), the
// one-body half step, and the reference-shaped ABI entry points.
//
// Tile: 64 x 64 outputs per 256-thread workgroup, K staged 8 at a time;
// each thread owns a 4 x 4 complex micro-tile at rows ty + 16 r and columns
// tx + 16 c so that stores of a row segment are coalesced.
#include "afq_gemm.h"

namespace afq {

namespace {

constexpr int BM = 64, BN = 64, BK = 8, NT = 256;

template <bool AC, bool BC>
__global__ __launch_bounds__(NT) void k_gemm(GemmArgs g) {
  __shared__ cplx As[BK][BM + 1];
  __shared__ cplx Bs[BK][BN + 1];
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const long b = blockIdx.z;
  const double* Ar = (const double*)g.A;
  const cplx* Ac = (const cplx*)g.A;
  const double* Br = (const double*)g.B;
  const cplx* Bc = (const cplx*)g.B;
  const long aoff = b * g.sA, boff = b * g.sB;
  const bool a_kfast = (g.lda_k == 1);
  const bool b_nfast = (g.ldb_n == 1);

  cplx acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = mkc(0.0, 0.0);

  for (int k0 = 0; k0 < g.k; k0 += BK) {
    // A tile (BM x BK): 512 elements, 2 per thread
#pragma unroll
    for (int i = 0; i < (BM * BK) / NT; ++i) {
      int e = t + i * NT;
      int mm, kk;
      if (a_kfast) { mm = e / BK; kk = e % BK; } else { mm = e % BM; kk = e / BM; }
      int gm = m0 + mm, gk = k0 + kk;
      cplx v = mkc(0.0, 0.0);
      if (gm < g.m && gk < g.k) {
        long off = aoff + (long)gm * g.lda_m + (long)gk * g.lda_k;
        if (AC) { v = Ac[off]; if (g.conjA) v.y = -v.y; } else { v.x = Ar[off]; }
      }
      As[kk][mm] = v;
    }
#pragma unroll
    for (int i = 0; i < (BN * BK) / NT; ++i) {
      int e = t + i * NT;
      int nn, kk;
      if (b_nfast) { nn = e % BN; kk = e / BN; } else { nn = e / BK; kk = e % BK; }
      int gn = n0 + nn, gk = k0 + kk;
      cplx v = mkc(0.0, 0.0);
      if (gn < g.n && gk < g.k) {
        long off = boff + (long)gk * g.ldb_k + (long)gn * g.ldb_n;
        if (BC) { v = Bc[off]; if (g.conjB) v.y = -v.y; } else { v.x = Br[off]; }
      }
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
      cplx a[4], bb[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) bb[c] = Bs[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (AC && BC) {
            cfma(acc[r][c], a[r], bb[c]);
          } else if (AC) {
            acc[r][c].x = fma(a[r].x, bb[c].x, acc[r][c].x);
            acc[r][c].y = fma(a[r].y, bb[c].x, acc[r][c].y);
          } else if (BC) {
            acc[r][c].x = fma(a[r].x, bb[c].x, acc[r][c].x);
            acc[r][c].y = fma(a[r].x, bb[c].y, acc[r][c].y);
          } else {
            acc[r][c].x = fma(a[r].x, bb[c].x, acc[r][c].x);
          }
        }
    }
    __syncthreads();
  }

  const long coff = b * g.sC;
  const long roff = b * g.sR;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int gm = m0 + ty + 16 * r;
    if (gm >= g.m) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int gn = n0 + tx + 16 * c;
      if (gn >= g.n) continue;
      cplx v = cmul(g.alpha, acc[r][c]);
      long ci = coff + (long)gm * g.ldc + gn;
      if (g.R) {
        g.C[ci] = v;
        long ri = roff + (long)gm * g.ldr + gn;
        g.R[ri] = cadd(g.R[ri], v);
      } else {
        if (g.beta != 0.0) v = cadd(v, cscale(g.C[ci], g.beta));
        g.C[ci] = v;
      }
    }
  }
}

}  // namespace

int gemm_fp64(hipStream_t stream, const GemmArgs& g) {
  if (g.m <= 0 || g.n <= 0 || g.batch <= 0) return AFQ_OK;
  dim3 grid(cdiv_int(g.m, BM), cdiv_int(g.n, BN), g.batch);
  if (grid.y > 65535 || grid.z > 65535) return fail(AFQ_EUNSUPPORTED, "gemm grid too large");
  if (g.k <= 0) {
    // empty contraction: acc = 0
  }
  if (g.a_cplx && g.b_cplx) hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(NT), 0, stream, g);
  else if (g.a_cplx) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(NT), 0, stream, g);
  else if (g.b_cplx) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(NT), 0, stream, g);
  else hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(NT), 0, stream, g);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace afq
