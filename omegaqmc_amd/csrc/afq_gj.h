// Wavefront-level dense kernels shared by the single- and multi-determinant
// Green's functions: in-register Gauss-Jordan inversion with implicit row
// pivoting, and Ghalf = S^-1 phi^T on FP64 MFMA by one wavefront.
#pragma once
#include "afq_common.h"

namespace afq {
namespace gj {

typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma64(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double cabs1(cplx a) { return fabs(a.x) + fabs(a.y); }

// 64-bit DPP move (lanes the control does not write keep their value)
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ unsigned long long dpp64(unsigned long long x) {
  const int lo = __builtin_amdgcn_update_dpp((int)x, (int)x, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(x >> 32), (int)(x >> 32), CTRL, ROWS, 0xF, false);
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) {
  return a > b ? a : b;
}

// value of lane `lane` (uniform) of a complex register, as scalars
__device__ __forceinline__ double rl64(double v, int lane) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ cplx rlc(cplx v, int lane) { return mkc(rl64(v.x, lane), rl64(v.y, lane)); }

// Orders the LDS accesses of one wavefront (lanes exchange data through LDS).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int GJ_NMAX = 40;
constexpr int GJ_LD = GJ_NMAX + 1;

// ----------------------------------------------------------------------------
// Register Gauss-Jordan path (identity trials, N <= 40): one wavefront per
// spin, lane i holds row i of S in registers.  Step k picks the pivot row p_k
// of column k among the rows not used yet (first maximal |re| + |im|, as
// getrf, by a DPP max-reduction over one 64-bit key per lane) and eliminates column k from every other row in place (in-place
// Gauss-Jordan inversion with the row permutation kept implicit): afterwards
// row p_k holds row k of S^-1, whose column j sits in the storage column of
// column p_j.  det S = sign(p) prod_k d_k with d_k the pivots, accumulated as
// a mantissa-exponent pair so that no partial product over- or underflows.
//
// The row is kept rotated so that the active column is always register 0:
// the update of step k writes column j+1 into register j, and the new column
// k (the pivot row's 1/d, the others' -f/d) goes to the last register.  Every
// register index is static, so the step is straight-line VALU code.  With
// g = f/d for the other rows and g = 1 - 1/d for the pivot row, the single
// update a_j <- a_j - g * (unscaled pivot row)_j covers both.  The pivot row
// reaches the other lanes by readlane (scalar operands of the update), not
// through LDS: one-lane LDS writes and broadcast reads cost a full LDS
// instruction each, and the eight wavefronts of a CU would queue on them.
//
// DET: the determinant only (LU form): step k updates just the N - k - 1
// columns right of the pivot.  The rows still to be pivoted see exactly the
// operations of the full elimination, so the pivots and the determinant are
// bit-identical to it; the other rows and columns are left stale.
// ----------------------------------------------------------------------------
template <int NB, bool DET = false>
__device__ __forceinline__ void wave_gj(cplx (&a)[NB], int N, int* order,
                                        int& mystep, cplx* ov) {
  const int l = threadIdx.x & 63;
  bool active = l < N;
  mystep = 0;
  // det = sign(p) * pr * 2^e2, pr kept in [0.5, 1) in its larger component
  cplx pr = mkc(1.0, 0.0);
  int e2 = 0;
  int parity = 0;
  bool sing = false;  // an exact zero pivot: det 0 (slogdet sign 0) whatever follows
  for (int k = 0; k < N; ++k) {
    const cplx f = a[0];
    // pivot key: |re| + |im| with its low 6 mantissa bits replaced by 63 - lane
    // (the first of equal maxima wins); a NaN row ranks below every number
    // and is taken only if nothing else is left; finished rows rank last
    unsigned long long key = 0;
    if (active) {
      const double v = cabs1(f);
      key = (v != v) ? (1ull << 6)
                     : (((__builtin_bit_cast(unsigned long long, v) >> 6) + 2) << 6);
      key |= (unsigned long long)(63 - l);
    }
    key = umax64(key, dpp64<0xB1>(key));         // quad_perm [1, 0, 3, 2]
    key = umax64(key, dpp64<0x4E>(key));         // quad_perm [2, 3, 0, 1]
    key = umax64(key, dpp64<0x141>(key));        // half-row mirror
    key = umax64(key, dpp64<0x140>(key));        // row mirror
    key = umax64(key, dpp64<0x142, 0xA>(key));   // row_bcast15 into rows 1, 3
    key = umax64(key, dpp64<0x143, 0xC>(key));   // row_bcast31 into rows 2, 3
    const unsigned klo = __builtin_amdgcn_readlane((unsigned)key, 63);
    const int p = 63 - (int)(klo & 63u);
    const bool me = l == p;
    if (me) {
      active = false;
      mystep = k;
    }
    if (l == 0) order[k] = p;
    // sign of the pivot permutation from its Lehmer code
    const unsigned long long am = __ballot(active || me);
    parity ^= __popcll(am & ((1ull << p) - 1ull)) & 1;
    const cplx d = rlc(f, p);
    sing = sing || (d.x == 0.0 && d.y == 0.0);
    pr = cmul(pr, d);
    {
      int e;
      (void)frexp(fmax(fabs(pr.x), fabs(pr.y)), &e);
      pr = mkc(ldexp(pr.x, -e), ldexp(pr.y, -e));
      e2 += e;
    }
    // 1/d = conj(d) / |d|^2 with d first scaled by a power of two (exact), so
    // that |d|^2 neither over- nor underflows: one division
    cplx id;
    {
      int ed;
      (void)frexp(fmax(fabs(d.x), fabs(d.y)), &ed);
      const double sx = ldexp(d.x, -ed), sy = ldexp(d.y, -ed);
      const double q = 1.0 / fma(sx, sx, sy * sy);
      id = mkc(ldexp(sx * q, -ed), ldexp(-sy * q, -ed));
    }
    const cplx g = me ? mkc(1.0 - id.x, -id.y) : cmul(f, id);
    const cplx col = me ? id : mkc(-g.x, -g.y);
#pragma unroll
    for (int j = 0; j + 1 < NB; ++j) {
      if (!DET || j + 1 < N - k) {  // uniform; the loop stays unrolled
        cplx x = a[j + 1];
        cfms(x, g, rlc(a[j + 1], p));
        a[j] = x;
      }
    }
    if (!DET) a[NB - 1] = col;
  }
  wave_sync();  // order[] for gj_store_inverse
  if (l == 0) {
    const cplx r = sing ? mkc(0.0, 0.0) : mkc(ldexp(pr.x, e2), ldexp(pr.y, e2));
    *ov = parity ? mkc(-r.x, -r.y) : r;
  }
}

// Ghalf (N x M) = Inv (N x N) . phi^T on FP64 MFMA by one wavefront, N <= 40.
// Column tiles (16 orbitals) go in pairs with two static fragment buffers: the
// phi fragments of the next tile are loaded while the current one runs its
// N/16 row tiles (each phi fragment feeds every row tile).
template <int KQ, int TI>
__device__ __forceinline__ void ghalf_tile(const cplx (&ia)[TI][KQ], int N, int M, int tj,
                                           const cplx (&b)[KQ], cplx* __restrict__ Ghalf,
                                           bf16_t* __restrict__ gb, long Ks, long gplane,
                                           int ldq, bf16_t* sb) {
  const int l = threadIdx.x & 63;
  const int li = l & 15, lq = l >> 4;
  const int RTn = (N + 15) >> 4;
  const int q = 16 * tj + li;
#pragma unroll
  for (int ti = 0; ti < TI; ++ti) {
    if (ti >= RTn) break;  // uniform
    d4v cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < KQ; ++kk) {
      const cplx a = ia[ti][kk];
      cr = mfma64(a.x, b[kk].x, cr);
      ci = mfma64(a.x, b[kk].y, ci);
      cr = mfma64(-a.y, b[kk].y, cr);
      ci = mfma64(a.y, b[kk].x, ci);
    }
    if (Ghalf && q < M) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * ti + lq + 4 * r;
        if (i < N) Ghalf[(long)i * M + q] = mkc(cr[r], ci[r]);
      }
    }
    if (gb) {
      // split planes of the 16 x 16 tile through the wave's LDS staging
      // (sb[plane][row][col]), then 16-byte row chunks to global memory;
      // columns past M are written as the zeros the padding must hold
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int il = lq + 4 * r;
        bf16_t h0, l0, h1, l1;
        split2(q < M ? cr[r] : 0.0, h0, l0);
        split2(q < M ? ci[r] : 0.0, h1, l1);
        sb[(0 * 16 + il) * 16 + li] = h0;
        sb[(1 * 16 + il) * 16 + li] = l0;
        sb[(2 * 16 + il) * 16 + li] = h1;
        sb[(3 * 16 + il) * 16 + li] = l1;
      }
      wave_sync();
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int chunk = l + 64 * hh;  // (plane, row, half) of 8 bf16
        const int p = chunk >> 5, il = (chunk >> 1) & 15, half = chunk & 1;
        const int i = 16 * ti + il;
        if (i < N) {
          const long off = (p & 1 ? gplane : 0) + (p & 2 ? Ks : 0) + (long)i * ldq + 16 * tj +
                           8 * half;
          *(uint4*)(gb + off) = *(const uint4*)(sb + (p * 16 + il) * 16 + 8 * half);
        }
      }
      wave_sync();
    }
  }
}

template <int KQ>
__device__ __forceinline__ void ghalf_frag(const cplx* __restrict__ phi, int N, int M, int tj,
                                           cplx (&b)[KQ]) {
  const int l = threadIdx.x & 63;
  const int li = l & 15, lq = l >> 4;
  const int bj = min(16 * tj + li, M - 1);  // rows past M only feed unused outputs
#pragma unroll
  for (int kk = 0; kk < KQ; ++kk) {
    const int k = 4 * kk + lq;
    b[kk] = k < N ? phi[(long)bj * N + k] : mkc(0.0, 0.0);
  }
}

// The S^-1 fragments of every row tile (row 16 ti + li, column 4 kk + lq) are
// read once (LDS or global memory) and held in registers for all column
// tiles tj0, tj0 + tstep, ...  DB: the phi fragments of the next tile are in
// flight while this one runs (two fragment buffers; without, the registers
// leave room for a second wavefront per SIMD to cover the loads).  gb (split
// planes) needs sb: 2048 bytes of LDS per wavefront.
template <int KQ, bool DB = true>
__device__ __forceinline__ void ghalf_mfma_wave(const cplx* Inv, int LD, int N, int M,
                                                const cplx* __restrict__ phi,
                                                cplx* __restrict__ Ghalf, bf16_t* __restrict__ gb,
                                                long Ks, long gplane, int ldq, bf16_t* sb,
                                                int tj0 = 0, int tstep = 1) {
  constexpr int TI = (4 * KQ + 15) / 16;
  const int l = threadIdx.x & 63;
  const int li = l & 15, lq = l >> 4;
  cplx ia[TI][KQ];
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int kk = 0; kk < KQ; ++kk) {
      const int ai = 16 * ti + li, k = 4 * kk + lq;
      ia[ti][kk] = (ai < N && k < N) ? Inv[ai * LD + k] : mkc(0.0, 0.0);
    }
  const int RTm = (M + 15) >> 4;
  if (!DB) {
    cplx b[KQ];
    for (int tj = tj0; tj < RTm; tj += tstep) {
      ghalf_frag<KQ>(phi, N, M, tj, b);
      ghalf_tile<KQ, TI>(ia, N, M, tj, b, Ghalf, gb, Ks, gplane, ldq, sb);
    }
    return;
  }
  cplx b0[KQ], b1[KQ];
  if (tj0 >= RTm) return;
  ghalf_frag<KQ>(phi, N, M, tj0, b0);
  for (int tj = tj0; tj < RTm; tj += 2 * tstep) {
    if (tj + tstep < RTm) ghalf_frag<KQ>(phi, N, M, tj + tstep, b1);
    ghalf_tile<KQ, TI>(ia, N, M, tj, b0, Ghalf, gb, Ks, gplane, ldq, sb);
    if (tj + tstep >= RTm) break;
    if (tj + 2 * tstep < RTm) ghalf_frag<KQ>(phi, N, M, tj + 2 * tstep, b0);
    ghalf_tile<KQ, TI>(ia, N, M, tj + tstep, b1, Ghalf, gb, Ks, gplane, ldq, sb);
  }
}

// Scatter the Gauss-Jordan result (registers, rotated N times) into S^-1
// (row stride ld, in LDS or global memory): row p_k of the result is row k of
// S^-1; storage column c (now in register (c - N) mod NB) is column p_c.
template <int NB>
__device__ __forceinline__ void gj_store_inverse(const cplx (&a)[NB], int N, int mystep,
                                                 const int* order, cplx* Inv, int ld = GJ_LD) {
  const int l = threadIdx.x & 63;
  if (l < N) {
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      int c = q + N;
      if (c >= NB) c -= NB;
      if (c < N) Inv[(long)mystep * ld + order[c]] = a[q];
    }
  }
  wave_sync();
}

}  // namespace gj
}  // namespace afq
