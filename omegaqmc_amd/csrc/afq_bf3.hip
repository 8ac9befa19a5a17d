// Split-BF16 kernels of the reduced-precision step (v_mfma_f32_16x16x32_bf16).
//
// Every FP64 operand x is stored as two bf16 planes, hi = bf16(x) and
// lo = bf16(x - hi) (round to nearest even), and a product is accumulated in
// FP32 as ah*bh + ah*bl + al*bh; the dropped al*bl term is ~2^-16 of the
// product.  The walkers, determinants, QR, weights and energy reductions
// stay in FP64.
//
//   k_bf3_gemm<VHS>   VHS = i sqrt(dt) sum_g x_g L_g over the upper-triangle
//                     16x16 blocks of the symmetric result; each workgroup
//                     owns one block for 128 walkers and writes it and its
//                     transpose in the block-major layout the propagator reads.
//   k_bf3_gemm<FB>    split-K force bias, FP32 partials summed in FP64 later
//                     (Lhalf . Ghalf, or packed L . packed G for multi-det
//                     trials).  Both GEMMs stage their operands by LDS DMA.
//   k_prop_bf3        expH1 . sum_n VHS^n/n! . expH1 (Horner form) on a 48-column
//                     block of one spin of one walker, the block kept in LDS as
//                     bf16 planes; each one-body half step as y + (expH1 - I) y
//                     with y in the FP32 accumulators, so the fixed rounding of
//                     expH1's near-1 diagonal does not compound step after step.
//   k_exch_bf3        exchange energy sum_ij T_ij T_ji, T = Lhalf_g . Ghalf^T,
//                     T kept in registers and LDS only.
//
// MFMA layout (16x16x32 bf16, checked on the device): A[i][k] in lane
// i + 16 (k / 8), element k % 8; B[k][j] in lane j + 16 (k / 8); D[i][j] in lane
// j + 16 (i / 4), register i % 4.  With GEMM rows ordered (walker, re/im) a
// lane's four output rows are (w, re), (w, im), (w+1, re), (w+1, im).
#include <type_traits>

#include "afq_bf3.h"
#include "afq_kernels.h"

namespace afq {

namespace {

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v mfma_bf(const uint4& a, const uint4& b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a),
                                                 __builtin_bit_cast(bf8v, b), c, 0, 0, 0);
}

// c += ah*bh + ah*bl + al*bh
__device__ __forceinline__ f4v mfma3(const uint4& ah, const uint4& al, const uint4& bh,
                                     const uint4& bl, f4v c) {
  c = mfma_bf(al, bh, c);
  c = mfma_bf(ah, bl, c);
  return mfma_bf(ah, bh, c);
}

typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

// bf16 pair (round to nearest even) packed in 32 bits, element a low
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f2v{a, b}, bf2v));
}
// four FP32 values -> hi = bf16(x), lo = bf16(x - hi) (x - hi is exact in FP32)
__device__ __forceinline__ void split4(f4v v, uint2& hi, uint2& lo) {
  hi.x = pk_bf16(v[0], v[1]);
  hi.y = pk_bf16(v[2], v[3]);
  const float r0 = v[0] - __uint_as_float(hi.x << 16);
  const float r1 = v[1] - __uint_as_float(hi.x & 0xffff0000u);
  const float r2 = v[2] - __uint_as_float(hi.y << 16);
  const float r3 = v[3] - __uint_as_float(hi.y & 0xffff0000u);
  lo.x = pk_bf16(r0, r1);
  lo.y = pk_bf16(r2, r3);
}

// 16-byte chunk position of K chunk c in row li of a 16-row block (rows of 64
// bytes).  Measured on gfx950 with a ds_read_b128 probe: reading one chunk of
// all 16 rows takes 4 LDS cycles per wave with this XOR, against 8 with no
// swizzle or with c ^ ((li >> 1) ^ (li >> 3)) & 3, although every one of
// these maps the 16 lanes of a quarter-wave to distinct 16-byte bank groups
__device__ __forceinline__ int xswz(int li, int c) { return c ^ ((li >> 1) & 3); }

__device__ __forceinline__ uint4 neg8(uint4 a) {
  return make_uint4(a.x ^ 0x80008000u, a.y ^ 0x80008000u, a.z ^ 0x80008000u, a.w ^ 0x80008000u);
}

__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int x = b & 7, slot = b >> 3;
  const int per = nb >> 3, rem = nb & 7;
  return (x < rem) ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}

// ----------------------------------------------------------------------------
// system preparation
// ----------------------------------------------------------------------------
__global__ void k_prep_L(long n, int M, int G, int Gp, long NC, const int2* __restrict__ pq,
                         const double* __restrict__ L, bf16_t* __restrict__ Lb) {
  const long MM = (long)M * M;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long c = e / Gp;
    const int g = (int)(e % Gp);
    const int b = (int)(c >> 8), i = (int)((c >> 4) & 15), j = (int)(c & 15);
    const int2 PQ = pq[b];
    const int p = 16 * PQ.x + i, q = 16 * PQ.y + j;
    double v = 0.0;
    if (g < G && p < M && q < M) v = L[(long)g * MM + (long)p * M + q];
    bf16_t h, l;
    split2(v, h, l);
    const long o = ((c >> 4) * (Gp >> 5) + (g >> 5)) * 512 + (c & 15) * 32 + (g & 31);
    Lb[o] = h;
    Lb[NC * Gp + o] = l;
  }
}

// rows x cols -> [plane][rows_p][cols_p] with zero padding; diag_sub is
// taken off the diagonal first (A - diag_sub I).  tiled: K-tiled planes,
// (r, c) at ((r / 16) (cols_p / 32) + c / 32) 512 + (r % 16) 32 + c % 32
// (rows_p a multiple of 16, cols_p of 32)
__global__ void k_prep_mat(long n, int rows, int cols, int cols_p, long plane,
                           const double* __restrict__ A, double diag_sub,
                           bf16_t* __restrict__ out, int tiled) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long r = e / cols_p;
    const int c = (int)(e % cols_p);
    const double v = (r < rows && c < cols) ? A[r * cols + c] - (r == c ? diag_sub : 0.0) : 0.0;
    bf16_t h, l;
    split2(v, h, l);
    const long o = tiled ? ((r >> 4) * (cols_p >> 5) + (c >> 5)) * 512 + (r & 15) * 32 + (c & 31) : e;
    out[o] = h;
    out[plane + o] = l;
  }
}

// Lhalf (G, N, M) -> [plane][G][K] with element (g, i, q) at g K + i ldq + q
__global__ void k_prep_lh(long n, int N, int M, int ldq, long K, long plane,
                          const double* __restrict__ R, bf16_t* __restrict__ out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long g = e / K;
    const int c = (int)(e % K);
    const int i = c / ldq, q = c % ldq;
    const double v = (i < N && q < M) ? R[(g * N + i) * M + q] : 0.0;
    bf16_t h, l;
    split2(v, h, l);
    out[e] = h;
    out[plane + e] = l;
  }
}

typedef int i32x4 __attribute__((ext_vector_type(4)));
// buffer load of 16 bytes per lane straight into LDS: lane l's bytes land at
// lds + 16 l (lds uniform across the wavefront).  Written as inline asm so
// that the compiler does not see an LDS store: for the intrinsic form it
// waits for every outstanding DMA before the next LDS read, which would
// serialise the prefetch of the next K step with the reads of this one.  The
// caller orders the DMA with its own vmcnt wait and barrier.  No VGPR is
// written; the operands are read when the instruction issues.  M0 holds the
// LDS base and needs one wait state before the load.
__device__ __forceinline__ void dma16(i32x4 rsrc, unsigned lds, int voffset, int soffset) {
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tbuffer_load_dwordx4 %1, %2, %3 offen lds"
               :
               : "s"(__builtin_amdgcn_readfirstlane(lds)), "v"(voffset), "s"(rsrc), "s"(soffset)
               : "memory", "m0");
}

// buffer resource over `bytes` bytes at p (raw buffer, 32-bit offsets)
__device__ __forceinline__ i32x4 buf_rsrc(const void* p, unsigned bytes) {
  const unsigned long long ad = (unsigned long long)p;
  i32x4 r = {(int)(unsigned)ad, (int)((ad >> 32) & 0xffff), (int)bytes, 0x00020000};
  r.x = __builtin_amdgcn_readfirstlane(r.x);
  r.y = __builtin_amdgcn_readfirstlane(r.y);
  r.z = __builtin_amdgcn_readfirstlane(r.z);
  return r;
}

// LDS writes of this wavefront visible to its own later reads (no other wave)
__device__ __forceinline__ void xwave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ----------------------------------------------------------------------------
// the GEMM core: 256 x 256 tile, K step 32, 8 waves of 64 x 128
// ----------------------------------------------------------------------------
constexpr int GB = 256, GK = 32, GT = 512;

// Operand planes are row-major (element (r, k) at r ld + k) or K-tiled: rows
// in blocks of 16, each block's K in steps of 32, every 16 x 32 piece 1 KB of
// contiguous memory, (r, k) at ((r / 16) (ld / 32) + k / 32) 512 + (r % 16) 32
// + k % 32 (ld a multiple of 32, rows padded to 16).  A K-tiled K step of 16
// rows is one contiguous kilobyte for the DMA: 42.6 against 23.8 B per clock
// per CU for 16 separate 64-byte rows, measured on gfx950.
struct Bf3Split {
  const bf16_t* a[2];  // A planes (rows x lda)
  const bf16_t* b[2];  // B planes (cols x ldb)
  long lda, ldb;
  int kbeg, kend;
  unsigned abytes, bbytes;  // bytes of one A / B plane (< 2^32)
  int tiled;                // both operands K-tiled
};

struct Bf3Gemm {
  Bf3Split sp[8];
  int nsplit;
  int rows, cols, ntm, ntn;
  // VHS epilogue
  const int2* blk_pq;
  int RT, W;
  long vstride;
  bf16_t* Vb;
  float sqdt;
  // FB epilogue
  float2* part;
  int G;
};

template <int EPI>
__global__ __launch_bounds__(GT) void k_bf3_gemm(Bf3Gemm a) {
  extern __shared__ __align__(16) bf16_t lds[];
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int wr = wv >> 1, wc = wv & 1;
  const int li = l & 15, lq = l >> 4;
  const int v = xcd_remap(blockIdx.x, a.ntm * a.ntn);
  const int tm = v % a.ntm, tn = v / a.ntm;
  const int row0 = tm * GB, col0 = tn * GB;
  const Bf3Split& S = a.sp[blockIdx.y];

  // two stage buffers: the next K step lands in one while this one is
  // consumed; rows of GK with the 16-byte chunks XOR-swizzled within 16-row
  // blocks (conflict-free MFMA operand reads)
  constexpr int GSTAGE = 4 * GB * GK;

  // Operand chunks go from global memory straight into LDS by buffer-load
  // DMA (no staging registers): per stage 64 instructions of 16 rows each,
  // 16 blocks of 16 rows in each of the planes A hi, A lo, B hi, B lo.  Lane l
  // fetches row l / 4 of its block, K chunk xswz(l / 4, l % 4), so the chunk
  // lands at its swizzled position l % 4.  A row or column outside the matrix
  // reads row 0: it only feeds output rows or columns that the epilogue
  // discards.
  const i32x4 rs[4] = {buf_rsrc(S.a[0], S.abytes), buf_rsrc(S.a[1], S.abytes),
                       buf_rsrc(S.b[0], S.bbytes), buf_rsrc(S.b[1], S.bbytes)};
  const int lrow = l >> 2, lkc = xswz(lrow, l & 3);
  unsigned off[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    // instruction i of this wave: operand plane q = i / 2 (A hi, A lo, B hi,
    // B lo), 16-row block b = wv + 8 (i % 2)
    const int q = i >> 1, b = wv + 8 * (i & 1);
    const int r = 16 * b + lrow;
    const bool isa = q < 2;
    const int gr = (isa ? row0 : col0) + r;
    const bool ok = gr < (isa ? a.rows : a.cols);
    const long ld = isa ? S.lda : S.ldb;
    // K-tiled: row block gr / 16 is 32 ld bytes, the lane's row (gr % 16 = lrow)
    // 64 bytes into each kilobyte
    off[i] = S.tiled ? (unsigned)((ok ? (long)(gr >> 4) : 0L) * ld * 32 + lrow * 64 + 16 * lkc)
                     : (unsigned)((ok ? (long)gr : 0L) * ld * 2 + 16 * lkc);
  }
  const int kbytes = S.tiled ? 32 : 2;  // byte offset of K step k0 per unit of k0
  auto fetch1 = [&](int buf, int k0, int i) {
    const int q = i >> 1, b = wv + 8 * (i & 1);
    const unsigned dst =
        (unsigned)(uintptr_t)(lds + buf * GSTAGE + q * GB * GK + b * 16 * GK);
    dma16(rs[q], dst, (int)off[i], kbytes * k0);
  };
  auto dma_wait = []() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); };
  const int frag = li * GK + 8 * xswz(li, lq);  // (row li, chunk lq) of a 16-row block

  f4v acc[4][8];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int w = 0; w < 8; ++w) acc[u][w] = f4v{0.f, 0.f, 0.f, 0.f};

  const int nk = (S.kend - S.kbeg) / GK;
  if (nk > 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) fetch1(0, S.kbeg, i);
    dma_wait();
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    const bf16_t* st = lds + buf * GSTAGE + frag;
    // the next K step into the other buffer (last read before the previous
    // barrier), all of it at the start so that every chunk has the whole
    // step's products to land in
    if (kt + 1 < nk) {
#pragma unroll
      for (int i = 0; i < 8; ++i) fetch1(buf ^ 1, S.kbeg + (kt + 1) * GK, i);
    }
    uint4 ah[4], al[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = 64 * wr + 16 * u;  // block-aligned row
      ah[u] = *(const uint4*)(st + r * GK);
      al[u] = *(const uint4*)(st + (GB + r) * GK);
    }
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      const int c = 128 * wc + 16 * w;
      const uint4 bh = *(const uint4*)(st + (2 * GB + c) * GK);
      const uint4 bl = *(const uint4*)(st + (3 * GB + c) * GK);
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u][w] = mfma3(ah[u], al[u], bh, bl, acc[u][w]);
    }
    dma_wait();
    __syncthreads();
  }

  if (EPI == 0) {
    // one 16x16 block (P,Q) of VHS per column tile; this wave holds block
    // rows i = 8 wc + x (x < 8), column j = li, for the walkers of its rows.
    // The transposed block (Q,P) takes a lane's 8 rows as one 16-byte store;
    // the block (P,Q) itself goes through a wave-private LDS area
    // [lq][plane][x][j] (the stages are free after the K loop), read back as
    // 16-byte row chunks: 4 stores of 16 bytes per lane instead of 32 of 2.
    const int2 PQ = a.blk_pq[tn];
    const long offd = ((long)PQ.x * a.RT + PQ.y) * 1024;
    const long offt = ((long)PQ.y * a.RT + PQ.x) * 1024;
    bf16_t* stg = lds + wv * 2048;  // 4 walkers x 4 planes x 8 rows x 16 columns
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int w = (row0 + 64 * wr + 16 * u + 4 * lq) / 2 + h;
        if (w < a.W) {
          bf16_t* base = a.Vb + (long)w * a.vstride;
          bf16_t pl[4][8];
#pragma unroll
          for (int x = 0; x < 8; ++x) {
            // FP32 product (the nearest float to the exact one) and its exact
            // FP32 split
            const float re = -a.sqdt * acc[u][x][2 * h + 1];
            const float im = a.sqdt * acc[u][x][2 * h];
            pl[0][x] = bf_bits(re);
            pl[1][x] = bf_bits(re - bf_val(pl[0][x]));
            pl[2][x] = bf_bits(im);
            pl[3][x] = bf_bits(im - bf_val(pl[2][x]));
#pragma unroll
            for (int p = 0; p < 4; ++p) stg[((lq * 4 + p) * 8 + x) * 16 + li] = pl[p][x];
          }
          if (PQ.x != PQ.y) {
#pragma unroll
            for (int p = 0; p < 4; ++p) {
              uint4 pk;
              pk.x = pl[p][0] | ((unsigned)pl[p][1] << 16);
              pk.y = pl[p][2] | ((unsigned)pl[p][3] << 16);
              pk.z = pl[p][4] | ((unsigned)pl[p][5] << 16);
              pk.w = pl[p][6] | ((unsigned)pl[p][7] << 16);
              *(uint4*)(base + offt + p * 256 + li * 16 + 8 * wc) = pk;
            }
          }
        }
        xwave_sync();
        // chunk c = ((q * 4 + p) * 8 + x) * 2 + half of the area: walker
        // group q, plane p, block row 8 wc + x, columns 8 half .. 8 half + 7
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int c = l + 64 * k;
          const int q = c >> 6, p = (c >> 4) & 3, x = (c >> 1) & 7, half = c & 1;
          const int wq = (row0 + 64 * wr + 16 * u + 4 * q) / 2 + h;
          const uint4 v = *(const uint4*)(stg + c * 8);
          if (wq < a.W)
            *(uint4*)(a.Vb + (long)wq * a.vstride + offd + p * 256 + (8 * wc + x) * 16 + 8 * half) = v;
        }
        xwave_sync();
      }
    }
  } else {
    const long soff = (long)blockIdx.y * a.W * a.G;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int w = (row0 + 64 * wr + 16 * u + 4 * lq) / 2 + h;
        if (w >= a.W) continue;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
          const int g = col0 + 128 * wc + 16 * x + li;
          if (g < a.G) a.part[soff + (long)w * a.G + g] = make_float2(acc[u][x][2 * h], acc[u][x][2 * h + 1]);
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------
// fused propagator on a 48-column block of one spin
// ----------------------------------------------------------------------------
// The Taylor sum runs in Horner form, y <- T0 + V y / n for n = nmax..1 with
// y = T0 = expH1 phi to start (the same polynomial as afqmc_gto.py:451-469):
// T0 = phi + D phi with D = expH1 - I.
// only T0 has to be kept besides the current y, and it waits in a global
// scratch in the accumulators' own lane order (t0s, RT * 3 * 64 * 8 floats per
// workgroup), so the registers hold one accumulator set and two waves fit
// per SIMD.
constexpr int PW = 8, PT = 64 * PW, PC = 48;
// Horner stage n (y <- T0 + V y / n) reaches the result through n - 1 more
// products by V / k, so a relative error e in its product V y moves phi by
// about e |V|^n / n!.  Stages n >= TAYLOR_BF16_FROM use plain bf16 products
// (e ~ 2^-8, hi planes only): at |V|_2 = 0.5 (M=300, Nchol=1800, dt=0.005)
// that is 8e-7 (n = 5) and 7e-8 (n = 6) of phi, below the ~4e-6 of the split
// products of the first stage.
constexpr int TAYLOR_BF16_FROM = 5;

// A walker's columns are the concatenation of its two spins (column cg < Na
// of phia, then Nb of phib), cut into nblk blocks of equal width bs <= PC =
// 48: with N = 40 + 40 two blocks of 40, with 20 + 18 one of 38.  Equal
// blocks keep a walker's workgroups in step, so the second one finds the
// walker's VHS rows in L2.  A block's 16-column tiles past its last column
// (x >= nxl) are skipped, a uniform branch per workgroup.
template <int MT>
__global__ __launch_bounds__(PT) void k_prop_bf3(int W, int M, int RT, int Kp, int Na, int Nb,
                                                 int nblk, int bs, const bf16_t* __restrict__ E,
                                                 const bf16_t* __restrict__ Vb,
                                                 float4* __restrict__ t0s, double* phia,
                                                 double* phib, int nmax) {
  extern __shared__ __align__(16) bf16_t Ts[];  // [4][PC][Kp + 8]
  const int KR = Kp + 8;
  const int vb = xcd_remap(blockIdx.x, W * nblk);
  const int w = vb / nblk, cbx = vb % nblk;
  const int c0 = cbx * bs;
  const int Nt = Na + Nb;
  const int ncol = min(bs, Nt - c0);            // columns of this block
  const int nxl = min(3, (ncol + 15) >> 4);     // live 16-column tiles (uniform)
  // (spin array, column, row stride) of block column c; nullptr past the block
  auto colmap = [&](int c, int* col, int* ld) -> double* {
    const int cg = c0 + c;
    *col = 0; *ld = 1;
    if (c >= ncol) return nullptr;
    if (cg < Na) { *col = cg; *ld = Na; return phia; }
    *col = cg - Na; *ld = Nb;
    return phib;
  };
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int li = l & 15, lq = l >> 4;
  const long MM = (long)M * M;
  (void)MM;
  const bf16_t* Vw = Vb + (long)w * RT * RT * 1024;
  const long Eplane = (long)RT * 16 * Kp;

  auto T = [&](int p, int c) { return Ts + (p * PC + c) * KR; };

  // stage phi (FP64) as four bf16 planes, zero outside M x N: an item is a
  // row pair (2 rp, 2 rp + 1) of one column, written as one 32-bit word per
  // plane; SB items per thread have their loads in flight together (clamped
  // in-bounds addresses, zeroed after the load)
  {
    typedef double d2v __attribute__((ext_vector_type(2)));
    constexpr int SB = 5;
    const int np = PC * (Kp / 2);
    for (int e0 = t; e0 < np; e0 += PT * SB) {
      d2v v[SB][2];
#pragma unroll
      for (int k = 0; k < SB; ++k) {
        const int e = e0 + PT * k;
        const int c = e % PC, rp = e / PC;
        int col, ld;
        const double* base = colmap(min(c, ncol - 1), &col, &ld);
        const d2v* pw = (const d2v*)base + (long)w * M * ld + col;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int r = 2 * rp + h;
          const d2v x = pw[(long)min(r, M - 1) * ld];
          const bool ok = e < np && r < M && c < ncol;
          v[k][h] = ok ? x : d2v{0.0, 0.0};
        }
      }
#pragma unroll
      for (int k = 0; k < SB; ++k) {
        const int e = e0 + PT * k;
        if (e >= np) continue;
        const int c = e % PC, rp = e / PC;
        bf16_t q[4][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          split2(v[k][h].x, q[0][h], q[1][h]);
          split2(v[k][h].y, q[2][h], q[3][h]);
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
          *(unsigned*)(T(p, c) + 2 * rp) = q[p][0] | ((unsigned)q[p][1] << 16);
      }
    }
  }
  __syncthreads();

  f4v are[MT][3], aim[MT][3];
  float4* T0 = t0s + (long)(w * nblk + cbx) * RT * 3 * 64 * 2;
  auto t0_at = [&](int u, int x) { return T0 + ((long)((wv + PW * u) * 3 + x) * 64 + l) * 2; };
  auto owns = [&](int u) { return wv + PW * u < RT; };
  auto zero = [&]() {
#pragma unroll
    for (int u = 0; u < MT; ++u)
#pragma unroll
      for (int x = 0; x < 3; ++x) { are[u][x] = f4v{0, 0, 0, 0}; aim[u][x] = f4v{0, 0, 0, 0}; }
  };
  auto bfrag = [&](int x, int ks, uint4* b) {
    if (x >= nxl) return;
#pragma unroll
    for (int p = 0; p < 4; ++p) b[p] = *(const uint4*)(T(p, 16 * x + li) + 32 * ks + 8 * lq);
  };
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  typedef __attribute__((address_space(1))) const v4u gv4u;
  auto gld = [](const bf16_t* p) {
    const v4u x = *(const gv4u*)p;
    return make_uint4(x.x, x.y, x.z, x.w);
  };
  // acc = E . T (E real, K-tiled planes [2][RT][Kp / 32][16][32]: a wave's
  // fragment of one row tile and k-step is one contiguous kilobyte); row
  // tiles this wave does not own read a valid row and are not used
  auto ldE = [&](uint4* eh, uint4* el, int ks) {
#pragma unroll
    for (int u = 0; u < MT; ++u) {
      const long o = ((long)min(wv + PW * u, RT - 1) * (Kp >> 5) + ks) * 512 + li * 32 + 8 * lq;
      eh[u] = gld(E + o);
      el[u] = gld(E + Eplane + o);
    }
  };
  auto estep = [&](const uint4* eh, const uint4* el, int ks) {
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      if (x >= nxl) continue;
      uint4 b[4];
      bfrag(x, ks, b);
#pragma unroll
      for (int u = 0; u < MT; ++u) {
        if (!owns(u)) continue;
        are[u][x] = mfma3(eh[u], el[u], b[0], b[1], are[u][x]);
        aim[u][x] = mfma3(eh[u], el[u], b[2], b[3], aim[u][x]);
      }
    }
  };
  // acc <- phi (FP64 -> FP32) on the owned tiles, zero outside M x N: the
  // identity part of the first one-body half step
  auto load_phi = [&]() {
#pragma unroll
    for (int u = 0; u < MT; ++u) {
      const int r0 = 16 * min(wv + PW * u, RT - 1) + 4 * lq;
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        int col, ld;
        const double* base = colmap(16 * x + li, &col, &ld);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const bool ok = owns(u) && r0 + r < M && base != nullptr;
          const long i = ok ? ((long)w * M + r0 + r) * ld + col : 0;
          const double2 z = ok ? *(const double2*)(base + 2 * i) : make_double2(0.0, 0.0);
          are[u][x][r] = (float)z.x;
          aim[u][x][r] = (float)z.y;
        }
      }
    }
  };
  // acc += D . T with D = expH1 - I, so that a one-body half step is
  // acc = y + D y when acc enters holding y (FP32): the identity part never
  // passes through the BF16 split.  k-steps in pairs with two static
  // fragment buffers: the next k-step's D rows are in flight while this one
  // runs.
  auto onebody = [&]() {
    const int nks = Kp / 32;
    uint4 e0h[MT], e0l[MT], e1h[MT], e1l[MT];
    ldE(e0h, e0l, 0);
    for (int ks = 0; ks < nks; ks += 2) {
      ldE(e1h, e1l, ks + 1 < nks ? ks + 1 : nks - 1);
      estep(e0h, e0l, ks);
      if (ks + 1 >= nks) break;
      ldE(e0h, e0l, ks + 2 < nks ? ks + 2 : nks - 1);
      estep(e1h, e1l, ks + 1);
    }
  };
  // acc = V . T (V complex, block-major planes).  VHS fragments of row tile u
  // at k-step ks; indices are clamped into the matrix: a clamped row tile is
  // not owned, and a clamped column block Q = RT meets zero rows of T.
  auto ldv = [&](uint4* v, int u, int ks) {
    const int Q = min(2 * ks + (lq >> 1), RT - 1), j0 = 8 * (lq & 1);
    const int P = min(wv + PW * u, RT - 1);
    const bf16_t* blk = Vw + ((long)P * RT + Q) * 1024 + li * 16 + j0;
#pragma unroll
    for (int p = 0; p < 4; ++p) v[p] = gld(blk + p * 256);
  };
  // plain bf16 (hi planes only) for the high-order Horner stages
  auto ldv_hi = [&](uint4* v, int u, int ks) {
    const int Q = min(2 * ks + (lq >> 1), RT - 1), j0 = 8 * (lq & 1);
    const int P = min(wv + PW * u, RT - 1);
    const bf16_t* blk = Vw + ((long)P * RT + Q) * 1024 + li * 16 + j0;
    v[0] = gld(blk);
    v[2] = gld(blk + 512);
  };
  auto bfrag_hi = [&](int x, int ks, uint4* b) {
    if (x >= nxl) return;
    b[0] = *(const uint4*)(T(0, 16 * x + li) + 32 * ks + 8 * lq);
    b[2] = *(const uint4*)(T(2, 16 * x + li) + 32 * ks + 8 * lq);
  };
  auto vstep_hi = [&](const uint4* vc, uint4 (*b)[4], int u) {
    if (owns(u)) {
      const uint4 ni = neg8(vc[2]);
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        if (x >= nxl) continue;
        are[u][x] = mfma_bf(vc[0], b[x][0], are[u][x]);
        are[u][x] = mfma_bf(ni, b[x][2], are[u][x]);
        aim[u][x] = mfma_bf(vc[0], b[x][2], aim[u][x]);
        aim[u][x] = mfma_bf(vc[2], b[x][0], aim[u][x]);
      }
    }
  };
  auto vstep = [&](const uint4* vc, uint4 (*b)[4], int u) {
    if (owns(u)) {
      const uint4 ni = neg8(vc[2]), nl = neg8(vc[3]);
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        if (x >= nxl) continue;
        are[u][x] = mfma3(vc[0], vc[1], b[x][0], b[x][1], are[u][x]);
        are[u][x] = mfma3(ni, nl, b[x][2], b[x][3], are[u][x]);
        aim[u][x] = mfma3(vc[0], vc[1], b[x][2], b[x][3], aim[u][x]);
        aim[u][x] = mfma3(vc[2], vc[3], b[x][0], b[x][1], aim[u][x]);
      }
    }
  };
  // Row tiles are walked one at a time with the next tile's fragments in
  // flight.  Two fragment buffers alternate by the parity of the sub-step
  // (row tile, k-step); the k-steps go in pairs so that, with the row-tile
  // loop unrolled, every buffer choice is static and no loaded register is
  // copied (a copy would wait for the load).
  // Row tiles are walked one at a time with the fragments of the tile after
  // next in flight.  Three fragment buffers rotate by sub-step (row tile,
  // k-step); the k-steps go in threes so that, with the row-tile loop
  // unrolled, every buffer choice is static and no loaded register is copied
  // (a copy would wait for the load).
  // LOW: plain bf16 products (the Horner stages n >= TAYLOR_BF16_FROM)
  auto term = [&](auto lowc) {
    constexpr bool LOW = decltype(lowc)::value;
    zero();
    const int nks = Kp / 32;
    uint4 vb[3][4];
    auto ld = [&](uint4* v, int u, int ks) {
      if constexpr (LOW) ldv_hi(v, u, ks);
      else ldv(v, u, ks);
    };
    ld(vb[0], 0, 0);
    ld(vb[1], 1 % MT, (1 / MT) < nks ? 1 / MT : nks - 1);
    for (int ks = 0; ks < nks; ks += 3) {
#pragma unroll
      for (int h = 0; h < 3; ++h) {
        const int kk = ks + h;
        const bool live = kk < nks;  // uniform
        uint4 b[3][4];
        if (live) {
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            if constexpr (LOW) bfrag_hi(x, kk, b[x]);
            else bfrag(x, kk, b[x]);
          }
        }
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int sidx = h * MT + u;
          const int t2 = sidx + 2;
          const int kn = ks + t2 / MT;
          ld(vb[t2 % 3], t2 % MT, kn < nks ? kn : nks - 1);
          if (live) {
            if constexpr (LOW) vstep_hi(vb[sidx % 3], b, u);
            else vstep(vb[sidx % 3], b, u);
          }
        }
      }
    }
  };
  // acc -> LDS planes (rows < M), four rows of a lane at a time, split in
  // FP32 (x - bf16(x) is exact in FP32: the same planes as the FP64 split)
  auto to_lds = [&](f4v (*xr)[3], f4v (*xi)[3]) {
#pragma unroll
    for (int u = 0; u < MT; ++u) {
      if (!owns(u)) continue;
      const int r0 = 16 * (wv + PW * u) + 4 * lq;
      const bool edge = 16 * (wv + PW * u) + 16 > M;  // uniform: the tile crosses row M
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        f4v vr = xr[u][x], vi = xi[u][x];
        if (edge) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const bool ok = r0 + r < M;
            vr[r] = ok ? vr[r] : 0.f;
            vi[r] = ok ? vi[r] : 0.f;
          }
        }
        uint2 hr, lr, hi, li2;
        split4(vr, hr, lr);
        split4(vi, hi, li2);
        *(uint2*)(T(0, 16 * x + li) + r0) = hr;
        *(uint2*)(T(1, 16 * x + li) + r0) = lr;
        *(uint2*)(T(2, 16 * x + li) + r0) = hi;
        *(uint2*)(T(3, 16 * x + li) + r0) = li2;
      }
    }
  };

  load_phi();
  onebody();
  __syncthreads();
  to_lds(are, aim);
#pragma unroll
  for (int u = 0; u < MT; ++u) {
    if (!owns(u)) continue;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      float4* d = t0_at(u, x);
      d[0] = make_float4(are[u][x][0], are[u][x][1], are[u][x][2], are[u][x][3]);
      d[1] = make_float4(aim[u][x][0], aim[u][x][1], aim[u][x][2], aim[u][x][3]);
    }
  }
  __syncthreads();
  for (int n = nmax; n >= 1; --n) {
    if (n >= TAYLOR_BF16_FROM) term(std::true_type{});
    else term(std::false_type{});
    const float rn = 1.0f / (float)n;
#pragma unroll
    for (int u = 0; u < MT; ++u) {
      if (!owns(u)) continue;
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        const float4* d = t0_at(u, x);
        const float4 tr = d[0], ti = d[1];
        are[u][x] = f4v{tr.x, tr.y, tr.z, tr.w} + are[u][x] * rn;
        aim[u][x] = f4v{ti.x, ti.y, ti.z, ti.w} + aim[u][x] * rn;
      }
    }
    __syncthreads();
    to_lds(are, aim);
    __syncthreads();
  }
  onebody();  // acc holds y_1: phi_out = y_1 + D y_1
#pragma unroll
  for (int u = 0; u < MT; ++u) {
    if (!owns(u)) continue;
    const int r0 = 16 * (wv + PW * u) + 4 * lq;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      int col, ld;
      double* base = colmap(16 * x + li, &col, &ld);
      if (!base) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r0 + r >= M) continue;
        double* z = base + 2 * (((long)w * M + r0 + r) * ld + col);
        *(double2*)z = make_double2((double)are[u][x][r], (double)aim[u][x][r]);
      }
    }
  }
}

}  // namespace

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------
size_t bf3_vhs_bytes(const Bf3System& b, int W) { return (size_t)W * b.RT * b.RT * 1024 * 2; }
static size_t x_rows(int W) { return (size_t)(2 * W + 15) / 16 * 16; }
size_t bf3_x_bytes(const Bf3System& b, int W) { return (size_t)2 * x_rows(W) * b.Gp * 2; }
size_t bf3_gh_bytes(const Bf3System& b, int W) {
  return (size_t)2 * 2 * W * ((size_t)b.Ka + b.Kb) * 2;
}

int bf3_prepare(afq_system* s, Bf3System* b) {
  hipStream_t st = s->ctx->stream;
  const int M = s->M, G = s->G;
  b->M = M; b->Na = s->Na; b->Nb = s->Nb; b->G = G;
  b->RT = (M + 15) / 16;
  b->Gp = (G + 31) / 32 * 32;
  b->nblk = b->RT * (b->RT + 1) / 2;
  b->Kp = (16 * b->RT + 31) / 32 * 32;
  b->ldq = (M + GK - 1) / GK * GK;  // whole K steps of the exchange GEMM
  b->Ka = (s->Na * b->ldq + 31) / 32 * 32;
  b->Kb = (s->Nb * b->ldq + 31) / 32 * 32;
  b->Lb = nullptr; b->blk_pq = nullptr; b->Lha = nullptr; b->Lhb = nullptr; b->E = nullptr;
  if (b->RT > 32) return fail(AFQ_EUNSUPPORTED, "bf16x3 propagator: M=%d > 512", M);
  const long NC = (long)b->nblk * 256;
  int2* hpq = new int2[b->nblk];
  int k = 0;
  for (int P = 0; P < b->RT; ++P)
    for (int Q = P; Q < b->RT; ++Q) hpq[k++] = make_int2(P, Q);
  hipError_t e = hipMalloc(&b->blk_pq, sizeof(int2) * b->nblk);
  if (e == hipSuccess) e = hipMemcpyAsync(b->blk_pq, hpq, sizeof(int2) * b->nblk, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  delete[] hpq;
  if (e != hipSuccess) return hip_fail(e, "bf16x3 block table");
  if (hipMalloc(&b->Lb, sizeof(bf16_t) * 2 * NC * b->Gp) != hipSuccess ||
      hipMalloc(&b->E, sizeof(bf16_t) * 2 * (size_t)16 * b->RT * b->Kp) != hipSuccess ||
      hipMalloc(&b->Lha, sizeof(bf16_t) * 2 * (size_t)G * b->Ka + 64) != hipSuccess ||
      hipMalloc(&b->Lhb, sizeof(bf16_t) * 2 * (size_t)G * b->Kb + 64) != hipSuccess)
    return fail(AFQ_ENOMEM, "bf16x3 operands");
  if (G > 0) {
    const long n = NC * b->Gp;
    hipLaunchKernelGGL(k_prep_L, dim3(8192), dim3(256), 0, st, n, M, G, b->Gp, NC, b->blk_pq,
                       s->chol, b->Lb);
    AFQ_CHECK_LAUNCH();
  }
  {
    const long rows = 16L * b->RT, n = rows * b->Kp;
    // the one-body half steps run as phi + (expH1 - I) phi: only the O(dt)
    // part D = expH1 - I goes through the split-BF16 products (exact in FP64:
    // the diagonal of expH1 lies within a factor 2 of 1)
    hipLaunchKernelGGL(k_prep_mat, dim3(1024), dim3(256), 0, st, n, M, M, b->Kp, n, s->expH1, 1.0,
                       b->E, 1);
    AFQ_CHECK_LAUNCH();
  }
  if (G > 0 && s->Na > 0 && s->rchol_a) {
    const long n = (long)G * b->Ka;
    hipLaunchKernelGGL(k_prep_lh, dim3(4096), dim3(256), 0, st, n, s->Na, M, b->ldq, (long)b->Ka,
                       n, s->rchol_a, b->Lha);
    AFQ_CHECK_LAUNCH();
  }
  if (G > 0 && s->Nb > 0 && s->rchol_b) {
    const long n = (long)G * b->Kb;
    hipLaunchKernelGGL(k_prep_lh, dim3(4096), dim3(256), 0, st, n, s->Nb, M, b->ldq, (long)b->Kb,
                       n, s->rchol_b, b->Lhb);
    AFQ_CHECK_LAUNCH();
  }
  AFQ_CHECK_HIP(hipStreamSynchronize(st));
  return AFQ_OK;
}

void bf3_release(Bf3System* b) {
  if (b->Lb) (void)hipFree(b->Lb);
  if (b->blk_pq) (void)hipFree(b->blk_pq);
  if (b->Lha) (void)hipFree(b->Lha);
  if (b->Lhb) (void)hipFree(b->Lhb);
  if (b->E) (void)hipFree(b->E);
  b->Lb = nullptr; b->blk_pq = nullptr; b->Lha = b->Lhb = nullptr; b->E = nullptr;
}

static size_t gemm_lds() { return (size_t)2 * 4 * GB * GK * sizeof(bf16_t); }

int bf3_vhs(hipStream_t s, const Bf3System& b, int W, const bf16_t* X, double dt, bf16_t* Vb) {
  if (W <= 0) return AFQ_OK;
  if (gemm_lds() > max_lds_bytes()) return fail(AFQ_EUNSUPPORTED, "bf16x3 GEMM: LDS %zu", gemm_lds());
  Bf3Gemm a{};
  a.nsplit = 1;
  a.sp[0].a[0] = X;
  a.sp[0].a[1] = X + x_rows(W) * b.Gp;
  a.sp[0].lda = b.Gp;
  a.sp[0].tiled = 1;
  a.sp[0].b[0] = b.Lb;
  a.sp[0].b[1] = b.Lb + (size_t)b.nblk * 256 * b.Gp;
  a.sp[0].ldb = b.Gp;
  a.sp[0].kbeg = 0;
  a.sp[0].kend = b.Gp;
  const size_t ab = x_rows(W) * b.Gp * 2, bb = (size_t)b.nblk * 256 * b.Gp * 2;
  if (ab >= (1ull << 32) || bb >= (1ull << 32))
    return fail(AFQ_EUNSUPPORTED, "bf16x3 VHS: operand plane of %zu / %zu bytes (limit 4 GiB)", ab, bb);
  a.sp[0].abytes = (unsigned)ab;
  a.sp[0].bbytes = (unsigned)bb;
  a.rows = 2 * W;
  a.cols = b.nblk * 256;
  a.ntm = cdiv_int(2L * W, GB);
  a.ntn = b.nblk;
  a.blk_pq = b.blk_pq;
  a.RT = b.RT;
  a.W = W;
  a.vstride = (long)b.RT * b.RT * 1024;
  a.Vb = Vb;
  a.sqdt = (float)sqrt(dt);
  hipLaunchKernelGGL((k_bf3_gemm<0>), dim3(a.ntm * a.ntn, 1), dim3(GT), gemm_lds(), s, a);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int bf3_force_bias(hipStream_t s, const Bf3System& b, int W, const bf16_t* Gha, const bf16_t* Ghb,
                   float* part, int* nsplit_out, int per_spin) {
  Bf3Gemm a{};
  int n = 0;
  per_spin = std::min(2, std::max(1, per_spin));
  const int Ks[2] = {b.Ka, b.Kb};
  const int Nrow[2] = {b.Na, b.Nb};
  const bf16_t* gh[2] = {Gha, Ghb};
  const bf16_t* lh[2] = {b.Lha, b.Lhb};
  for (int sp = 0; sp < 2; ++sp) {
    if (Nrow[sp] == 0 || Ks[sp] == 0) continue;
    const int K = Ks[sp];
    const int chunk = ((K / 32 + per_spin - 1) / per_spin) * 32;
    for (int k0 = 0; k0 < K; k0 += chunk) {
      Bf3Split& S = a.sp[n++];
      S.a[0] = gh[sp];
      S.a[1] = gh[sp] + (size_t)2 * W * K;
      S.lda = K;
      S.b[0] = lh[sp];
      S.b[1] = lh[sp] + (size_t)b.G * K;
      S.ldb = K;
      S.kbeg = k0;
      S.kend = std::min(K, k0 + chunk);
      const size_t ab = (size_t)2 * W * K * 2, bb = (size_t)b.G * K * 2;
      if (ab >= (1ull << 32) || bb >= (1ull << 32))
        return fail(AFQ_EUNSUPPORTED, "bf16x3 force bias: operand plane of %zu / %zu bytes (limit 4 GiB)",
                    ab, bb);
      S.abytes = (unsigned)ab;
      S.bbytes = (unsigned)bb;
    }
  }
  // no vectors: no partials (the caller zero-fills the force bias)
  *nsplit_out = b.G > 0 ? n : 0;
  if (W <= 0 || n == 0 || b.G <= 0) return AFQ_OK;
  if (gemm_lds() > max_lds_bytes()) return fail(AFQ_EUNSUPPORTED, "bf16x3 GEMM: LDS %zu", gemm_lds());
  a.nsplit = n;
  a.rows = 2 * W;
  a.cols = b.G;
  a.ntm = cdiv_int(2L * W, GB);
  a.ntn = cdiv_int(b.G, GB);
  a.W = W;
  a.G = b.G;
  a.part = (float2*)part;
  hipLaunchKernelGGL((k_bf3_gemm<1>), dim3(a.ntm * a.ntn, n), dim3(GT), gemm_lds(), s, a);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int bf3_split_rows(hipStream_t s, const double* A, int rows, int cols, int cols_p, bf16_t* out) {
  const long n = (long)rows * cols_p;
  if (n <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_prep_mat, dim3(1024), dim3(256), 0, s, n, rows, cols, cols_p, n, A, 0.0, out, 0);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int bf3_packed_force_bias(hipStream_t s, int W, int G, int Kp, const bf16_t* gpl,
                          const bf16_t* lpl, float* part, int* nsplit_out) {
  *nsplit_out = 0;
  if (W <= 0 || G <= 0 || Kp <= 0) return AFQ_OK;
  if (Kp % GK) return fail(AFQ_EUNSUPPORTED, "bf16x3 packed force bias: K %d", Kp);
  if (gemm_lds() > max_lds_bytes()) return fail(AFQ_EUNSUPPORTED, "bf16x3 GEMM: LDS %zu", gemm_lds());
  const size_t ab = (size_t)2 * W * Kp * 2, bb = (size_t)G * Kp * 2;
  if (ab >= (1ull << 32) || bb >= (1ull << 32))
    return fail(AFQ_EUNSUPPORTED, "bf16x3 packed force bias: operand plane of %zu / %zu bytes", ab,
                bb);
  Bf3Gemm a{};
  constexpr int nsp = 4;
  const int chunk = ((Kp / GK + nsp - 1) / nsp) * GK;
  int n = 0;
  for (int k0 = 0; k0 < Kp; k0 += chunk) {
    Bf3Split& S = a.sp[n++];
    S.a[0] = gpl;
    S.a[1] = gpl + (size_t)2 * W * Kp;
    S.lda = Kp;
    S.b[0] = lpl;
    S.b[1] = lpl + (size_t)G * Kp;
    S.ldb = Kp;
    S.kbeg = k0;
    S.kend = std::min(Kp, k0 + chunk);
    S.abytes = (unsigned)ab;
    S.bbytes = (unsigned)bb;
  }
  a.nsplit = n;
  a.rows = 2 * W;
  a.cols = G;
  a.ntm = cdiv_int(2L * W, GB);
  a.ntn = cdiv_int(G, GB);
  a.W = W;
  a.G = G;
  a.part = (float2*)part;
  hipLaunchKernelGGL((k_bf3_gemm<1>), dim3(a.ntm * a.ntn, n), dim3(GT), gemm_lds(), s, a);
  AFQ_CHECK_LAUNCH();
  *nsplit_out = n;
  return AFQ_OK;
}

size_t bf3_prop_scratch_bytes(const Bf3System& b, int W) {
  const long ncb = (b.Na + b.Nb + b.pcols - 1) / b.pcols;
  return (size_t)W * ncb * b.RT * 3 * 64 * 2 * sizeof(float4);
}

int bf3_propagate(hipStream_t s, const Bf3System& b, int W, const bf16_t* Vb, float* t0,
                  double* phia, double* phib, int nmax) {
  const int Nt = b.Na + b.Nb;
  if (b.pcols < 16 || b.pcols > PC) return fail(AFQ_EINVAL, "bf16x3 propagator: block %d", b.pcols);
  const int nblk = (Nt + b.pcols - 1) / b.pcols;
  if (W <= 0 || nblk == 0) return AFQ_OK;
  const int bs = (Nt + nblk - 1) / nblk;
  const size_t sm = (size_t)4 * PC * (b.Kp + 8) * sizeof(bf16_t);
  if (sm > max_lds_bytes()) return fail(AFQ_EUNSUPPORTED, "bf16x3 propagator: LDS %zu", sm);
  if (b.RT > 4 * PW) return fail(AFQ_EUNSUPPORTED, "bf16x3 propagator: M=%d", b.M);
  if (!t0) return fail(AFQ_EINVAL, "bf16x3 propagator: no scratch");
  const dim3 grid(W * nblk), blk(PT);
#define AFQ_PROP_BF3(MT)                                                                      \
  hipLaunchKernelGGL((k_prop_bf3<MT>), grid, blk, sm, s, W, b.M, b.RT, b.Kp, b.Na, b.Nb, nblk, bs, \
                     b.E, Vb, (float4*)t0, phia, phib, nmax)
  if (b.RT <= PW) AFQ_PROP_BF3(1);
  else if (b.RT <= 2 * PW) AFQ_PROP_BF3(2);
  else if (b.RT <= 3 * PW) AFQ_PROP_BF3(3);
  else AFQ_PROP_BF3(4);
#undef AFQ_PROP_BF3
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}


// ----------------------------------------------------------------------------
// exchange energy (energy.py:51-91): for every (g, walker) of one spin
//   T = Lhalf_g . Ghalf^T  (N x N complex),   E_x += sum_ij T_ij T_ji
// as one GEMM with rows (g, i) of Lhalf and columns (walker, re/im, j) of
// Ghalf, K = q.  A wave owns an 80 x 80 tile holding ng = 80/N whole g blocks
// and nw = 80/(2N) whole walkers, so every N x N block is complete in its
// accumulators; the contraction T_ij T_ji goes through a wave-private LDS
// block.  Sums per (row tile, walker) are written out and reduced in a fixed
// order by k_exch_reduce (no atomics: the result is deterministic).
// ----------------------------------------------------------------------------
namespace {

// A wave owns an XWT x XWT tile of the GEMM with rows (g, i) and columns
// (walker, re/im, j); a workgroup holds XCfg::WR x XWC waves.  Two shapes:
//   wide (4 x 2 waves: 320 x 160, one workgroup per CU, 131 KB of LDS), for
//     long K with at most two g blocks per wave: the epilogue runs in two
//     plane passes in blocks after staging buffer 0, which receives the next
//     tile's first K step meanwhile;
//   narrow (2 x 2 waves: 160 x 160, two workgroups per CU, 80 KB each), for
//     short K or many g blocks per wave: one epilogue pass per g block with
//     both planes, over the staging buffers, while the other workgroup
//     computes.
constexpr int XWT = 80, XWC = 2, XGR = 8;
constexpr int XMAXN = XWT / 2;
// epilogue row pitch in floats: 4 mod 8 keeps the column writes of 16 lanes
// and the 16-byte row reads on distinct banks
constexpr int XP = 44;
template <bool WIDE>
struct XCfg {
  static constexpr int WR = WIDE ? 4 : 2, NW = WR * XWC, NT = 64 * NW, WPC = WIDE ? 1 : 2;
  static constexpr int RT = WR * XWT, CT = XWC * XWT;
  // 16-row blocks of one staging buffer: A hi, A lo (RT rows each), then B hi,
  // B lo (CT rows each), rows of GK bf16 (64 bytes), each block written by one
  // LDS-DMA instruction; wave wv fills blocks wv + NW i
  static constexpr int ABLK = 2 * RT / 16, NBLK = ABLK + 2 * CT / 16;
  static constexpr int SLOT = (NBLK + NW - 1) / NW;
  static constexpr int STAGE = NBLK * 16 * GK;  // bf16
  // one wave's epilogue block: the T planes of one g block, T_c[wi][i][j] at
  // ((2 wi + c) N + i) XP + j (narrow: both planes, 2 nw N <= XWT rows) or
  // one plane at (wi N + i) XP + j (wide: nw N <= XWT / 2 rows), then a dummy
  // area of 256 floats for lanes whose values are not stored
  static constexpr int EROWS = WIDE ? XWT / 2 : XWT;
  static constexpr int DUM = EROWS * XP, EPI = DUM + 256;
  static constexpr int EOFF = WIDE ? STAGE : 0;  // bf16 offset of the epilogue blocks
  static size_t lds_bytes() {
    const size_t stages = (size_t)2 * STAGE * sizeof(bf16_t);
    const size_t epi = (size_t)EOFF * sizeof(bf16_t) + (size_t)NW * EPI * sizeof(float) +
                       (size_t)NW * XMAXN * 2 * sizeof(double);  // + wsum [NW][nw <= XMAXN][2]
    return std::max(stages, epi);
  }
};
static_assert(XP >= XMAXN && XP % 8 == 4 && (XWT / 2) * XP % 4 == 0, "epilogue layout");

struct Bf3Exch {
  const bf16_t* a;  // Lhalf hi plane (lo plane at +aplane): row (g, i) at g ka + i ldq
  const bf16_t* b;  // Ghalf hi plane (lo plane at +bplane): row (w, c, j) at (2 w + c) kb + j ldq
  int aplane, bplane;
  int ka, kb;
  int N, G, W, ldq;
  int ng, nw, ntr, ntc;
  int nb;        // logical tiles: row tiles padded to groups of XGR, times ntc
  int mN, m2N;   // ceil(2^16 / N), ceil(2^16 / 2N): x / N = (x mN) >> 16 for x < 512
  double* part;  // [ntr][W] complex
};

// Persistent grid: XCD x owns a contiguous range of logical tiles and its
// workgroups take them in turn, so at any time an XCD works on neighbouring
// tiles (inside a group of XGR row tiles the row tile runs fastest, so the L2
// holds both operands).  Operand chunks go from global memory straight into
// two LDS stages by buffer-load DMA (no staging registers), one barrier per
// K step; the next K step's chunks are issued between the first product
// groups of this one (issued all at once they back up the vector memory queue
// and stall the wave before its products).
template <int VW, bool WIDE>
__global__ __launch_bounds__(XCfg<WIDE>::NT, XCfg<WIDE>::WPC) void k_exch_bf3(Bf3Exch a) {
  using C = XCfg<WIDE>;
  extern __shared__ __align__(16) bf16_t lds[];
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int wr = wv / XWC, wc = wv % XWC;
  const int li = l & 15, lq = l >> 4;
  const int N = a.N, NR = a.ng * N, NC = a.nw * 2 * N;
  const int nx = gridDim.x >> 3, jx = blockIdx.x >> 3;
  const int per = (a.nb + 7) >> 3;
  const int v0 = (blockIdx.x & 7) * per, v1 = min(a.nb, v0 + per);
  auto tile_of = [&](int v, int& tr, int& tc) {
    const int grp = v / (XGR * a.ntc), in = v % (XGR * a.ntc);
    tr = grp * XGR + in % XGR;
    tc = in / XGR;
  };
  // the first tile at or after v with a real row tile (padding only in the last group)
  auto next_tile = [&](int v) {
    for (; v < v1; v += nx) {
      int tr, tc;
      tile_of(v, tr, tc);
      if (tr < a.ntr) break;
    }
    return v;
  };
  int v = next_tile(v0 + jx);
  if (v >= v1) return;  // whole workgroup: before any barrier

  const i32x4 rsa = buf_rsrc(a.a, 4u * a.aplane), rsb = buf_rsrc(a.b, 4u * a.bplane);

  // DMA slots: wave wv's instruction i fills 16-row block b = wv + C::NW i of
  // the stage (blocks < C::ABLK: A rows (plane, row), then B rows); lane l
  // fetches row l / 4 of the block, K chunk xswz(l / 4, l % 4).  An empty
  // slot (row outside the tile's g blocks or walkers) reads offset 0: it only
  // feeds outputs the epilogue ignores.
  const int wvu = __builtin_amdgcn_readfirstlane(wv);
  const int lrow = l >> 2, lkc = xswz(lrow, l & 3);
  int off[C::SLOT];
  auto slots = [&](int tr, int tc) {
    int tt = t;  // opaque per tile: keeps the index arithmetic out of registers between tiles
    asm volatile("" : "+v"(tt));
    const int lr = (tt & 63) >> 2;
#pragma unroll
    for (int i = 0; i < C::SLOT; ++i) {
      const int b = wvu + C::NW * i;  // uniform
      off[i] = 0;
      if (b < C::ABLK) {
        const int p = b >= C::RT / 16, r = 16 * (b - p * (C::RT / 16)) + lr;
        const int wrr = r / XWT, rr = r - wrr * XWT;  // row wave, row in it
        const int gq = (rr * a.mN) >> 16, ii = rr - gq * N;
        const int g = (tr * C::WR + wrr) * a.ng + gq;
        const bool ok = rr < NR && g < a.G;
        off[i] = ok ? 2 * (p * a.aplane + g * a.ka + ii * a.ldq + 8 * lkc) : 0;
      } else if (b < C::NBLK) {
        const int bb = b - C::ABLK, p = bb >= C::CT / 16, r = 16 * (bb - p * (C::CT / 16)) + lr;
        const int wcc = r / XWT, rr = r - wcc * XWT;  // column wave, column in it
        const int wq = (rr * a.m2N) >> 16, rem = rr - wq * 2 * N;
        const int ri = rem >= N, j = rem - ri * N;
        const int w = (tc * XWC + wcc) * a.nw + wq;
        const bool ok = rr < NC && w < a.W;
        off[i] = ok ? 2 * (p * a.bplane + (2 * w + ri) * a.kb + j * a.ldq + 8 * lkc) : 0;
      }
    }
  };
  // chunks at K offset k0 into stage st (rows are ldq long, a multiple of GK,
  // zero past M: chunks never run into the next row)
  auto fetch_slot = [&](int st, int k0, int i) {
    const unsigned base = (unsigned)(uintptr_t)(lds + st * C::STAGE);
    const int b = wvu + C::NW * i;  // 16-row block, uniform
    if (b < C::ABLK) dma16(rsa, base + b * 1024, off[i], 2 * k0);
    else if (b < C::NBLK) dma16(rsb, base + b * 1024, off[i], 2 * k0);
  };
  auto fetch = [&](int st, int k0) {
#pragma unroll
    for (int i = 0; i < C::SLOT; ++i) fetch_slot(st, k0, i);
  };
  auto dma_wait = []() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); };
  const int nk = a.ldq / GK;

  float* Tw = (float*)(lds + C::EOFF) + wv * C::EPI;
  double* wsum = (double*)((float*)(lds + C::EOFF) + C::NW * C::EPI);  // [NW][nw][2]
  // lane offset of (row li, chunk lq) inside a 16-row block, in bf16
  const int frag = li * GK + 8 * xswz(li, lq);

  if (WIDE) {
    int tr, tc;
    tile_of(v, tr, tc);
    slots(tr, tc);
    fetch(0, 0);
  }
  while (v < v1) {
    int tr, tc;
    tile_of(v, tr, tc);
    if (!WIDE) {
      slots(tr, tc);
      fetch(0, 0);
    }
    // stage 0 holds this tile's first K step; in the wide shape the barrier
    // also ends the previous tile's epilogue, whose blocks the K loop's DMA
    // overwrites
    dma_wait();
    __syncthreads();

    f4v acc[5][5];
#pragma unroll
    for (int u = 0; u < 5; ++u)
#pragma unroll
      for (int w = 0; w < 5; ++w) acc[u][w] = f4v{0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < nk; ++kt) {
      const bf16_t* st = lds + (kt & 1) * C::STAGE + frag;
      uint4 ah[5], al[5];
#pragma unroll
      for (int u = 0; u < 5; ++u) {
        const int r = XWT * wr + 16 * u;  // block-aligned row
        ah[u] = *(const uint4*)(st + r * GK);
        al[u] = *(const uint4*)(st + (C::RT + r) * GK);
      }
#pragma unroll
      for (int w = 0; w < 5; ++w) {
        const int c = XWT * wc + 16 * w;
        const uint4 bh = *(const uint4*)(st + (2 * C::RT + c) * GK);
        const uint4 bl = *(const uint4*)(st + (2 * C::RT + C::CT + c) * GK);
#pragma unroll
        for (int u = 0; u < 5; ++u) {
          acc[u][w] = mfma3(ah[u], al[u], bh, bl, acc[u][w]);
          const int q = 5 * w + u;  // static: one DMA slot after each of the first groups
          if (q < C::SLOT && kt + 1 < nk) fetch_slot((kt + 1) & 1, (kt + 1) * GK, q);
        }
      }
      dma_wait();
      __syncthreads();
    }

    // wide: the next tile's first K step goes to stage 0 while the epilogue runs
    const int vn = next_tile(v + nx);
    if (WIDE && vn < v1) {
      int trn, tcn;
      tile_of(vn, trn, tcn);
      slots(trn, tcn);
      fetch(0, 0);
    }

    // Epilogue: E(g, w) = sum_ij T_ij T_ji for the g blocks and walkers of
    // the wave tile.  With T = Tr + i Ti,
    //   Re E = sum Tr_ij Tr_ji - sum Ti_ij Ti_ji,
    //   Im E = sum Tr_ij Ti_ji + sum Ti_ij Tr_ji,
    // so a lane needs the transposed partners of its elements in its own
    // plane (sum A) and in the other one (sum B).  A lane holds rows
    // 4 lq .. 4 lq + 3 of one column; it writes them into the wave's block,
    // and after a wave barrier reads the partners T[wi][j][i .. i + 3], one
    // 16-byte load (one load per element when N % 4 != 0, where a lane's rows
    // can span two g blocks).  Wide: pass (gi, c) stores plane c of g block gi
    // only, a lane of plane c adds its products to A, one of the other plane
    // to B; narrow: pass gi stores both planes.  Products and a lane's partial
    // sums over one g block (at most 16 terms per sum) are FP32, the sums over
    // g blocks, lanes and tiles FP64, all in a fixed order.
    int wiv[5], jv[5];
    bool cvv[5], cimv[5];
    {
      int lj = li;  // opaque per tile, as in slots()
      asm volatile("" : "+v"(lj));
#pragma unroll
      for (int w = 0; w < 5; ++w) {
        const int col = 16 * w + lj;
        const bool cv = col < NC;
        const int wi = cv ? (col * a.m2N) >> 16 : 0;
        const int rem = col - wi * 2 * N;
        const bool cim = cv && rem >= N;
        cvv[w] = cv;
        cimv[w] = cim;
        wiv[w] = wi;
        jv[w] = cv ? rem - (cim ? N : 0) : 0;
      }
    }
    double cA[5], cB[5];
#pragma unroll
    for (int w = 0; w < 5; ++w) cA[w] = cB[w] = 0.0;
    const int gw0 = (tr * C::WR + wr) * a.ng;
    if constexpr (WIDE) {
      int wb[5], rb[5];
#pragma unroll
      for (int w = 0; w < 5; ++w) {
        wb[w] = wiv[w] * N * XP + jv[w];    // + i XP: this lane's element (i, j)
        rb[w] = (wiv[w] * N + jv[w]) * XP;  // + i: its partner (j, i)
      }
      if (VW == 4) {
        // a lane's rows 4 lq + 16 u + (0..3) are rows i0 .. i0 + 3 of g block gi
        // (i0 = 16 u + 4 lq - gi N, whole when N % 4 == 0); its element (i, j)
        // sits at wpos + (16 u - gi N) XP, the partners (j, i0 .. i0 + 3) at
        // rpos + 16 u - gi N
        int wpos[5], rpos[5];
  #pragma unroll
        for (int w = 0; w < 5; ++w) {
          wpos[w] = wb[w] + 4 * lq * XP;
          rpos[w] = cvv[w] ? rb[w] + 4 * lq : C::DUM + 64;  // outside the tile: dummy area
        }
        for (int gi = 0; gi < a.ng; ++gi) {
          if (gw0 + gi >= a.G) break;  // wave-uniform
          const int lo = gi * N, hi = lo + N;  // rows of this g block in the wave tile
          float S[2][5];  // S[c][w]: the pass-c sums (plane c partners)
  #pragma unroll
          for (int c = 0; c < 2; ++c)
  #pragma unroll
            for (int w = 0; w < 5; ++w) S[c][w] = 0.f;
  #pragma unroll
          for (int c = 0; c < 2; ++c) {
  #pragma unroll
            for (int u = 0; u < 5; ++u) {
              if (16 * u + 16 <= lo || 16 * u >= hi) continue;  // uniform
              const int r0 = 16 * u + 4 * lq, du = 16 * u - lo;
              if (r0 >= lo && r0 < hi) {
  #pragma unroll
                for (int w = 0; w < 5; ++w) {
                  if (cvv[w] && (int)cimv[w] == c) {  // only the lanes of plane c write
                    float* d = Tw + wpos[w] + du * XP;
  #pragma unroll
                    for (int r = 0; r < 4; ++r) d[r * XP] = acc[u][w][r];
                  }
                }
              }
            }
            xwave_sync();
  #pragma unroll
            for (int u = 0; u < 5; ++u) {
              if (16 * u + 16 <= lo || 16 * u >= hi) continue;  // uniform
              const int r0 = 16 * u + 4 * lq, du = 16 * u - lo;
              if (r0 >= lo && r0 < hi) {
                float4 ps[5];  // all five partner loads in flight together
  #pragma unroll
                for (int w = 0; w < 5; ++w) ps[w] = *(const float4*)(Tw + rpos[w] + du);
  #pragma unroll
                for (int w = 0; w < 5; ++w) {
                  const f4v x = acc[u][w];
                  S[c][w] = fmaf(x[0], ps[w].x, S[c][w]);
                  S[c][w] = fmaf(x[1], ps[w].y, S[c][w]);
                  S[c][w] = fmaf(x[2], ps[w].z, S[c][w]);
                  S[c][w] = fmaf(x[3], ps[w].w, S[c][w]);
                }
              }
            }
            xwave_sync();
          }
          // own-plane partners feed A, the other plane's feed B
  #pragma unroll
          for (int w = 0; w < 5; ++w) {
            cA[w] += (double)(cimv[w] ? S[1][w] : S[0][w]);
            cB[w] += (double)(cimv[w] ? S[0][w] : S[1][w]);
          }
        }
      } else {
        for (int gi = 0; gi < a.ng; ++gi) {
          if (gw0 + gi >= a.G) break;  // wave-uniform
          const int lo = gi * N, hi = lo + N;  // rows of this g block in the wave tile
          float A[5], B[5];
  #pragma unroll
          for (int w = 0; w < 5; ++w) A[w] = B[w] = 0.f;
  #pragma unroll
          for (int c = 0; c < 2; ++c) {
  #pragma unroll
            for (int u = 0; u < 5; ++u) {
              if (16 * u + 16 <= lo || 16 * u >= hi) continue;  // uniform
  #pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int row = 16 * u + 4 * lq + r;
                const bool inb = row >= lo && row < hi;
  #pragma unroll
                for (int w = 0; w < 5; ++w) {
                  const bool own = inb && cvv[w] && (int)cimv[w] == c;
                  Tw[own ? wb[w] + (row - lo) * XP : C::DUM + l] = acc[u][w][r];
                }
              }
            }
            xwave_sync();
  #pragma unroll
            for (int u = 0; u < 5; ++u) {
              if (16 * u + 16 <= lo || 16 * u >= hi) continue;  // uniform
  #pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int row = 16 * u + 4 * lq + r;
                const bool inb = row >= lo && row < hi;
  #pragma unroll
                for (int w = 0; w < 5; ++w) {
                  const bool ok = inb && cvv[w];
                  const float x = ok ? acc[u][w][r] : 0.f;
                  const float pv = Tw[ok ? rb[w] + row - lo : C::DUM + l];
                  if ((int)cimv[w] == c) A[w] = fmaf(x, pv, A[w]);
                  else B[w] = fmaf(x, pv, B[w]);
                }
              }
            }
            xwave_sync();
          }
  #pragma unroll
          for (int w = 0; w < 5; ++w) {
            cA[w] += (double)A[w];
            cB[w] += (double)B[w];
          }
        }
      }
    } else {
      // this lane's element (i, j) at wb + i XP, its partners (j, i) in its
      // own plane at ro + i and in the other one at rx + i
      int wb[5], ro[5], rx[5];
#pragma unroll
      for (int w = 0; w < 5; ++w) {
        const int c = cimv[w] ? 1 : 0;
        wb[w] = (2 * wiv[w] + c) * N * XP + jv[w];
        ro[w] = ((2 * wiv[w] + c) * N + jv[w]) * XP;
        rx[w] = ((2 * wiv[w] + 1 - c) * N + jv[w]) * XP;
      }
      if (VW > 1) {
        // a lane's rows 4 lq .. 4 lq + 3 in groups of VW (N % VW == 0, so a
        // group never spans two g blocks); partner reads of columns outside
        // the tile go to the dummy area
        constexpr int NGR = 4 / VW;
        typedef float fvw __attribute__((ext_vector_type(VW)));
        int wpos[5], opos[5], xpos[5];
#pragma unroll
        for (int w = 0; w < 5; ++w) {
          wpos[w] = wb[w] + 4 * lq * XP;
          opos[w] = cvv[w] ? ro[w] + 4 * lq : C::DUM + 64;
          xpos[w] = cvv[w] ? rx[w] + 4 * lq : C::DUM + 64;
        }
        for (int gi = 0; gi < a.ng; ++gi) {
          if (gw0 + gi >= a.G) break;  // wave-uniform
          const int lo = gi * N, hi = lo + N;  // rows of this g block in the wave tile
          float A[5], B[5];
#pragma unroll
          for (int w = 0; w < 5; ++w) A[w] = B[w] = 0.f;
#pragma unroll
          for (int u = 0; u < 5; ++u) {
            if (16 * u + 16 <= lo || 16 * u >= hi) continue;  // uniform
#pragma unroll
            for (int h = 0; h < NGR; ++h) {
              const int rg = 16 * u + 4 * lq + VW * h, du = 16 * u - lo + VW * h;
              if (rg >= lo && rg < hi) {
#pragma unroll
                for (int w = 0; w < 5; ++w) {
                  if (cvv[w]) {
                    float* d = Tw + wpos[w] + du * XP;
#pragma unroll
                    for (int r = 0; r < VW; ++r) d[r * XP] = acc[u][w][VW * h + r];
                  }
                }
              }
            }
          }
          xwave_sync();
#pragma unroll
          for (int u = 0; u < 5; ++u) {
            if (16 * u + 16 <= lo || 16 * u >= hi) continue;  // uniform
#pragma unroll
            for (int h = 0; h < NGR; ++h) {
              const int rg = 16 * u + 4 * lq + VW * h, du = 16 * u - lo + VW * h;
              if (rg >= lo && rg < hi) {
                fvw po[5], px[5];  // all ten partner loads in flight together
#pragma unroll
                for (int w = 0; w < 5; ++w) {
                  po[w] = *(const fvw*)(Tw + opos[w] + du);
                  px[w] = *(const fvw*)(Tw + xpos[w] + du);
                }
#pragma unroll
                for (int w = 0; w < 5; ++w) {
#pragma unroll
                  for (int r = 0; r < VW; ++r) {
                    const float x = acc[u][w][VW * h + r];
                    A[w] = fmaf(x, po[w][r], A[w]);
                    B[w] = fmaf(x, px[w][r], B[w]);
                  }
                }
              }
            }
          }
          xwave_sync();
#pragma unroll
          for (int w = 0; w < 5; ++w) {
            cA[w] += (double)A[w];
            cB[w] += (double)B[w];
          }
        }
      } else {
#pragma unroll
        for (int w = 0; w < 5; ++w) {
          if (!cvv[w]) ro[w] = rx[w] = C::DUM + 4 * l;
        }
        for (int gi = 0; gi < a.ng; ++gi) {
          if (gw0 + gi >= a.G) break;  // wave-uniform
          const int lo = gi * N, hi = lo + N;  // rows of this g block in the wave tile
#pragma unroll
          for (int u = 0; u < 5; ++u) {
            if (16 * u + 16 <= lo || 16 * u >= hi) continue;  // uniform
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = 16 * u + 4 * lq + r;
              const bool inb = row >= lo && row < hi;
#pragma unroll
              for (int w = 0; w < 5; ++w)
                Tw[inb && cvv[w] ? wb[w] + (row - lo) * XP : C::DUM + l] = acc[u][w][r];
            }
          }
          xwave_sync();
          float A[5], B[5];
#pragma unroll
          for (int w = 0; w < 5; ++w) A[w] = B[w] = 0.f;
#pragma unroll
          for (int u = 0; u < 5; ++u) {
            if (16 * u + 16 <= lo || 16 * u >= hi) continue;  // uniform
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = 16 * u + 4 * lq + r;
              const bool inb = row >= lo && row < hi;
#pragma unroll
              for (int w = 0; w < 5; ++w) {
                const bool ok = inb && cvv[w];
                const int di = ok ? row - lo : 0;
                const float x = ok ? acc[u][w][r] : 0.f;
                A[w] = fmaf(x, Tw[ro[w] + di], A[w]);
                B[w] = fmaf(x, Tw[rx[w] + di], B[w]);
              }
            }
          }
#pragma unroll
          for (int w = 0; w < 5; ++w) {
            cA[w] += (double)A[w];
            cB[w] += (double)B[w];
          }
          xwave_sync();
        }
      }
    }
    // per walker: the lane's columns in a fixed order, then the wave
    for (int k = 0; k < a.nw; ++k) {  // uniform
      double x = 0.0, y = 0.0;
#pragma unroll
      for (int w = 0; w < 5; ++w) {
        if (cvv[w] && wiv[w] == k) {
          x += cimv[w] ? -cA[w] : cA[w];
          y += cB[w];
        }
      }
      x = wave_sum(x);
      y = wave_sum(y);
      if (l == 0) {
        wsum[(wv * a.nw + k) * 2] = x;
        wsum[(wv * a.nw + k) * 2 + 1] = y;
      }
    }
    __syncthreads();
    // the row waves of a column hold disjoint g ranges of the same walkers
    if (t < XWC * a.nw) {
      const int wcc = t / a.nw, wi = t % a.nw;
      const int w = (tc * XWC + wcc) * a.nw + wi;
      if (w < a.W) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int r = 0; r < C::WR; ++r) {
          re += wsum[((r * XWC + wcc) * a.nw + wi) * 2];
          im += wsum[((r * XWC + wcc) * a.nw + wi) * 2 + 1];
        }
        a.part[((long)tr * a.W + w) * 2] = re;
        a.part[((long)tr * a.W + w) * 2 + 1] = im;
      }
    }
    // narrow: the epilogue blocks are the staging of the next tile
    if (!WIDE) __syncthreads();
    v = vn;  // wide: the barrier at the loop top protects wsum and the epilogue blocks
  }
}

// e_x[w] = 1/2 sum_s sum_tr part_s[tr][w], in a fixed order: a block of
// 64 walkers x 4 row-tile lanes; lane q sums the row tiles r = q mod 4 of
// both spins, then the four lane sums are added in order.
__global__ __launch_bounds__(256) void k_exch_reduce(int W, int ntra, const double* __restrict__ pa,
                                                     int ntrb, const double* __restrict__ pb,
                                                     double* __restrict__ ex) {
  __shared__ double red[4][64][2];
  const int wl = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int w = blockIdx.x * 64 + wl;
  double re = 0.0, im = 0.0;
  if (w < W) {
    for (int r = q; r < ntra; r += 4) {
      re += pa[((long)r * W + w) * 2];
      im += pa[((long)r * W + w) * 2 + 1];
    }
    for (int r = q; r < ntrb; r += 4) {
      re += pb[((long)r * W + w) * 2];
      im += pb[((long)r * W + w) * 2 + 1];
    }
  }
  red[q][wl][0] = re;
  red[q][wl][1] = im;
  __syncthreads();
  if (q == 0 && w < W) {
    ex[2 * w] = 0.5 * (((red[0][wl][0] + red[1][wl][0]) + red[2][wl][0]) + red[3][wl][0]);
    ex[2 * w + 1] = 0.5 * (((red[0][wl][1] + red[1][wl][1]) + red[2][wl][1]) + red[3][wl][1]);
  }
}

// e_c[w] = 1/2 sum_g X_g^2 with X_g = sum of the split-K force-bias partials
__global__ __launch_bounds__(256) void k_coul_parts(int W, int G, int nsplit,
                                                    const float2* __restrict__ part,
                                                    double* __restrict__ ec) {
  __shared__ double red[32];
  const int w = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int g = threadIdx.x; g < G; g += 256) {
    double xr = 0.0, xi = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) {
      const float2 p = part[((long)sp * W + w) * G + g];
      xr += p.x;
      xi += p.y;
    }
    re += xr * xr - xi * xi;
    im += 2.0 * xr * xi;
  }
  const cplx c = block_sum_c(mkc(re, im), red);
  if (threadIdx.x == 0) {
    ec[2 * w] = 0.5 * c.x;
    ec[2 * w + 1] = 0.5 * c.y;
  }
}

// the wide shape for long K (>= 8 K steps) and at most two g blocks per wave
// with N % 4 == 0; measured at M=300, N=40 (wide 10.7 ms per spin, narrow
// 11.7) and M=150, N=20 / 18 (wide 1.04 / 1.87 ms, narrow 0.88 / 1.20)
bool exch_wide(int N, int ldq) { return N > 0 && N % 4 == 0 && XWT / N <= 2 && ldq / GK >= 8; }

void exch_geom(int N, int G, int W, int ldq, int* ng, int* nw, int* ntr, int* ntc) {
  const int wr = exch_wide(N, ldq) ? XCfg<true>::WR : XCfg<false>::WR;
  *ng = N > 0 ? XWT / N : 1;
  *nw = N > 0 ? std::max(1, XWT / (2 * N)) : 1;
  *ntr = cdiv_int(G, wr * *ng);
  *ntc = cdiv_int(W, XWC * *nw);
}

}  // namespace

// Lhalf planes of one determinant with occupation string occ: element
// (g, i, q) = L_g[occ[i]][q] (real L, identity-column determinants), laid out
// as k_prep_lh's ([plane][G][K], (i, q) at i ldq + q, zero padding)
__global__ void k_prep_lh_occ(long n, int N, int M, int ldq, long K, long plane,
                              const int* __restrict__ occ, const double* __restrict__ L,
                              bf16_t* __restrict__ out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long g = e / K;
    const int c = (int)(e % K);
    const int i = c / ldq, q = c % ldq;
    const double v = (i < N && q < M) ? L[(g * M + occ[i]) * M + q] : 0.0;
    bf16_t h, l;
    split2(v, h, l);
    out[e] = h;
    out[plane + e] = l;
  }
}

// Ghalf (W, N, M) complex -> split planes [2][2W][K], rows (w, re/im),
// (i, q) at i ldq + q, zero padding (the layout of GhalfBf16)
__global__ void k_prep_gh(long n, int N, int M, int ldq, long K, long plane,
                          const double* __restrict__ Gh, bf16_t* __restrict__ out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long r = e / K;
    const int c = (int)(e % K);
    const int i = c / ldq, q = c % ldq;
    const long w = r >> 1;
    const int ri = (int)(r & 1);
    const double v = (i < N && q < M) ? Gh[((w * N + i) * M + q) * 2 + ri] : 0.0;
    bf16_t h, l;
    split2(v, h, l);
    out[e] = h;
    out[plane + e] = l;
  }
}

bool bf3_exchange_supported(const Bf3System& b) { return b.Na <= XMAXN && b.Nb <= XMAXN; }

size_t bf3_exchange_part_bytes(const Bf3System& b, int W) {
  int ng, nw, ntra = 0, ntrb = 0, ntc;
  if (b.Na) exch_geom(b.Na, b.G, W, b.ldq, &ng, &nw, &ntra, &ntc);
  if (b.Nb) exch_geom(b.Nb, b.G, W, b.ldq, &ng, &nw, &ntrb, &ntc);
  return (size_t)(ntra + ntrb) * W * 2 * sizeof(double);
}

int bf3_exchange(hipStream_t s, const Bf3System& b, int W, const bf16_t* gha, const bf16_t* ghb,
                 double* part, double* ex) {
  if (W <= 0) return AFQ_OK;
  if (!bf3_exchange_supported(b))
    return fail(AFQ_EUNSUPPORTED, "bf16x3 exchange: nocc %d/%d > %d", b.Na, b.Nb, XMAXN);
  const size_t smax = std::max(XCfg<true>::lds_bytes(), XCfg<false>::lds_bytes());
  if (smax > max_lds_bytes()) return fail(AFQ_EUNSUPPORTED, "bf16x3 exchange: LDS %zu", smax);
  const int Ns[2] = {b.Na, b.Nb};
  const long Ks[2] = {b.Ka, b.Kb};
  const bf16_t* gh[2] = {gha, ghb};
  const bf16_t* lh[2] = {b.Lha, b.Lhb};
  int ntrs[2] = {0, 0};
  double* parts[2] = {part, part};
  for (int sp = 0; sp < 2; ++sp) {
    if (Ns[sp] == 0 || b.G == 0) continue;
    Bf3Exch a{};
    const long aplane = (long)b.G * Ks[sp], bplane = 2L * W * Ks[sp];
    if (4 * aplane >= (1L << 31) || 4 * bplane >= (1L << 31))
      return fail(AFQ_EUNSUPPORTED, "bf16x3 exchange: operand planes exceed 2^31 bytes");
    a.a = lh[sp];
    a.b = gh[sp];
    a.aplane = (int)aplane;
    a.bplane = (int)bplane;
    a.ka = (int)Ks[sp];
    a.kb = (int)Ks[sp];
    a.N = Ns[sp];
    a.G = b.G;
    a.W = W;
    a.ldq = b.ldq;
    exch_geom(a.N, b.G, W, b.ldq, &a.ng, &a.nw, &a.ntr, &a.ntc);
    a.part = part + (sp == 1 ? (size_t)ntrs[0] * W * 2 : 0);
    parts[sp] = a.part;
    ntrs[sp] = a.ntr;
    const long nb = (long)cdiv_int(a.ntr, XGR) * XGR * a.ntc;
    if (nb > 0x7fffffffL) return fail(AFQ_EUNSUPPORTED, "bf16x3 exchange: too many tiles");
    a.nb = (int)nb;
    a.mN = (65536 + a.N - 1) / a.N;
    a.m2N = (65536 + 2 * a.N - 1) / (2 * a.N);
    if (a.ldq % GK) return fail(AFQ_EUNSUPPORTED, "bf16x3 exchange: row stride %d", a.ldq);
    // persistent: WPC workgroups per CU, a multiple of 8 (one share per XCD)
    const bool wide = exch_wide(a.N, b.ldq);
    const int wpc = wide ? XCfg<true>::WPC : XCfg<false>::WPC;
    const long nwg = std::min<long>((long)num_cus() * wpc, (nb + 7) / 8 * 8);
    const dim3 grid((unsigned)std::max<long>(8, nwg / 8 * 8));
    // epilogue row groups: 4 rows when N % 4 == 0, 2 when N is even, else 1
    const dim3 bn(XCfg<false>::NT);
    const size_t sn = XCfg<false>::lds_bytes();
    if (wide)
      hipLaunchKernelGGL((k_exch_bf3<4, true>), grid, dim3(XCfg<true>::NT), XCfg<true>::lds_bytes(), s, a);
    else if (a.N % 4 == 0)
      hipLaunchKernelGGL((k_exch_bf3<4, false>), grid, bn, sn, s, a);
    else if (a.N % 2 == 0)
      hipLaunchKernelGGL((k_exch_bf3<2, false>), grid, bn, sn, s, a);
    else
      hipLaunchKernelGGL((k_exch_bf3<1, false>), grid, bn, sn, s, a);
    AFQ_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_exch_reduce, dim3(cdiv_int(W, 64)), dim3(256), 0, s, W, ntrs[0], parts[0],
                     ntrs[1], parts[1], ex);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

// force-bias partials of one determinant's Coulomb term: one split per spin
static size_t md_coul_part_bytes(const Bf3System& b, int W) {
  return (size_t)2 * W * std::max(b.G, 1) * 2 * sizeof(float);
}

size_t bf3_md_two_body_bytes(const Bf3System& b, int W) {
  return align256((size_t)2 * b.G * b.Ka * sizeof(bf16_t)) +
         align256((size_t)2 * b.G * b.Kb * sizeof(bf16_t)) +
         align256((size_t)2 * 2 * W * b.Ka * sizeof(bf16_t)) +
         align256((size_t)2 * 2 * W * b.Kb * sizeof(bf16_t)) +
         align256(bf3_exchange_part_bytes(b, W)) + align256(md_coul_part_bytes(b, W));
}

int bf3_md_two_body(hipStream_t s, const Bf3System& b, int W, const int* occa, const int* occb,
                    const double* chol, const double* gha, const double* ghb, void* scratch,
                    double* ec, double* ex) {
  if (W <= 0) return AFQ_OK;
  char* p = (char*)scratch;
  auto take = [&](size_t n) {
    char* r = p;
    p += align256(n);
    return r;
  };
  Bf3System d = b;
  d.Lha = (bf16_t*)take((size_t)2 * b.G * b.Ka * sizeof(bf16_t));
  d.Lhb = (bf16_t*)take((size_t)2 * b.G * b.Kb * sizeof(bf16_t));
  bf16_t* pa = (bf16_t*)take((size_t)2 * 2 * W * b.Ka * sizeof(bf16_t));
  bf16_t* pb = (bf16_t*)take((size_t)2 * 2 * W * b.Kb * sizeof(bf16_t));
  double* part = (double*)take(bf3_exchange_part_bytes(b, W));
  float* cpart = (float*)take(md_coul_part_bytes(b, W));
  const int Ns[2] = {b.Na, b.Nb};
  const long Ks[2] = {b.Ka, b.Kb};
  const int* occ[2] = {occa, occb};
  const double* gh[2] = {gha, ghb};
  bf16_t* lh[2] = {d.Lha, d.Lhb};
  bf16_t* gp[2] = {pa, pb};
  for (int sp = 0; sp < 2; ++sp) {
    if (Ns[sp] == 0 || b.G == 0) continue;
    const long nl = (long)b.G * Ks[sp];
    hipLaunchKernelGGL(k_prep_lh_occ, dim3(2048), dim3(256), 0, s, nl, Ns[sp], b.M, b.ldq, Ks[sp],
                       nl, occ[sp], chol, lh[sp]);
    AFQ_CHECK_LAUNCH();
    const long ng = 2L * W * Ks[sp];
    hipLaunchKernelGGL(k_prep_gh, dim3(2048), dim3(256), 0, s, ng, Ns[sp], b.M, b.ldq, Ks[sp], ng,
                       gh[sp], gp[sp]);
    AFQ_CHECK_LAUNCH();
  }
  // Coulomb: X_g = sum_s sum_(i,q) Lhalf_s[g][i q] Ghalf_s[w][i q], the force
  // bias contraction of this determinant's planes (no vectors or no
  // electrons: X = 0)
  int nsplit = 0;
  int rc = bf3_force_bias(s, d, W, pa, pb, cpart, &nsplit, 1);
  if (rc) return rc;
  if (nsplit > 0) {
    if ((rc = bf3_coulomb(s, W, b.G, nsplit, cpart, ec))) return rc;
  } else {
    if ((rc = zero_async(s, ec, (size_t)2 * W * sizeof(double)))) return rc;
  }
  return bf3_exchange(s, d, W, pa, pb, part, ex);
}

int bf3_coulomb(hipStream_t s, int W, int G, int nsplit, const float* part, double* ec) {
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_coul_parts, dim3(W), dim3(256), 0, s, W, G, nsplit, (const float2*)part, ec);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace afq
