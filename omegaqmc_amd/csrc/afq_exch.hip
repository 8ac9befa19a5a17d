// Fused FP64 exchange energy (OmegaQMC/observables/energy.py:51-91, the
// exchange half; afqmc_pw_heg.py:414-457 with vector signs):
//   E_x[w] = 1/2 sum_g sign_g sum_ij T_ij T_ji,   T = Lhalf_g . Ghalf_w^T (N x N)
// as one real GEMM on v_mfma_f64_16x16x4f64: rows (g, i) of Lhalf, columns
// (walker, j, re/im) of Ghalf, K = the orbital index.  A wave owns a 48 x 80
// tile: ng = 48/N whole g blocks by nw = 80/(2N) whole walkers, so every
// N x N block T_gw is complete in its accumulators, and the pairing
// T_ij T_ji goes through a wave-private LDS block one g block at a time.  T
// never reaches HBM (the slab path wrote gchunk N x W N blocks of it).  With
// T = Tr + i Ti:  Re E = sum Tr_ij Tr_ji - Ti_ij Ti_ji,
//                 Im E = sum Tr_ij Ti_ji + Ti_ij Tr_ji,
// so a lane pairs its value with the transposed partner in its own plane
// (real part) and in the other plane (imaginary part).  Sums per (row tile,
// walker) go to a scratch array and are reduced in a fixed order: the result
// is deterministic.
#include "afq_kernels.h"

namespace afq {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int ER = 3, EC = 5;            // 16-row / 16-column tiles per wave
constexpr int ERW = 16 * ER, ECW = 16 * EC;
constexpr int ETP = EXCH_F64_NMAX + 1;   // LDS row pitch (doubles)

struct ExchF64 {
  const double* A;     // Lhalf (G, N, M) real
  const double* B;     // Ghalf (W, N, M) complex
  const double* sign;  // (G) or NULL
  int W, M, N, G, ng, nw;
  int rt0, nrt, nct;   // row tiles [rt0, rt0 + nrt) in this launch, column tiles
  double* part;        // [nrt][W] complex
};

// consecutive virtual indices run on one XCD (the hardware deals workgroups
// round-robin over the 8 XCDs)
__device__ __forceinline__ int exch_xcd_remap(int b, int nb) {
  const int x = b & 7, slot = b >> 3;
  const int per = nb >> 3, rem = nb & 7;
  return (x < rem) ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}

__device__ __forceinline__ void exch_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_exch_f64(ExchF64 a) {
  // one plane: [wl][i][j] at (wl N + i) ETP + j, then a dummy slot that the
  // lanes of empty columns write and read
  __shared__ double Tl[(ECW / 2) * ETP + 2];
  const int l = threadIdx.x, li = l & 15, lq = l >> 4;
  const int N = a.N, M = a.M;
  const int nt = a.nrt * a.nct;
  const int v = exch_xcd_remap(blockIdx.x, nt);
  // groups of 8 row tiles; inside a group the row tile runs fastest, so
  // neighbouring workgroups share the walkers' Ghalf and the group's Lhalf
  // rows stay in L2 across the column tiles
  const int grp = v / (8 * a.nct);
  const int gs = min(8, a.nrt - grp * 8);
  const int in = v - grp * 8 * a.nct;
  const int trl = grp * 8 + in % gs, tc = in / gs;
  const int tr = a.rt0 + trl;

  // operand rows/columns of this lane (fixed over K)
  long aoff[ER], boff[EC];
  bool av[ER], bv[EC];
  int cwl[EC], cj[EC], cc[EC];
#pragma unroll
  for (int u = 0; u < ER; ++u) {
    const int r = 16 * u + li, gb = r / N, i = r - gb * N;
    const int g = tr * a.ng + gb;
    av[u] = gb < a.ng && g < a.G;
    aoff[u] = av[u] ? ((long)g * N + i) * M : 0;
  }
#pragma unroll
  for (int x = 0; x < EC; ++x) {
    const int cl = 16 * x + li, p = cl >> 1, c = cl & 1;
    const int wl = p / N, j = p - wl * N;
    const int w = tc * a.nw + wl;
    bv[x] = wl < a.nw && w < a.W;
    boff[x] = bv[x] ? (((long)w * N + j) * M) * 2 + c : 0;
    cwl[x] = wl;
    cj[x] = j;
    cc[x] = c;
  }

  d4v acc[ER][EC];
#pragma unroll
  for (int u = 0; u < ER; ++u)
#pragma unroll
    for (int x = 0; x < EC; ++x) acc[u][x] = d4v{0.0, 0.0, 0.0, 0.0};

  // K loop: lane (li, lq) feeds A[row li][k0 + lq] and B[k0 + lq][col li];
  // the next k-step's loads are in flight while this one multiplies (the
  // step past M loads zeros)
  // Loads are unconditional (an empty row or column reads element 0 and a
  // k past M reads k = M - 1), then masked by selects: no branches in the
  // loop.  An empty column's products land only in its own accumulators,
  // which the epilogue skips.
  // 32-bit element offsets from the wave-uniform bases (the host checks the
  // operand sizes), so every load is base + one VGPR offset
  const double* __restrict__ Ab = a.A;
  const double* __restrict__ Bb = a.B;
  int oa[ER], ob[EC];
#pragma unroll
  for (int u = 0; u < ER; ++u) oa[u] = (int)aoff[u];
#pragma unroll
  for (int x = 0; x < EC; ++x) ob[x] = (int)boff[x];
  double ra0[ER], rb0[EC], ra1[ER], rb1[EC];
  auto load = [&](int k0, double* xa, double* xb) {
    const int kk = k0 + lq;
    const bool kin = kk < M;
    const int k = kin ? kk : M - 1;
#pragma unroll
    for (int u = 0; u < ER; ++u) {
      const double t = Ab[oa[u] + k];
      xa[u] = (av[u] && kin) ? t : 0.0;
    }
#pragma unroll
    for (int x = 0; x < EC; ++x) {
      const double t = Bb[ob[x] + 2 * k];
      xb[x] = kin ? t : 0.0;
    }
  };
  auto mma = [&](const double* xa, const double* xb) {
#pragma unroll
    for (int u = 0; u < ER; ++u)
#pragma unroll
      for (int x = 0; x < EC; ++x)
        acc[u][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[u], xb[x], acc[u][x], 0, 0, 0);
  };
  load(0, ra0, rb0);
#pragma unroll 1
  for (int k0 = 0; k0 < M; k0 += 4) {
    load(k0 + 4, ra1, rb1);
    mma(ra0, rb0);
#pragma unroll
    for (int u = 0; u < ER; ++u) ra0[u] = ra1[u];
#pragma unroll
    for (int x = 0; x < EC; ++x) rb0[x] = rb1[x];
  }

  // epilogue, one g block and one plane at a time: D[i][j] of tile (u, x)
  // sits in lane li + 16 lq, register r, row i = lq + 4 r.  Plane p's values
  // go to LDS; every lane then reads the transposed partner T_p[j][i] of its
  // own value: the same plane feeds the real part, the other plane the
  // imaginary part.
  double cre[EC], cim[EC];
#pragma unroll
  for (int x = 0; x < EC; ++x) cre[x] = cim[x] = 0.0;
  // LDS element of T_p[wl][i][j] for the lane's column x: write base (+ i
  // ETP) and transposed-read base (+ i); empty columns point at a dummy slot
  int wbs[EC], rbs[EC];
#pragma unroll
  for (int x = 0; x < EC; ++x) {
    wbs[x] = cwl[x] < a.nw ? cwl[x] * N * ETP + cj[x] : (ECW / 2) * ETP;
    rbs[x] = bv[x] ? (cwl[x] * N + cj[x]) * ETP : (ECW / 2) * ETP;
  }
  for (int gb = 0; gb < a.ng; ++gb) {
    const int g = tr * a.ng + gb;
    if (g >= a.G) break;  // uniform
    const int lo = gb * N, hi = lo + N;
    const double sg = a.sign ? a.sign[g] : 1.0;
    for (int p = 0; p < 2; ++p) {
      // same-plane pairs: +Tr Tr (p = 0), -Ti Ti (p = 1), times the g sign
      const double fa = p ? -sg : sg;
      // opaque per pass: keeps the per-entry address arithmetic inside the
      // pass instead of hoisted (and held in registers) across passes
      int wb[EC], rb[EC];
#pragma unroll
      for (int x = 0; x < EC; ++x) {
        wb[x] = wbs[x];
        rb[x] = rbs[x];
        asm volatile("" : "+v"(wb[x]), "+v"(rb[x]));
      }
#pragma unroll
      for (int u = 0; u < ER; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * u + lq + 4 * r;
          if (row < lo || row >= hi) continue;
          const int ro = (row - lo) * ETP;
#pragma unroll
          for (int x = 0; x < EC; ++x)
            if (cc[x] == p) Tl[wb[x] + ro] = acc[u][x][r];
        }
      exch_wave_sync();
#pragma unroll
      for (int u = 0; u < ER; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * u + lq + 4 * r;
          if (row < lo || row >= hi) continue;
          const int i = row - lo;
#pragma unroll
          for (int x = 0; x < EC; ++x) {
            const double q = acc[u][x][r] * Tl[rb[x] + i];
            if (cc[x] == p) cre[x] = fma(fa, q, cre[x]);
            else cim[x] = fma(sg, q, cim[x]);
          }
        }
      exch_wave_sync();
    }
  }
  // per walker: the lane's columns in a fixed order, then the wave
  for (int wl = 0; wl < a.nw; ++wl) {  // uniform
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int x = 0; x < EC; ++x)
      if (bv[x] && cwl[x] == wl) {
        re += cre[x];
        im += cim[x];
      }
    re = wave_sum(re);
    im = wave_sum(im);
    const int w = tc * a.nw + wl;
    if (l == 0 && w < a.W) {
      a.part[((long)trl * a.W + w) * 2] = re;
      a.part[((long)trl * a.W + w) * 2 + 1] = im;
    }
  }
}

// e[w] (+)= 1/2 sum_tr part[tr][w], row tiles in order
__global__ void k_exch_f64_reduce(int W, int nrt, const double* __restrict__ part,
                                  double* __restrict__ e, int accumulate) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  double re = 0.0, im = 0.0;
  for (int r = 0; r < nrt; ++r) {
    re += part[((long)r * W + w) * 2];
    im += part[((long)r * W + w) * 2 + 1];
  }
  const double x = 0.5 * re, y = 0.5 * im;
  e[2 * w] = accumulate ? e[2 * w] + x : x;
  e[2 * w + 1] = accumulate ? e[2 * w + 1] + y : y;
}

}  // namespace

bool exchange_f64_fused_fits(int N) { return N > 0 && N <= EXCH_F64_NMAX; }

int exchange_f64_fused(hipStream_t s, int W, int M, int N, int G, const double* Ghalf,
                       const double* rchol, double* partial, size_t partial_bytes,
                       double* e_exch, bool accumulate, const double* sign) {
  if (W <= 0) return AFQ_OK;
  if (!exchange_f64_fused_fits(N)) return fail(AFQ_EUNSUPPORTED, "fused exchange: N=%d", N);
  ExchF64 a;
  a.A = rchol; a.B = Ghalf; a.sign = sign;
  a.W = W; a.M = M; a.N = N; a.G = G;
  a.ng = ERW / N;
  a.nw = std::max(1, ECW / (2 * N));
  a.nct = cdiv_int(W, a.nw);
  a.part = partial;
  if ((long)G * N * M >= 0x7fffffffL || 2L * W * N * M >= 0x7fffffffL)
    return fail(AFQ_EUNSUPPORTED, "fused exchange: operands beyond 2^31 elements");
  const int ntr = cdiv_int(G, a.ng);
  // row tiles per launch: as many as the partial buffer holds
  const long cap = (long)(partial_bytes / (16UL * W));
  if (cap < 1) return fail(AFQ_EINVAL, "fused exchange: partial buffer of %zu bytes", partial_bytes);
  bool acc = accumulate;
  if (ntr == 0 && !acc) AFQ_CHECK_HIP(hipMemsetAsync(e_exch, 0, 16UL * W, s));
  for (int rt0 = 0; rt0 < ntr; rt0 += (int)std::min<long>(cap, ntr)) {
    a.rt0 = rt0;
    a.nrt = (int)std::min<long>(cap, ntr - rt0);
    const long nb = (long)a.nrt * a.nct;
    if (nb > 0x7fffffffL) return fail(AFQ_EUNSUPPORTED, "fused exchange: grid too large");
    hipLaunchKernelGGL(k_exch_f64, dim3((unsigned)nb), dim3(64), 0, s, a);
    AFQ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_exch_f64_reduce, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, a.nrt,
                       partial, e_exch, acc ? 1 : 0);
    AFQ_CHECK_LAUNCH();
    acc = true;
  }
  return AFQ_OK;
}

}  // namespace afq
