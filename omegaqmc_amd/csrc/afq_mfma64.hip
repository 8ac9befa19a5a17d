// FP64 (v_mfma_f64_16x16x4f64) and FP32 (v_mfma_f32_16x16x4f32) MFMA kernels
// for the three dense contractions of the phaseless step
// (OmegaQMC/afqmc_gto.py:451-482, 682-683), templated on the scalar type T:
//
//   k_vhs<T>    VHS[w] = i sqrt(dt) sum_g x[w,g] L_g, as a real GEMM whose rows
//               are (walker, re/im) pairs and whose columns are the packed upper
//               triangle of the symmetric L_g (both triangles of VHS written).
//   k_fb<T>     vbias[w,g] = sum_k Ghalf[w,k] Lhalf[g,k], split over K with
//               per-split partial sums (FP64) reduced in a fixed order afterwards.
//   k_prop<T>   phi <- expH1 . sum_{n<=nmax} VHS^n/n! . expH1 . phi, fused: one
//               workgroup owns a 16-column block of one walker for the whole
//               chain, keeping that block in LDS between the 2 + nmax products.
//               The FP32 form applies the one-body half steps as y + D y with
//               D = expH1 - I, y in the accumulators.
//
// Walkers, Ghalf and the shifted fields stay FP64 in HBM; the FP32 kernels
// convert them on load and write FP64 walkers.
//
// MFMA operand layout (gfx950, both checked on the device): A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k; D[i][j] in lane j + 16q, register r with
// i = q + 4r (FP64) or i = 4q + r (FP32).  GEMM rows that pair (walker re,
// walker im) are ordered (Mf<T>::prow) so that a lane's four output registers
// are (w, re), (w, im), (w', re), (w', im).
//
// Workgroups are renumbered so that the hardware's round-robin assignment of
// consecutive workgroups to the 8 XCDs puts tiles that share an operand
// (same L column tile / same walker) on one XCD, where they share its L2.
#include "afq_mfma.h"

namespace afq {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <class T> struct Mf;
template <> struct Mf<double> {
  typedef d4v V4;
  typedef cplx C;
  static constexpr bool DFORM = false;
  __device__ static V4 mma(double a, double b, V4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // tile row held by lane group q in register r
  __device__ static int orow(int q, int r) { return q + 4 * r; }
  // row of component c (0 re, 1 im) of walker wi (0..7) of a 16-row tile
  __device__ static int prow(int wi, int c) { return (wi & 3) + 8 * (wi >> 2) + 4 * c; }
  // walker (0..7 of the tile) in registers 2h, 2h + 1 of lane group q
  __device__ static int wlane(int q, int h) { return q + 4 * h; }
  __device__ static C mk(double r, double i) { return mkc(r, i); }
};
template <> struct Mf<float> {
  typedef f4v V4;
  typedef float2 C;
  static constexpr bool DFORM = true;
  __device__ static V4 mma(float a, float b, V4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int orow(int q, int r) { return 4 * q + r; }
  __device__ static int prow(int wi, int c) { return 2 * wi + c; }
  __device__ static int wlane(int q, int h) { return 2 * q + h; }
  __device__ static C mk(double r, double i) { return make_float2((float)r, (float)i); }
};

// Virtual block index such that consecutive virtual indices run on one XCD.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int x = b & 7, slot = b >> 3;
  const int per = nb >> 3, rem = nb & 7;
  return (x < rem) ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}

constexpr int TM = 128, TN = 128, TK = 16, NTH = 256;
constexpr int PNW = 8, PTH = 64 * PNW;  // waves / threads of the fused propagator

// --------------------------------------------------------------------------
// VHS build
// --------------------------------------------------------------------------
template <class T, bool PACKED>
__global__ __launch_bounds__(NTH) void k_vhs(int W, int G, int P, long MM,
                                             const cplx* __restrict__ xs,
                                             const T* __restrict__ L,
                                             const int2* __restrict__ colmap, double sqdt,
                                             typename Mf<T>::C* __restrict__ VHS, int ntm,
                                             int ntn) {
  typedef Mf<T> F;
  __shared__ T As[TK][TM + 4];
  __shared__ T Bs[TK][TN + 4];
  const int v = xcd_remap(blockIdx.x, ntm * ntn);
  const int tm = v % ntm, tn = v / ntm;
  const int w0 = tm * (TM / 2), n0 = tn * TN;
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int wr = wv >> 1, wc = wv & 1;

  cplx ra[4];
  T rb[8];
  auto load = [&](int g0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = t + NTH * i, wl = e >> 4, gk = e & 15;
      const int w = w0 + wl, g = g0 + gk;
      ra[i] = (w < W && g < G) ? xs[(long)w * G + g] : mkc(0.0, 0.0);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = t + NTH * i, kk = e >> 7, nn = e & 127;
      const int g = g0 + kk, n = n0 + nn;
      rb[i] = (g < G && n < P) ? L[(long)g * P + n] : T(0);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = t + NTH * i, wl = e >> 4, gk = e & 15;
      As[gk][16 * (wl >> 3) + F::prow(wl & 7, 0)] = (T)ra[i].x;
      As[gk][16 * (wl >> 3) + F::prow(wl & 7, 1)] = (T)ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = t + NTH * i;
      Bs[e >> 7][e & 127] = rb[i];
    }
  };

  typename F::V4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = typename F::V4{0, 0, 0, 0};

  const int nk = (G + TK - 1) / TK;
  load(0);
  store();
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load((kt + 1) * TK);
#pragma unroll
    for (int ks = 0; ks < TK / 4; ++ks) {
      const int kl = 4 * ks + (l >> 4);
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kl][64 * wr + 16 * i + (l & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kl][64 * wc + 16 * j + (l & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = F::mma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
    if (kt + 1 < nk) {
      store();
      __syncthreads();
    }
  }

  const int q = l >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int tile = 4 * wr + i;
    const int wA = w0 + 8 * tile + F::wlane(q, 0), wB = w0 + 8 * tile + F::wlane(q, 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + 64 * wc + 16 * j + (l & 15);
      if (n >= P) continue;
      long o1 = n, o2 = -1;
      if (PACKED) {
        const int2 cm = colmap[n];
        o1 = cm.x;
        o2 = cm.y;
      }
      const auto c = acc[i][j];
      if (wA < W) {
        const auto val = F::mk(-sqdt * c[1], sqdt * c[0]);
        VHS[wA * MM + o1] = val;
        if (o2 >= 0) VHS[wA * MM + o2] = val;
      }
      if (wB < W) {
        const auto val = F::mk(-sqdt * c[3], sqdt * c[2]);
        VHS[wB * MM + o1] = val;
        if (o2 >= 0) VHS[wB * MM + o2] = val;
      }
    }
  }
}

// --------------------------------------------------------------------------
// force bias, split K
// --------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(NTH) void k_fb(int W, int G, FbSplits sp, cplx* __restrict__ part,
                                            int ntm, int ntn) {
  typedef Mf<T> F;
  __shared__ T As[TK][TM + 4];
  __shared__ T Bs[TK][TN + 4];
  const int s = blockIdx.y;
  const int v = xcd_remap(blockIdx.x, ntm * ntn);
  const int tm = v % ntm, tn = v / ntm;
  const int w0 = tm * (TM / 2), n0 = tn * TN;
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int wr = wv >> 1, wc = wv & 1;
  const cplx* __restrict__ Gh = sp.gh[s];
  const T* __restrict__ Lh = (const T*)sp.lh[s];
  const long K = sp.K[s];
  const int kb = sp.kbeg[s], ke = sp.kend[s];

  cplx ra[4];
  T rb[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = t + NTH * i, wl = e >> 4, kk = e & 15;
      const int w = w0 + wl, k = k0 + kk;
      ra[i] = (w < W && k < ke) ? Gh[(long)w * K + k] : mkc(0.0, 0.0);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = t + NTH * i, gl = e >> 4, kk = e & 15;
      const int g = n0 + gl, k = k0 + kk;
      rb[i] = (g < G && k < ke) ? Lh[(long)g * K + k] : T(0);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = t + NTH * i, wl = e >> 4, kk = e & 15;
      As[kk][16 * (wl >> 3) + F::prow(wl & 7, 0)] = (T)ra[i].x;
      As[kk][16 * (wl >> 3) + F::prow(wl & 7, 1)] = (T)ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = t + NTH * i;
      Bs[e & 15][e >> 4] = rb[i];
    }
  };

  typename F::V4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = typename F::V4{0, 0, 0, 0};

  const int nk = (ke - kb + TK - 1) / TK;
  if (nk > 0) {
    load(kb);
    store();
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load(kb + (kt + 1) * TK);
#pragma unroll
    for (int ks = 0; ks < TK / 4; ++ks) {
      const int kl = 4 * ks + (l >> 4);
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kl][64 * wr + 16 * i + (l & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kl][64 * wc + 16 * j + (l & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = F::mma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
    if (kt + 1 < nk) {
      store();
      __syncthreads();
    }
  }

  const int q = l >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int tile = 4 * wr + i;
    const int wA = w0 + 8 * tile + F::wlane(q, 0), wB = w0 + 8 * tile + F::wlane(q, 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int g = n0 + 64 * wc + 16 * j + (l & 15);
      if (g >= G) continue;
      const auto c = acc[i][j];
      if (wA < W) part[((long)s * W + wA) * G + g] = mkc(c[0], c[1]);
      if (wB < W) part[((long)s * W + wB) * G + g] = mkc(c[2], c[3]);
    }
  }
}

// --------------------------------------------------------------------------
// fused one-body / Taylor / one-body on a 16-column block of one walker
// --------------------------------------------------------------------------
// E == nullptr skips both one-body products (plain _apply_exp_vhs).  The
// output may alias the input: a workgroup reads and writes only its own
// columns, reading all of them before writing any.  With Mf<T>::DFORM the
// matrix E is expH1 - I and each one-body product is acc = y + E y.
// SCALED: walker w's matrix is scoef[w] * smat (both real, (W,) and (M, M))
// instead of V[w].
template <class T, int MT, bool SCALED = false>
__global__ __launch_bounds__(PTH) void k_prop(int W, int M, int Na, int Nb, int ncb,
                                              const T* __restrict__ E,
                                              const typename Mf<T>::C* __restrict__ V,
                                              const cplx* pa_in, const cplx* pb_in,
                                              cplx* pa_out, cplx* pb_out, int nmax,
                                              const double* __restrict__ smat = nullptr,
                                              const double* __restrict__ scoef = nullptr) {
  typedef Mf<T> F;
  typedef typename F::C C;
  typedef typename F::V4 V4;
  extern __shared__ __align__(16) char smem_raw[];
  C* Ts = (C*)smem_raw;  // [Mpad][16]
  const int RT = (M + 15) >> 4, Mpad = RT * 16;
  const int vb = xcd_remap(blockIdx.x, W * ncb);
  const int w = vb / ncb, cb = vb % ncb;
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int li = l & 15, lq = l >> 4;
  const int Nt = Na + Nb;
  const long MM = (long)M * M;
  const C* __restrict__ Vw = SCALED ? nullptr : V + (long)w * MM;
  const double sc = SCALED ? scoef[w] : 0.0;

  // element offset of (row, column c of this block) in the spin array, or -1
  auto coloff = [&](int c, long row, bool* spin_b) -> long {
    const int cg = 16 * cb + c;
    *spin_b = cg >= Na;
    if (cg < Na) return ((long)w * M + row) * Na + cg;
    if (cg < Nt) return ((long)w * M + row) * Nb + (cg - Na);
    return -1;
  };

  // stage the input block
  for (int e = t; e < Mpad * 16; e += PTH) {
    const int q = e >> 4, c = e & 15;
    bool sb = false;
    const long o = (q < M) ? coloff(c, q, &sb) : -1;
    const cplx x = (o >= 0) ? (sb ? pb_in[o] : pa_in[o]) : mkc(0.0, 0.0);
    Ts[e] = F::mk(x.x, x.y);
  }
  __syncthreads();

  V4 are[MT], aim[MT], rre[MT], rim[MT];
  auto owns = [&](int u) { return wv + PNW * u < RT; };
  auto arow = [&](int u) { return 16 * (wv + PNW * u) + li; };
  auto orow = [&](int u, int r) { return 16 * (wv + PNW * u) + F::orow(lq, r); };

  // acc = E (real, M x M) . Ts (DFORM: acc = Ts + E . Ts); the A fragments
  // of the next k-step are loaded before the MFMAs of the current one.
  auto onebody = [&]() {
#pragma unroll
    for (int u = 0; u < MT; ++u) {
      if (F::DFORM) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = orow(u, r);
          const C x = (owns(u) && row < M) ? Ts[row * 16 + li] : F::mk(0.0, 0.0);
          are[u][r] = x.x;
          aim[u][r] = x.y;
        }
      } else {
        are[u] = V4{0, 0, 0, 0};
        aim[u] = V4{0, 0, 0, 0};
      }
    }
    T a0[MT], a1[MT];
    auto ld = [&](T* a, int k0) {
      const int k = k0 + lq;
#pragma unroll
      for (int u = 0; u < MT; ++u) {
        const int row = arow(u);
        a[u] = (row < M && k < M) ? E[(long)row * M + k] : T(0);
      }
    };
    auto mm = [&](const T* a, int k0) {
      const C b = Ts[(k0 + lq) * 16 + li];
#pragma unroll
      for (int u = 0; u < MT; ++u)
        if (owns(u)) are[u] = F::mma(a[u], b.x, are[u]);
#pragma unroll
      for (int u = 0; u < MT; ++u)
        if (owns(u)) aim[u] = F::mma(a[u], b.y, aim[u]);
    };
    ld(a0, 0);
    for (int k0 = 0; k0 < Mpad; k0 += 8) {
      ld(a1, k0 + 4);
      mm(a0, k0);
      if (k0 + 8 < Mpad) ld(a0, k0 + 8);
      mm(a1, k0 + 4);
    }
  };
  // acc = V_w (complex) . Ts, same prefetch scheme
  auto taylor_term = [&]() {
#pragma unroll
    for (int u = 0; u < MT; ++u) { are[u] = V4{0, 0, 0, 0}; aim[u] = V4{0, 0, 0, 0}; }
    C a0[MT], a1[MT];
    auto ld = [&](C* a, int k0) {
      const int k = k0 + lq;
#pragma unroll
      for (int u = 0; u < MT; ++u) {
        const int row = arow(u);
        if (SCALED)
          a[u] = (row < M && k < M) ? F::mk(sc * smat[(long)row * M + k], 0.0) : F::mk(0.0, 0.0);
        else
          a[u] = (row < M && k < M) ? Vw[(long)row * M + k] : F::mk(0.0, 0.0);
      }
    };
    auto mm = [&](const C* a, int k0) {
      const C b = Ts[(k0 + lq) * 16 + li];
#pragma unroll
      for (int u = 0; u < MT; ++u)
        if (owns(u)) are[u] = F::mma(a[u].x, b.x, are[u]);
#pragma unroll
      for (int u = 0; u < MT; ++u)
        if (owns(u)) aim[u] = F::mma(a[u].x, b.y, aim[u]);
#pragma unroll
      for (int u = 0; u < MT; ++u)
        if (owns(u)) are[u] = F::mma(-a[u].y, b.y, are[u]);
#pragma unroll
      for (int u = 0; u < MT; ++u)
        if (owns(u)) aim[u] = F::mma(a[u].y, b.x, aim[u]);
    };
    ld(a0, 0);
    for (int k0 = 0; k0 < Mpad; k0 += 8) {
      ld(a1, k0 + 4);
      mm(a0, k0);
      if (k0 + 8 < Mpad) ld(a0, k0 + 8);
      mm(a1, k0 + 4);
    }
  };
  // write acc rows (< M) into Ts
  auto to_lds = [&](const V4* xr, const V4* xi) {
#pragma unroll
    for (int u = 0; u < MT; ++u) {
      if (!owns(u)) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = orow(u, r);
        if (row < M) Ts[row * 16 + li] = C{xr[u][r], xi[u][r]};
      }
    }
  };

  if (E) {
    onebody();
    __syncthreads();
    to_lds(are, aim);
  } else {
    // the result starts as the input block itself
#pragma unroll
    for (int u = 0; u < MT; ++u) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = orow(u, r);
        const C x = owns(u) ? Ts[row * 16 + li] : F::mk(0.0, 0.0);
        are[u][r] = x.x;
        aim[u][r] = x.y;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < MT; ++u) { rre[u] = are[u]; rim[u] = aim[u]; }
  __syncthreads();
  for (int n = 1; n <= nmax; ++n) {
    taylor_term();
    const T dn = (T)n;
#pragma unroll
    for (int u = 0; u < MT; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        are[u][r] = are[u][r] / dn;
        aim[u][r] = aim[u][r] / dn;
        rre[u][r] = rre[u][r] + are[u][r];
        rim[u][r] = rim[u][r] + aim[u][r];
      }
    __syncthreads();
    to_lds(are, aim);
    __syncthreads();
  }
  if (E) {
    to_lds(rre, rim);
    __syncthreads();
    onebody();
  } else {
#pragma unroll
    for (int u = 0; u < MT; ++u) { are[u] = rre[u]; aim[u] = rim[u]; }
  }
#pragma unroll
  for (int u = 0; u < MT; ++u) {
    if (!owns(u)) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = orow(u, r);
      bool sb = false;
      const long o = (row < M) ? coloff(li, row, &sb) : -1;
      if (o >= 0) (sb ? pb_out : pa_out)[o] = mkc(are[u][r], aim[u][r]);
    }
  }
}

}  // namespace

template <class T>
static int vhs_launch(hipStream_t s, int W, int M, int G, const double* xs, const T* L, int P,
                      const int2* colmap, double dt, void* VHS) {
  if (W <= 0) return AFQ_OK;
  const int ntm = cdiv_int(2L * W, TM), ntn = cdiv_int(P, TN);
  const int nb = ntm * ntn;
  typedef typename Mf<T>::C C;
  if (colmap)
    hipLaunchKernelGGL((k_vhs<T, true>), dim3(nb), dim3(NTH), 0, s, W, G, P, (long)M * M,
                       (const cplx*)xs, L, colmap, sqrt(dt), (C*)VHS, ntm, ntn);
  else
    hipLaunchKernelGGL((k_vhs<T, false>), dim3(nb), dim3(NTH), 0, s, W, G, P, (long)M * M,
                       (const cplx*)xs, L, colmap, sqrt(dt), (C*)VHS, ntm, ntn);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int vhs_mfma_f64(hipStream_t s, int W, int M, int G, const double* xs, const double* L, int P,
                 const int2* colmap, double dt, double* VHS) {
  return vhs_launch<double>(s, W, M, G, xs, L, P, colmap, dt, VHS);
}

int vhs_mfma_f32(hipStream_t s, int W, int M, int G, const double* xs, const float* L, int P,
                 const int2* colmap, double dt, float* VHS) {
  return vhs_launch<float>(s, W, M, G, xs, L, P, colmap, dt, VHS);
}

FbSplits fb_splits(int M, int Na, int Nb, const double* Gha, const double* Ghb, const void* Lha,
                   const void* Lhb, int per_spin) {
  FbSplits sp{};
  int n = 0;
  const int Ns[2] = {Na, Nb};
  const double* gh[2] = {Gha, Ghb};
  const void* lh[2] = {Lha, Lhb};
  for (int s = 0; s < 2; ++s) {
    const long K = (long)Ns[s] * M;
    if (K == 0) continue;
    const long chunk = ((K + per_spin - 1) / per_spin + TK - 1) / TK * TK;
    for (long k0 = 0; k0 < K && n < FB_MAX_SPLITS; k0 += chunk) {
      sp.gh[n] = (const cplx*)gh[s];
      sp.lh[n] = lh[s];
      sp.K[n] = K;
      sp.kbeg[n] = (int)k0;
      sp.kend[n] = (int)std::min(K, k0 + chunk);
      ++n;
    }
  }
  sp.n = n;
  return sp;
}

template <class T>
static int fb_launch(hipStream_t s, int W, int G, const FbSplits& sp, double* part) {
  if (W <= 0 || G <= 0 || sp.n == 0) return AFQ_OK;
  const int ntm = cdiv_int(2L * W, TM), ntn = cdiv_int(G, TN);
  hipLaunchKernelGGL(k_fb<T>, dim3(ntm * ntn, sp.n), dim3(NTH), 0, s, W, G, sp, (cplx*)part,
                     ntm, ntn);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int fb_mfma_f64(hipStream_t s, int W, int G, const FbSplits& sp, double* part) {
  return fb_launch<double>(s, W, G, sp, part);
}

int fb_mfma_f32(hipStream_t s, int W, int G, const FbSplits& sp, double* part) {
  return fb_launch<float>(s, W, G, sp, part);
}

size_t prop_f64_smem(int M) { return (size_t)((M + 15) / 16) * 16 * 16 * sizeof(cplx); }
bool prop_f64_fused_fits(int M) { return M <= 512 && prop_f64_smem(M) <= max_lds_bytes(); }

template <class T>
static int prop_launch(hipStream_t s, int W, int M, int Na, int Nb, const T* E, const void* VHS,
                       const double* pa_in, const double* pb_in, double* pa_out, double* pb_out,
                       int nmax) {
  const int Nt = Na + Nb;
  if (W <= 0 || Nt == 0) return AFQ_OK;
  typedef typename Mf<T>::C C;
  const int RT = (M + 15) / 16;
  const int ncb = (Nt + 15) / 16;
  const size_t sm = (size_t)RT * 16 * 16 * sizeof(C);
  if (sm > max_lds_bytes())
    return fail(AFQ_EUNSUPPORTED, "fused propagator: M=%d needs %zu bytes of LDS", M, sm);
  const int nb = W * ncb;
  const dim3 grid(nb), blk(PTH);
#define AFQ_LAUNCH_PROP(MT)                                                                 \
  hipLaunchKernelGGL((k_prop<T, MT>), grid, blk, sm, s, W, M, Na, Nb, ncb, E,              \
                     (const C*)VHS, (const cplx*)pa_in, (const cplx*)pb_in, (cplx*)pa_out, \
                     (cplx*)pb_out, nmax)
  if (RT <= PNW) AFQ_LAUNCH_PROP(1);
  else if (RT <= 2 * PNW) AFQ_LAUNCH_PROP(2);
  else if (RT <= 3 * PNW) AFQ_LAUNCH_PROP(3);
  else if (RT <= 4 * PNW) AFQ_LAUNCH_PROP(4);
  else return fail(AFQ_EUNSUPPORTED, "fused propagator: M=%d > 512", M);
#undef AFQ_LAUNCH_PROP
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int prop_mfma_f64(hipStream_t s, int W, int M, int Na, int Nb, const double* expH1,
                  const double* VHS, const double* pa_in, const double* pb_in, double* pa_out,
                  double* pb_out, int nmax) {
  return prop_launch<double>(s, W, M, Na, Nb, expH1, VHS, pa_in, pb_in, pa_out, pb_out, nmax);
}

int prop_scaled_f64(hipStream_t s, int W, int M, int N, const double* mat, const double* coef,
                    const double* phi_in, double* phi_out, int nmax) {
  if (W <= 0 || N == 0) return AFQ_OK;
  const int RT = (M + 15) / 16;
  const int ncb = (N + 15) / 16;
  const size_t sm = (size_t)RT * 16 * 16 * sizeof(cplx);
  if (sm > max_lds_bytes())
    return fail(AFQ_EUNSUPPORTED, "scaled propagator: M=%d needs %zu bytes of LDS", M, sm);
  const dim3 grid(W * ncb), blk(PTH);
#define AFQ_LAUNCH_SC(MT)                                                                      \
  hipLaunchKernelGGL((k_prop<double, MT, true>), grid, blk, sm, s, W, M, N, 0, ncb,              \
                     (const double*)nullptr, (const cplx*)nullptr, (const cplx*)phi_in,          \
                     (const cplx*)nullptr, (cplx*)phi_out, (cplx*)nullptr, nmax, mat, coef)
  if (RT <= PNW) AFQ_LAUNCH_SC(1);
  else if (RT <= 2 * PNW) AFQ_LAUNCH_SC(2);
  else if (RT <= 3 * PNW) AFQ_LAUNCH_SC(3);
  else if (RT <= 4 * PNW) AFQ_LAUNCH_SC(4);
  else return fail(AFQ_EUNSUPPORTED, "scaled propagator: M=%d > 512", M);
#undef AFQ_LAUNCH_SC
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int prop_mfma_f32(hipStream_t s, int W, int M, int Na, int Nb, const float* expH1_minus_I,
                  const float* VHS, const double* pa_in, const double* pb_in, double* pa_out,
                  double* pb_out, int nmax) {
  return prop_launch<float>(s, W, M, Na, Nb, expH1_minus_I, VHS, pa_in, pb_in, pa_out, pb_out,
                            nmax);
}

}  // namespace afq
