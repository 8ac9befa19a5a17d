// Per-walker numerical status: one flag word per walker from its orbitals,
// overlap, weight and hybrid energy, and a two-word summary (OR of every
// flag, number of flagged walkers).  The reference lets NaN/inf propagate
// through the step and guards only the multi-determinant overlaps
// (observables/greens.py:264-276, greens_th4.py:17-44); these flags give a
// caller the same information per walker without scanning the arrays on the
// host.
#include "afq_kernels.h"

namespace afq {

namespace {

constexpr int STH = 256;

__device__ __forceinline__ bool finite2(cplx a) { return isfinite(a.x) && isfinite(a.y); }

// one workgroup per walker: scan phi_a and phi_b, then lane 0 adds the
// scalar checks
__global__ __launch_bounds__(STH) void k_walker_status(int W, long na, long nb,
                                                       const cplx* __restrict__ pa,
                                                       const cplx* __restrict__ pb,
                                                       const cplx* __restrict__ ov,
                                                       const double* __restrict__ w,
                                                       const cplx* __restrict__ eh,
                                                       double tiny,
                                                       uint32_t* __restrict__ flags) {
  const int wk = blockIdx.x;
  if (wk >= W) return;
  __shared__ int bad[STH / 64];
  int b = 0;
  const cplx* a = pa + (long)wk * na;
  for (long e = threadIdx.x; e < na; e += STH) b |= !finite2(a[e]);
  if (pb) {
    const cplx* c = pb + (long)wk * nb;
    for (long e = threadIdx.x; e < nb; e += STH) b |= !finite2(c[e]);
  }
  // wave OR, then across the four waves
  const unsigned long long m = __ballot(b);
  if ((threadIdx.x & 63) == 0) bad[threadIdx.x >> 6] = m != 0ull;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t f = 0;
    for (int i = 0; i < STH / 64; ++i)
      if (bad[i]) f |= AFQ_WALKER_NONFINITE_PHI;
    if (ov) {
      const cplx o = ov[wk];
      if (!finite2(o))
        f |= AFQ_WALKER_NONFINITE_OVERLAP;
      else if (hypot(o.x, o.y) <= tiny)
        f |= AFQ_WALKER_ZERO_OVERLAP;
    }
    if (w) {
      const double x = w[wk];
      if (!isfinite(x) || x < 0.0)
        f |= AFQ_WALKER_BAD_WEIGHT;
      else if (x == 0.0)
        f |= AFQ_WALKER_ZERO_WEIGHT;
    }
    if (eh && !finite2(eh[wk])) f |= AFQ_WALKER_NONFINITE_EHYBRID;
    flags[wk] = f;
  }
}

// one workgroup: summary[0] = OR of the flags (zero-weight walkers count as
// flagged only through their own bit), summary[1] = walkers with any flag
// other than AFQ_WALKER_ZERO_WEIGHT (a walker the phaseless projection
// removed is not a numerical failure)
__global__ __launch_bounds__(STH) void k_status_reduce(int W, const uint32_t* __restrict__ flags,
                                                       uint32_t* __restrict__ summary) {
  __shared__ uint32_t so[STH], sc[STH];
  uint32_t o = 0, c = 0;
  for (int i = threadIdx.x; i < W; i += STH) {
    const uint32_t f = flags[i];
    o |= f;
    c += (f & ~(uint32_t)AFQ_WALKER_ZERO_WEIGHT) != 0u;
  }
  so[threadIdx.x] = o;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int s = STH / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      so[threadIdx.x] |= so[threadIdx.x + s];
      sc[threadIdx.x] += sc[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    summary[0] = so[0];
    summary[1] = sc[0];
  }
}

}  // namespace

}  // namespace afq

extern "C" int afq_walker_status(afq_ctx* ctx, int W, int M, int Na, int Nb, const double* phia,
                                 const double* phib, const double* overlap,
                                 const double* weights, const double* e_hybrid,
                                 double zero_overlap_tol, uint32_t* flags, uint32_t* summary) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0, "bad shape");
  AFQ_REQUIRE(zero_overlap_tol >= 0.0, "zero_overlap_tol must be >= 0");
  AFQ_REQUIRE(W == 0 || (flags && (phia || Na == 0) && (phib || Nb == 0)), "NULL array");
  if (W > 0) {
    hipLaunchKernelGGL(afq::k_walker_status, dim3(W), dim3(afq::STH), 0, ctx->stream, W,
                       (long)M * Na, (long)M * Nb, (const cplx*)(Na ? phia : nullptr),
                       (const cplx*)(Nb ? phib : nullptr), (const cplx*)overlap, weights,
                       (const cplx*)e_hybrid, zero_overlap_tol, flags);
    AFQ_CHECK_LAUNCH();
  }
  if (summary) {
    hipLaunchKernelGGL(afq::k_status_reduce, dim3(1), dim3(afq::STH), 0, ctx->stream, W, flags,
                       summary);
    AFQ_CHECK_LAUNCH();
  }
  return AFQ_OK;
}
