// Comb population control and the driver's per-step weight glue
// (OmegaQMC/afqmc_gto.py:184-238, 1428-1432, 1448-1451, 1488-1495).
//
// Comb: ws = w / (sum w / W); c = running sum of ws; tooth j = j + r;
// idx_j = first i with c_i >= tooth (searchsorted 'left'), clipped to W-1.
// The total and the running sum are accumulated left to right by a single
// lane in FP64, so the indices are bit-for-bit those of a sequential host
// loop; the searches and the walker gather are parallel.
#include <mutex>

#include "afq_kernels.h"

// Running sums of afq_population_control_comb_noalloc when W does not fit in
// LDS: a fixed array in device memory (each device holds its own copy for
// the life of the process), so that entry point never allocates.  Host code
// reaches it only through the address hipGetSymbolAddress returns.
static __device__ __align__(256) double scratch[AFQ_COMB_NOALLOC_MAX_WALKERS];

namespace afq {

int comb_static_scratch(int device, double** out) {
  static std::mutex mu;
  static double* addr[64] = {};
  if (device < 0 || device >= 64) return fail(AFQ_EINVAL, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* p = nullptr;
    // the caller (afq_ctx_create) has made `device` current
    AFQ_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    addr[device] = (double*)p;
  }
  *out = addr[device];
  return AFQ_OK;
}

namespace {

constexpr int NT = 1024;

__global__ __launch_bounds__(NT) void k_comb_index(int W, const double* __restrict__ w, double r_arg,
                                                   const double* __restrict__ r_dev,
                                                   double* __restrict__ cum_g,
                                                   int32_t* __restrict__ idx,
                                                   double* __restrict__ w_new) {
  extern __shared__ double cum_s[];
  const double r = r_dev ? *r_dev : r_arg;
  const bool in_lds = (cum_g == nullptr);
  double* cum = in_lds ? cum_s : cum_g;
  for (int i = threadIdx.x; i < W; i += NT) cum[i] = w[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double total = 0.0;
    for (int i = 0; i < W; ++i) total += cum[i];
    const double scale = total / (double)W;
    double c = 0.0;
    for (int i = 0; i < W; ++i) {
      c += cum[i] / scale;
      cum[i] = c;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < W; j += NT) {
    const double tooth = (double)j + r;
    int lo = 0, hi = W;  // first index with cum >= tooth
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (cum[mid] < tooth) lo = mid + 1; else hi = mid;
    }
    idx[j] = lo > W - 1 ? W - 1 : lo;
    w_new[j] = 1.0;
  }
}

// One workgroup per destination walker: copy both spins and e_hybrid.
__global__ __launch_bounds__(256) void k_comb_gather(int W, long na, long nb,
                                                     const int32_t* __restrict__ idx,
                                                     const double4* __restrict__ pa,
                                                     const double4* __restrict__ pb,
                                                     const cplx* __restrict__ eh,
                                                     double4* __restrict__ pa_new,
                                                     double4* __restrict__ pb_new,
                                                     cplx* __restrict__ eh_new) {
  const long dst = blockIdx.x;
  const long src = idx[dst];
  const cplx* pac = (const cplx*)pa;
  const cplx* pbc = (const cplx*)pb;
  cplx* pan = (cplx*)pa_new;
  cplx* pbn = (cplx*)pb_new;
  if (src < 0 || src >= W) {
    // caller-supplied index out of range: poison the row instead of reading outside
    const cplx nan2 = make_double2(__builtin_nan(""), __builtin_nan(""));
    for (long i = threadIdx.x; i < na; i += 256) pan[dst * na + i] = nan2;
    for (long i = threadIdx.x; i < nb; i += 256) pbn[dst * nb + i] = nan2;
    if (eh && threadIdx.x == 0) eh_new[dst] = nan2;
    return;
  }
  for (long i = threadIdx.x; i < na; i += 256) pan[dst * na + i] = pac[src * na + i];
  for (long i = threadIdx.x; i < nb; i += 256) pbn[dst * nb + i] = pbc[src * nb + i];
  if (eh && threadIdx.x == 0) eh_new[dst] = eh[src];
}

__global__ __launch_bounds__(1024) void k_sum_abs(int W, const double* __restrict__ w,
                                                  double* __restrict__ out,
                                                  const int64_t* __restrict__ gate) {
  __shared__ double red[32];
  if (gate && *gate <= 1) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < W; i += 1024) s += fabs(w[i]);
  s = block_sum(s, red);
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_clip(int W, double* __restrict__ w, const double* __restrict__ total, double frac,
                       const int64_t* __restrict__ gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W || (gate && *gate <= 1)) return;
  const double b = *total * frac;
  w[i] = jmin(jmax(w[i], 0.0), b);
}

__global__ __launch_bounds__(1024) void k_accumulate(int W, const double* __restrict__ w,
                                                     const cplx* __restrict__ eh,
                                                     double* __restrict__ acc) {
  __shared__ double red[32];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < W; i += 1024) {
    double a = fabs(w[i]);
    s0 += a;
    s1 += a * eh[i].x;
  }
  s0 = block_sum(s0, red);
  s1 = block_sum(s1, red + 16);
  if (threadIdx.x == 0) {
    acc[0] += s0;
    acc[1] += s1;
  }
}

__global__ __launch_bounds__(1024) void k_block_stats(int W, const double* __restrict__ w,
                                                      const cplx* __restrict__ e,
                                                      double* __restrict__ out) {
  __shared__ double red[32];
  __shared__ double mx[16];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, m = -INFINITY;
  for (int i = threadIdx.x; i < W; i += 1024) {
    double a = fabs(w[i]);
    s0 += a;
    s1 += a * e[i].x;
    s2 += a * a;
    m = jmax(m, w[i]);
  }
  s0 = block_sum(s0, red);
  s1 = block_sum(s1, red + 16);
  double s2b = block_sum(s2, red);
  for (int off = 32; off > 0; off >>= 1) m = jmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = mx[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) mm = jmax(mm, mx[i]);
    out[0] = s1 / s0;
    out[1] = s0;
    out[2] = s2b;
    out[3] = mm;
  }
}

}  // namespace

__global__ void k_fill1(int W, double* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < W) w[i] = 1.0;
}

// Gather walkers by given survivor indices, weights reset to 1.
int gather_launch(hipStream_t s, int W, int M, int Na, int Nb, const int32_t* idx,
                  const double* phia, const double* phib, const double* eh, double* w_new,
                  double* phia_new, double* phib_new, double* eh_new) {
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_comb_gather, dim3(W), dim3(256), 0, s, W, (long)M * Na, (long)M * Nb, idx,
                     (const double4*)phia, (const double4*)phib, (const cplx*)eh,
                     (double4*)phia_new, (double4*)phib_new, (cplx*)eh_new);
  AFQ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_fill1, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, w_new);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int comb_launch(hipStream_t s, int W, int M, int Na, int Nb, const double* weights,
                const double* phia, const double* phib, double r, const double* eh,
                double* w_new, double* phia_new, double* phib_new, double* eh_new, int32_t* idx,
                double* cum_scratch, const double* r_dev) {
  if (W <= 0) return AFQ_OK;
  const bool lds = sizeof(double) * (size_t)W + 1024 <= max_lds_bytes();
  if (!lds && !cum_scratch) return fail(AFQ_EINVAL, "comb: W=%d needs a scratch buffer", W);
  const size_t sm = lds ? sizeof(double) * W : 0;
  if (lds) cum_scratch = nullptr;
  hipLaunchKernelGGL(k_comb_index, dim3(1), dim3(NT), sm, s, W, weights, r, r_dev, cum_scratch, idx,
                     w_new);
  AFQ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_comb_gather, dim3(W), dim3(256), 0, s, W, (long)M * Na, (long)M * Nb, idx,
                     (const double4*)phia, (const double4*)phib, (const cplx*)eh,
                     (double4*)phia_new, (double4*)phib_new, (cplx*)eh_new);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int clip_weights_launch(hipStream_t s, int W, double* weights, double frac, double* tmp,
                        const int64_t* gate) {
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_sum_abs, dim3(1), dim3(1024), 0, s, W, weights, tmp, gate);
  AFQ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_clip, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, weights, tmp, frac, gate);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int accumulate_launch(hipStream_t s, int W, const double* weights, const double* eh, double* acc) {
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(1024), 0, s, W, weights, (const cplx*)eh, acc);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int block_stats_launch(hipStream_t s, int W, const double* weights, const double* e, double* out) {
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_block_stats, dim3(1), dim3(1024), 0, s, W, weights, (const cplx*)e, out);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace afq

extern "C" {

int afq_population_control_apply(afq_ctx* ctx, int W, int M, int Na, int Nb, const int32_t* idx,
                                 const double* phia, const double* phib, const double* eh,
                                 double* w_new, double* phia_new, double* phib_new,
                                 double* eh_new) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0, "bad shape");
  AFQ_REQUIRE(idx && w_new && (Na == 0 || phia_new) && (Nb == 0 || phib_new), "NULL argument");
  AFQ_REQUIRE((Na == 0 || phia_new != phia) && (Nb == 0 || phib_new != phib),
              "outputs must not alias inputs");
  AFQ_REQUIRE(eh == nullptr || eh_new != nullptr, "e_hybrid_new is required with e_hybrid");
  return afq::gather_launch(ctx->stream, W, M, Na, Nb, idx, phia, phib, eh, w_new, phia_new,
                            phib_new, eh_new);
}

int afq_population_control_comb(afq_ctx* ctx, int W, int M, int Na, int Nb, const double* weights,
                                const double* phia, const double* phib, double r,
                                const double* eh, double* w_new, double* phia_new,
                                double* phib_new, double* eh_new, int32_t* idx) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0, "bad shape");
  AFQ_REQUIRE(r >= 0.0 && r < 1.0, "comb offset r=%g outside [0,1)", r);
  AFQ_REQUIRE((Na == 0 || phia_new != phia) && (Nb == 0 || phib_new != phib),
              "comb outputs must not alias inputs");
  AFQ_REQUIRE(idx != nullptr && w_new != nullptr, "idx and weights_new are required");
  AFQ_REQUIRE(eh == nullptr || eh_new != nullptr, "e_hybrid_new is required with e_hybrid");
  double* cum = nullptr;
  if (sizeof(double) * (size_t)W + 1024 > afq::max_lds_bytes()) {
    cum = (double*)afq::scratch(ctx, sizeof(double) * W);
    if (!cum) return AFQ_ENOMEM;
  }
  return afq::comb_launch(ctx->stream, W, M, Na, Nb, weights, phia, phib, r, eh, w_new, phia_new,
                          phib_new, eh_new, idx, cum);
}

int afq_population_control_comb_noalloc(afq_ctx* ctx, int W, int M, int Na, int Nb,
                                        const double* weights, const double* phia,
                                        const double* phib, double r, const double* eh,
                                        double* w_new, double* phia_new, double* phib_new,
                                        double* eh_new, int32_t* idx) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0, "bad shape");
  AFQ_REQUIRE(W <= AFQ_COMB_NOALLOC_MAX_WALKERS,
              "comb without allocation: %d walkers exceed the static limit of %d", W,
              AFQ_COMB_NOALLOC_MAX_WALKERS);
  AFQ_REQUIRE(r >= 0.0 && r < 1.0, "comb offset r=%g outside [0,1)", r);
  AFQ_REQUIRE((Na == 0 || phia_new != phia) && (Nb == 0 || phib_new != phib),
              "comb outputs must not alias inputs");
  AFQ_REQUIRE(idx != nullptr && w_new != nullptr, "idx and weights_new are required");
  AFQ_REQUIRE(eh == nullptr || eh_new != nullptr, "e_hybrid_new is required with e_hybrid");
  AFQ_REQUIRE(ctx->comb_static, "context has no static comb buffer");
  // every running sum is written before it is read: nothing to clear
  return afq::comb_launch(ctx->stream, W, M, Na, Nb, weights, phia, phib, r, eh, w_new, phia_new,
                          phib_new, eh_new, idx, ctx->comb_static);
}

int afq_clip_weights(afq_ctx* ctx, int W, double* weights, double frac) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0, "bad shape");
  double* tmp = (double*)afq::scratch(ctx, 256);
  if (!tmp) return AFQ_ENOMEM;
  return afq::clip_weights_launch(ctx->stream, W, weights, frac, tmp);
}

int afq_accumulate_hybrid(afq_ctx* ctx, int W, const double* weights, const double* eh,
                          double* acc) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0, "bad shape");
  return afq::accumulate_launch(ctx->stream, W, weights, eh, acc);
}

int afq_block_stats(afq_ctx* ctx, int W, const double* weights, const double* e, double* out) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W > 0, "bad shape");
  return afq::block_stats_launch(ctx->stream, W, weights, e, out);
}

}  // extern "C"
