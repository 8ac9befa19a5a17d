// Counter-based auxiliary fields: Philox4x32-10 (Salmon, Moraes, Dror,
// Shaw, SC'11) with Box-Muller.  Counter (g/2, w, step, 0), key = seed; two
// 53-bit uniforms per call, u1 in (0,1] and u2 in [0,1), give
// xi[w,2k] = r cos(2 pi u2), xi[w,2k+1] = r sin(2 pi u2), r = sqrt(-2 log u1).
// The comb offset uses counter (0, 0, step, 1).  One thread per (walker,
// pair); writes are coalesced pairs of doubles.
#include "afq_kernels.h"

namespace afq {

namespace {

struct u4 { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ u4 philox(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__global__ void k_aux(uint32_t k0, uint32_t k1, uint32_t step_arg, const int64_t* __restrict__ step_dev,
                      int W, int G, double* __restrict__ xi) {
  const int npair = (G + 1) / 2;
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= (long)W * npair) return;
  const uint32_t step = step_dev ? (uint32_t)*step_dev : step_arg;
  const int w = (int)(tid / npair), k = (int)(tid % npair);
  const u4 o = philox(u4{(uint32_t)k, (uint32_t)w, step, 0u}, k0, k1);
  const uint64_t a = ((uint64_t)o.y << 32) | o.x;
  const uint64_t b = ((uint64_t)o.w << 32) | o.z;
  const double u1 = (double)((a >> 11) + 1) * 0x1.0p-53;
  const double u2 = (double)(b >> 11) * 0x1.0p-53;
  const double r = sqrt(-2.0 * log(u1));
  const double th = 2.0 * M_PI * u2;
  double sn, cs;
  sincos(th, &sn, &cs);
  const long base = (long)w * G + 2 * k;
  xi[base] = r * cs;
  if (2 * k + 1 < G) xi[base + 1] = r * sn;
}

__device__ __host__ __forceinline__ double comb_uniform(uint64_t seed, uint32_t step) {
  const u4 o = philox(u4{0u, 0u, step, 1u}, (uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32));
  const uint64_t a = ((uint64_t)o.y << 32) | o.x;
  return (double)(a >> 11) * 0x1.0p-53;
}

// comb offsets of n seeds at the step *step_dev (one thread each)
__global__ void k_comb_r(int n, const uint64_t* __restrict__ seeds, const int64_t* __restrict__ step_dev,
                         double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = comb_uniform(seeds[i], (uint32_t)*step_dev);
}

}  // namespace

int aux_fields_launch(hipStream_t s, uint64_t seed, int64_t step, int W, int G, double* xi,
                      const int64_t* step_dev) {
  const long n = (long)W * ((G + 1) / 2);
  if (n <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_aux, dim3(cdiv_int(n, 256)), dim3(256), 0, s, (uint32_t)(seed & 0xffffffffu),
                     (uint32_t)(seed >> 32), (uint32_t)step, step_dev, W, G, xi);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int comb_offsets_dev_launch(hipStream_t s, int n, const uint64_t* seeds, const int64_t* step_dev,
                            double* r) {
  if (n <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_comb_r, dim3(cdiv_int(n, 64)), dim3(64), 0, s, n, seeds, step_dev, r);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

double comb_offset_host(uint64_t seed, int64_t step) { return comb_uniform(seed, (uint32_t)step); }

}  // namespace afq

extern "C" {

int afq_aux_fields(afq_ctx* ctx, uint64_t seed, int64_t step, int W, int G, double* xi) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && G >= 0, "bad shape");
  return afq::aux_fields_launch(ctx->stream, seed, step, W, G, xi);
}

double afq_comb_offset(uint64_t seed, int64_t step) { return afq::comb_offset_host(seed, step); }

}  // extern "C"
