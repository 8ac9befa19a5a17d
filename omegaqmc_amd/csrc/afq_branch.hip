// Pair-branching population control (OmegaQMC/afqmc_gto.py:241-284).
//
// The walkers are ordered by weight and the k-th lightest is paired with the
// k-th heaviest (k < W/2); when the heavy weight exceeds 4 <w> and the light
// one is below 0.1 <w> (<w> = sum w / W), the light walker becomes a copy of
// the heavy one and both take the mean of the two weights.  The pairs are
// disjoint, so every pair is decided independently.
//
// One workgroup orders (weight, index) keys in LDS by a bitonic sort: ties
// keep index order, i.e. a stable argsort.  The total weight is summed left
// to right by one lane, as a sequential host loop does.  The copies then run
// in place: a walker that is copied from is never written.
#include "afq_kernels.h"

namespace afq {

namespace {

constexpr int PB_NT = 1024;

// order-preserving map of a double onto uint64 (negative values reversed)
__device__ __forceinline__ unsigned long long ord_key(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// w (W,) in place; src (W,) receives the walker each walker copies (itself
// when unchanged).  LDS: Wp keys and indices, Wp = next power of two >= W.
__global__ __launch_bounds__(PB_NT) void k_pair_branch(int W, int Wp, double* __restrict__ w,
                                                       int32_t* __restrict__ src) {
  extern __shared__ __align__(16) unsigned char pb_lds[];
  unsigned long long* key = (unsigned long long*)pb_lds;
  int* id = (int*)(key + Wp);
  __shared__ double avg;
  const int t = threadIdx.x;
  for (int i = t; i < Wp; i += PB_NT) {
    key[i] = i < W ? ord_key(w[i]) : ~0ull;
    id[i] = i;
  }
  if (t == 0) {
    double s = 0.0;
    for (int i = 0; i < W; ++i) s += w[i];
    avg = s / (double)W;
  }
  __syncthreads();
  // bitonic sort, ascending in (key, index)
  for (int k = 2; k <= Wp; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < Wp; i += PB_NT) {
        const int l = i ^ j;
        if (l > i) {
          const unsigned long long ki = key[i], kl = key[l];
          const int ii = id[i], il = id[l];
          const bool gt = ki > kl || (ki == kl && ii > il);
          if (((i & k) == 0) == gt) {
            key[i] = kl; key[l] = ki;
            id[i] = il; id[l] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
  const double wmin = 0.1 * avg, wmax = 4.0 * avg;
  for (int k = t; k < W - W / 2; k += PB_NT) {
    const int il = id[k], ih = id[W - 1 - k];
    if (k < W / 2) {
      const double wl = w[il], wh = w[ih];
      if (wh > wmax && wl < wmin) {
        const double m = 0.5 * (wl + wh);
        w[il] = m;
        w[ih] = m;
        src[il] = ih;
      } else {
        src[il] = il;
      }
      src[ih] = ih;
    } else {
      src[il] = il;  // the middle walker of an odd count
    }
  }
}

// phi[i] <- phi[src[i]] for every walker with src[i] != i (M * N complex each)
__global__ void k_branch_copy(int W, long MN, const int32_t* __restrict__ src,
                              double2* __restrict__ phi) {
  const long n = (long)W * MN;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / MN);
    const int s = src[i];
    if (s != i) phi[e] = phi[(long)s * MN + e % MN];
  }
}

}  // namespace

size_t pair_branch_lds(int W) {
  int Wp = 1;
  while (Wp < W) Wp <<= 1;
  return (size_t)Wp * (sizeof(unsigned long long) + sizeof(int));
}

}  // namespace afq

extern "C" {

int afq_population_control_pair_branch(afq_ctx* ctx, int W, int M, int Na, int Nb,
                                       double* weights, double* phia, double* phib,
                                       int32_t* src) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0, "bad shape");
  AFQ_REQUIRE(W <= AFQ_PAIR_BRANCH_MAX_WALKERS, "pair branching: %d walkers exceed %d", W,
              AFQ_PAIR_BRANCH_MAX_WALKERS);
  AFQ_REQUIRE(weights && (Na == 0 || phia) && (Nb == 0 || phib), "NULL argument");
  if (W == 0) return AFQ_OK;
  hipStream_t s = ctx->stream;
  int32_t* sidx = src;
  if (!sidx) {
    sidx = (int32_t*)afq::scratch(ctx, sizeof(int32_t) * (size_t)W);
    if (!sidx) return AFQ_ENOMEM;
  }
  int Wp = 1;
  while (Wp < W) Wp <<= 1;
  const size_t sm = afq::pair_branch_lds(W);
  if (sm > afq::max_lds_bytes()) return afq::fail(AFQ_EUNSUPPORTED, "pair branching: LDS %zu", sm);
  hipLaunchKernelGGL(afq::k_pair_branch, dim3(1), dim3(afq::PB_NT), sm, s, W, Wp, weights, sidx);
  AFQ_CHECK_LAUNCH();
  const long mn[2] = {(long)M * Na, (long)M * Nb};
  double* phi[2] = {phia, phib};
  for (int sp = 0; sp < 2; ++sp) {
    if (mn[sp] == 0) continue;
    const long n = (long)W * mn[sp];
    const int blocks = (int)std::min<long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(afq::k_branch_copy, dim3(blocks), dim3(256), 0, s, W, mn[sp], sidx,
                       (double2*)phi[sp]);
    AFQ_CHECK_LAUNCH();
  }
  return AFQ_OK;
}

}  // extern "C"
