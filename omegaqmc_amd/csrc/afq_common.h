// Shared helpers of the gfx950 AFQMC kernels: complex128 arithmetic on
// double2, the context/system structs and launch-checking macros.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>

#include "../../include/omega_afqmc.h"

// ---------------------------------------------------------------------------
// complex128 as double2 (x = re, y = im)
// ---------------------------------------------------------------------------
typedef double2 cplx;

__host__ __device__ __forceinline__ cplx mkc(double r, double i) { return make_double2(r, i); }
__host__ __device__ __forceinline__ cplx cadd(cplx a, cplx b) { return mkc(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ cplx csub(cplx a, cplx b) { return mkc(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ cplx cscale(cplx a, double s) { return mkc(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ cplx cconj(cplx a) { return mkc(a.x, -a.y); }
// a += b * c
__host__ __device__ __forceinline__ void cfma(cplx& a, cplx b, cplx c) {
  a.x = fma(b.x, c.x, a.x);
  a.x = fma(-b.y, c.y, a.x);
  a.y = fma(b.x, c.y, a.y);
  a.y = fma(b.y, c.x, a.y);
}
// a -= b * c
__host__ __device__ __forceinline__ void cfms(cplx& a, cplx b, cplx c) {
  a.x = fma(-b.x, c.x, a.x);
  a.x = fma(b.y, c.y, a.x);
  a.y = fma(-b.x, c.y, a.y);
  a.y = fma(-b.y, c.x, a.y);
}
__device__ __forceinline__ double cabs_(cplx a) { return hypot(a.x, a.y); }
// Smith's algorithm for a / b.
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  if (fabs(b.y) <= fabs(b.x)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return mkc((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.x * r + b.y;
    return mkc((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
// bf16 split of an FP64 value: hi = bf16(x), lo = bf16(x - hi), round to nearest even.
typedef unsigned short bf16_t;
__device__ __forceinline__ bf16_t bf_bits(float x) { return __builtin_bit_cast(bf16_t, (__bf16)x); }
__device__ __forceinline__ float bf_val(bf16_t b) { return (float)__builtin_bit_cast(__bf16, b); }
// FP64 -> FP32 (nearest), kept as its own step: the compiler otherwise folds
// the two conversions of split2 into one exact FP64 -> BF16 rounding, which
// it emulates with about eight more instructions (three of them FP64).
__device__ __forceinline__ float f32_of(double x) {
  float f = (float)x;
  asm("" : "+v"(f));
  return f;
}
// hi = bf16(f32(x)), lo = bf16(f32(x - hi)), each rounding to nearest even
__device__ __forceinline__ void split2(double x, bf16_t& h, bf16_t& l) {
  h = bf_bits(f32_of(x));
  l = bf_bits(f32_of(x - (double)bf_val(h)));
}

// Split-bf16 copies of Ghalf written by the Green's-function kernel: plane
// hi at a, lo at a + 2 W Ka; row (w, re/im) at (2 w + c) Ka; element (i, q)
// at i * ld + q (ld >= M; the padding q in [M, ld) is never written).
struct GhalfBf16 {
  bf16_t* a;
  bf16_t* b;
  long Ka, Kb;
  int ld;
};

// NaN-propagating max/min (jnp.maximum / jnp.minimum semantics).
__device__ __forceinline__ double jmax(double a, double b) {
  return (a != a || b != b) ? (a + b) : (a > b ? a : b);
}
__device__ __forceinline__ double jmin(double a, double b) {
  return (a != a || b != b) ? (a + b) : (a < b ? a : b);
}

// ---------------------------------------------------------------------------
// context / system
// ---------------------------------------------------------------------------
struct afq_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // device address of the comb's static running-sum array on this device
  // (afq_population_control_comb_noalloc), looked up at context creation
  double* comb_static = nullptr;
};

namespace afq {
struct MdRatio;
}

struct afq_system {
  afq_ctx* ctx = nullptr;
  int M = 0, Na = 0, Nb = 0, G = 0;
  const double* chol = nullptr;
  const double* rchol_a = nullptr;
  const double* rchol_b = nullptr;
  const double* expH1 = nullptr;
  const double* mf_shift = nullptr;
  const double* trial_up = nullptr;
  const double* trial_dn = nullptr;
  // optional sign of each Cholesky vector in the ERI, (pq|rs) = sum_g
  // sign_g L_g,pq L_g,rs (+1 symmetric, -1 antisymmetric vectors of the
  // plane-wave electron gas; caller-owned, device, (G,)); NULL: all +1
  const double* chol_sign = nullptr;
  double dt = 0.005, fbbound = 1.0;
  int nmax = 6;
  int precision = AFQ_FP64;
  // reduced-precision operand copies (owned)
  void* chol_lp = nullptr;     // precision-specific packed L
  void* rchol_lp = nullptr;    // precision-specific packed Lhalf (both spins)
  // FP64 VHS operand: upper triangle of every L_g (G x P) and its column map
  double* chol_packed = nullptr;
  int2* colmap = nullptr;
  int packed_P = 0;
  int fb_nsplit = 1;
  // greens mode bits: 2 if trial_up == eye(M, Na) exactly, 4 for trial_dn
  int eye_mask = 0;
  // step workspace (owned), sized for max_walkers
  int cap_walkers = 0;
  void* work = nullptr;
  size_t work_bytes = 0;
  // optional per-phase HIP-event timing of the fused step
  struct Timing* timing = nullptr;
  // multi-determinant trial (ndet > 0): occupation strings and CI
  // coefficients (owned device copies) and the workspace of its step
  int ndet = 0;
  int det_chunk = 5;
  int* occ_a = nullptr;
  int* occ_b = nullptr;
  double* ci = nullptr;
  int md_cap = 0;
  void* md_work = nullptr;
  size_t md_bytes = 0;
  // device-resident step inputs of a captured step (afq_batch): when set,
  // the auxiliary fields use the step number *step_dev and the weight update
  // the energy shift *eshift_dev instead of the call's arguments
  const int64_t* step_dev = nullptr;
  const double* eshift_dev = nullptr;
  // excitation tables of the multi-det dets relative to det 0 (owned; NULL
  // when some det is more than four orbitals away) and the method of the
  // step's multi-det overlaps and G: 0 one inverse per (walker, det), 1 ratio
  // updates relative to det 0
  afq::MdRatio* md_ratio = nullptr;
  int md_method = 0;
  // split-BF16 planes of chol_packed for the BF16x3 multi-det force bias
  // ([2][G][md_kp], built on first use, owned)
  bf16_t* md_lpk = nullptr;
  int md_kp = 0;
};

// Phases of the fused step timed by afq_system_timing_*.
enum afq_phase {
  PH_GREENS = 0, PH_ONEBODY, PH_FORCEBIAS, PH_SHIFT, PH_VHS, PH_TAYLOR, PH_OVERLAP, PH_WEIGHTS,
  PH_RNG, PH_NPHASE
};

namespace afq {

void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
// Ensure the context scratch holds at least `bytes`; returns pointer or null.
void* scratch(afq_ctx* ctx, size_t bytes);
// Device address of the comb's file-scope running-sum array on the current
// device (hipGetSymbolAddress once per device, kept for the process).
int comb_static_scratch(int device, double** out);
// Largest dynamic LDS allocation a workgroup may request on the current device.
size_t max_lds_bytes();
// compute units of the current device
int num_cus();

}  // namespace afq

#define AFQ_CHECK_HIP(expr)                                       \
  do {                                                            \
    hipError_t _e = (expr);                                       \
    if (_e != hipSuccess) return afq::hip_fail(_e, #expr);        \
  } while (0)

#define AFQ_CHECK_LAUNCH() AFQ_CHECK_HIP(hipGetLastError())

#define AFQ_REQUIRE(cond, ...)                                    \
  do {                                                            \
    if (!(cond)) return afq::fail(AFQ_EINVAL, __VA_ARGS__);       \
  } while (0)

static inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
static inline int cdiv_int(long a, long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Block reductions (blockDim.x a multiple of 64, <= 1024).  Fixed tree order:
// results are deterministic for a given launch shape.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// All threads receive the block sum.  `red` is LDS of >= 16 doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

__device__ __forceinline__ cplx block_sum_c(cplx v, double* red) {
  double re = block_sum(v.x, red);
  double im = block_sum(v.y, red + 16);
  return mkc(re, im);
}
