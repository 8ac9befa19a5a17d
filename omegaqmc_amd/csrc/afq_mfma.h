// MFMA kernel launchers (FP64 and FP32 in afq_mfma64.hip, split BF16 in afq_bf3.hip).
#pragma once
#include "afq_common.h"

namespace afq {

constexpr int FB_MAX_SPLITS = 8;

// K-splits of the force-bias contraction; each split reads one spin.
struct FbSplits {
  const cplx* gh[FB_MAX_SPLITS];
  const void* lh[FB_MAX_SPLITS];  // double (FP64 kernel) or float (FP32 kernel)
  long K[FB_MAX_SPLITS];
  int kbeg[FB_MAX_SPLITS];
  int kend[FB_MAX_SPLITS];
  int n;
};

FbSplits fb_splits(int M, int Na, int Nb, const double* Gha, const double* Ghb, const void* Lha,
                   const void* Lhb, int per_spin);

// VHS (W,M,M) complex from xs (W,G) complex and L (G,P) real; colmap (P) maps a
// packed column to its two (p,q)/(q,p) offsets (NULL: P = M*M, identity).
int vhs_mfma_f64(hipStream_t s, int W, int M, int G, const double* xs, const double* L, int P,
                 const int2* colmap, double dt, double* VHS);
// part[s] (W,G) complex partial force bias of split s.
int fb_mfma_f64(hipStream_t s, int W, int G, const FbSplits& sp, double* part);
// phi_out <- expH1 . Taylor_nmax(VHS) . expH1 . phi_in (both spins); expH1 NULL
// skips the one-body products.  phi_out may alias phi_in.
int prop_mfma_f64(hipStream_t s, int W, int M, int Na, int Nb, const double* expH1,
                  const double* VHS, const double* pa_in, const double* pb_in, double* pa_out,
                  double* pb_out, int nmax);
size_t prop_f64_smem(int M);
// the fused FP64 propagator handles M (its column block fits LDS and M <= 512)
bool prop_f64_fused_fits(int M);
// phi_out[w] = sum_{n<=nmax} (coef[w] mat)^n phi_in[w] / n! (FP64; coef (W,) and
// mat (M, M) real, phi (W, M, N) complex); phi_out may alias phi_in.
int prop_scaled_f64(hipStream_t s, int W, int M, int N, const double* mat, const double* coef,
                    const double* phi_in, double* phi_out, int nmax);
// FP32 MFMA forms: L (G,P) and Lhalf in FP32, VHS (W,M,M) complex64, the
// one-body matrix given as expH1 - I (FP32); walkers, Ghalf, xs and the
// force-bias partials stay FP64.
int vhs_mfma_f32(hipStream_t s, int W, int M, int G, const double* xs, const float* L, int P,
                 const int2* colmap, double dt, float* VHS);
int fb_mfma_f32(hipStream_t s, int W, int G, const FbSplits& sp, double* part);
int prop_mfma_f32(hipStream_t s, int W, int M, int Na, int Nb, const float* expH1_minus_I,
                  const float* VHS, const double* pa_in, const double* pb_in, double* pa_out,
                  double* pb_out, int nmax);

}  // namespace afq
