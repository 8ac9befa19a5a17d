// FP32-MFMA phaseless step: the force bias, the VHS build and the fused
// one-body / Taylor / one-body propagator run on v_mfma_f32_16x16x4f32 with
// FP32 operands (L, Lhalf, expH1 - I) and FP32 accumulation; the Green's
// function, determinants, the force-bias cap and shifts, the weights, QR and
// the walkers themselves stay FP64.  Same step order as the FP64 path
// (OmegaQMC/afqmc_gto.py:630-725); the block energy is the FP64 one.
#include "afq_gemm.h"
#include "afq_kernels.h"
#include "afq_mfma.h"
#include "afq_step.h"

namespace afq {

void phase_begin(afq_system* s, int ph);
void phase_end(afq_system* s, int ph);

namespace {

__global__ void k_to_f32(long n, const double* __restrict__ src, int diag_n,
                         float* __restrict__ dst) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    double v = src[i];
    // diag_n > 0: src is diag_n x diag_n, and the identity is taken off first
    if (diag_n > 0 && i / diag_n == i % diag_n) v -= 1.0;
    dst[i] = (float)v;
  }
}

int to_f32(hipStream_t st, const double* src, long n, int diag_n, float** out) {
  *out = nullptr;
  if (n <= 0) return AFQ_OK;
  if (hipMalloc(out, sizeof(float) * n) != hipSuccess) {
    *out = nullptr;
    return fail(AFQ_ENOMEM, "FP32 operand of %ld values", n);
  }
  hipLaunchKernelGGL(k_to_f32, dim3(2048), dim3(256), 0, st, n, src, diag_n, *out);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace

struct F32System {
  float* L = nullptr;    // packed upper triangle (G, P) or full (G, M*M)
  int P = 0;
  float* lha = nullptr;  // (G, Na*M)
  float* lhb = nullptr;  // (G, Nb*M)
  float* D = nullptr;    // expH1 - I (M, M)
};

int f32_prepare(afq_system* s) {
  F32System* f = new F32System();
  s->chol_lp = f;
  hipStream_t st = s->ctx->stream;
  const int M = s->M, G = s->G;
  int rc = AFQ_OK;
  f->P = M * M;  // no vectors: the VHS GEMM (K = 0) writes the full zero matrix
  if (G > 0) {
    if (s->chol_packed) {
      f->P = s->packed_P;
      rc = to_f32(st, s->chol_packed, (long)G * f->P, 0, &f->L);
    } else {
      f->P = M * M;
      rc = to_f32(st, s->chol, (long)G * f->P, 0, &f->L);
    }
    if (!rc && s->Na > 0) rc = to_f32(st, s->rchol_a, (long)G * s->Na * M, 0, &f->lha);
    if (!rc && s->Nb > 0) rc = to_f32(st, s->rchol_b, (long)G * s->Nb * M, 0, &f->lhb);
  }
  if (!rc) rc = to_f32(st, s->expH1, (long)M * M, M, &f->D);
  if (!rc) AFQ_CHECK_HIP(hipStreamSynchronize(st));
  return rc;
}

void f32_release(afq_system* s) {
  F32System* f = (F32System*)s->chol_lp;
  if (!f) return;
  for (float* p : {f->L, f->lha, f->lhb, f->D})
    if (p) (void)hipFree(p);
  delete f;
  s->chol_lp = nullptr;
}

int propagate_fp32(afq_system* sys, int W, double* phia, double* phib, double* weights,
                   double* e_hybrid, double eshift, const double* xi_in, uint64_t seed,
                   int64_t step, double* ovlp_new) {
  const F32System& f = *(const F32System*)sys->chol_lp;
  hipStream_t s = sys->ctx->stream;
  const int M = sys->M, Na = sys->Na, Nb = sys->Nb, G = sys->G;
  StepSlots w;
  int rc = step_slots(sys, W, &w);
  if (rc) return rc;
  const double* xi = xi_in;
  if (!xi) {
    phase_begin(sys, PH_RNG);
    rc = aux_fields_launch(s, seed, step, W, G, w.xi, sys->step_dev);
    phase_end(sys, PH_RNG);
    if (rc) return rc;
    xi = w.xi;
  }
  phase_begin(sys, PH_GREENS);
  rc = greens_launch(s, 1 | sys->eye_mask, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn,
                     w.gha, w.ghb, w.oo, nullptr, w.inv);
  phase_end(sys, PH_GREENS);
  if (rc) return rc;
  phase_begin(sys, PH_FORCEBIAS);
  FbSplits sp = fb_splits(M, Na, Nb, w.gha, w.ghb, f.lha, f.lhb, std::max(1, sys->fb_nsplit / 2));
  if ((rc = fb_mfma_f32(s, W, G, sp, w.vb))) return rc;
  phase_end(sys, PH_FORCEBIAS);
  phase_begin(sys, PH_SHIFT);
  if ((rc = shift_launch(s, W, G, w.vb, sys->mf_shift, xi, sys->dt, sys->fbbound, nullptr, w.xs,
                         w.cfb, w.cmf, sp.n, nullptr, nullptr, 0, sys->chol_sign)))
    return rc;
  phase_end(sys, PH_SHIFT);
  phase_begin(sys, PH_VHS);
  rc = vhs_mfma_f32(s, W, M, G, w.xs, f.L, f.P, sys->chol_packed ? sys->colmap : nullptr, sys->dt,
                    (float*)w.vhs);
  phase_end(sys, PH_VHS);
  if (rc) return rc;
  phase_begin(sys, PH_TAYLOR);
  rc = prop_mfma_f32(s, W, M, Na, Nb, f.D, (const float*)w.vhs, phia, phib, phia, phib,
                     sys->nmax);
  phase_end(sys, PH_TAYLOR);
  if (rc) return rc;
  phase_begin(sys, PH_OVERLAP);
  if ((rc = greens_launch(s, sys->eye_mask, W, M, Na, Nb, phia, phib, sys->trial_up,
                          sys->trial_dn, nullptr, nullptr, ovlp_new)))
    return rc;
  phase_end(sys, PH_OVERLAP);
  phase_begin(sys, PH_WEIGHTS);
  rc = weights_launch(s, W, weights, w.oo, ovlp_new, w.cfb, w.cmf, e_hybrid, eshift, sys->dt,
                      weights, e_hybrid, sys->eshift_dev);
  phase_end(sys, PH_WEIGHTS);
  return rc;
}

}  // namespace afq
