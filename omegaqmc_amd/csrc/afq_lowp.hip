// Reduced-precision phaseless step: the force bias, the VHS build and the
// propagator run in split-BF16 MFMA with FP32 accumulation (afq_bf3.hip);
// the Green's function, determinants, the force-bias cap and shifts, the
// weights and the walkers themselves stay FP64.  Same step order as the FP64
// path (OmegaQMC/afqmc_gto.py:630-725).
#include "afq_bf3.h"
#include "afq_kernels.h"
#include "afq_step.h"

namespace afq {

void phase_begin(afq_system* s, int ph);
void phase_end(afq_system* s, int ph);

namespace {

struct LowpWork {
  int cap = 0;
  void* buf = nullptr;
  size_t bytes = 0;
  // offsets
  size_t gh_a, gh_b, oo, part, X, cfb, cmf, xi, vhs, t0, inv;
};

constexpr int FB_SPLITS_MAX = 4;

size_t lay(const Bf3System& b, int W, LowpWork* L) {
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t r = o;
    o += align256(n);
    return r;
  };
  L->gh_a = take((size_t)2 * 2 * W * b.Ka * sizeof(bf16_t));
  L->gh_b = take((size_t)2 * 2 * W * b.Kb * sizeof(bf16_t));
  L->oo = take(16L * W);
  L->part = take((size_t)FB_SPLITS_MAX * W * b.G * sizeof(float2));
  L->X = take(bf3_x_bytes(b, W));
  L->cfb = take(16L * W);
  L->cmf = take(16L * W);
  L->xi = take(8L * W * b.G);
  L->vhs = take(bf3_vhs_bytes(b, W));
  L->t0 = take(bf3_prop_scratch_bytes(b, W));
  L->inv = take(greens_inv_bytes(W, b.Na, b.Nb));
  return o;
}

int ensure(afq_system* s, int W, LowpWork** out) {
  Bf3System* b = (Bf3System*)s->chol_lp;
  LowpWork* L = (LowpWork*)s->rchol_lp;
  if (!L) {
    L = new LowpWork();
    s->rchol_lp = L;
  }
  if (L->buf && W <= L->cap) {
    *out = L;
    return AFQ_OK;
  }
  if (L->buf) {
    AFQ_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    AFQ_CHECK_HIP(hipFree(L->buf));
    L->buf = nullptr;
  }
  const size_t n = lay(*b, W, L);
  if (hipMalloc(&L->buf, n) != hipSuccess) {
    L->buf = nullptr;
    L->cap = 0;
    return fail(AFQ_ENOMEM, "reduced-precision workspace of %zu bytes", n);
  }
  // zero once: the padding of the bf16 planes must read as zeros
  AFQ_CHECK_HIP(hipMemsetAsync(L->buf, 0, n, s->ctx->stream));
  L->bytes = n;
  L->cap = W;
  *out = L;
  return AFQ_OK;
}

}  // namespace

// The split-bf16 xshifted planes, VHS and propagator scratch of the
// reduced-precision workspace (used by the multi-determinant step, whose force
// bias is FP64).
int lowp_step_buffers(afq_system* sys, int W, bf16_t** X, bf16_t** vhs, float** t0) {
  LowpWork* L = nullptr;
  int rc = ensure(sys, W, &L);
  if (rc) return rc;
  *X = (bf16_t*)((char*)L->buf + L->X);
  *vhs = (bf16_t*)((char*)L->buf + L->vhs);
  *t0 = (float*)((char*)L->buf + L->t0);
  return AFQ_OK;
}

int system_prepare_lowp(afq_system* sys) {
  if (sys->precision == AFQ_FP64) return AFQ_OK;
  if (sys->precision == AFQ_FP32) return f32_prepare(sys);
  if (sys->precision != AFQ_BF16X3)
    return fail(AFQ_EUNSUPPORTED, "precision %d is not built", sys->precision);
  // the split-BF16 VHS is built on the upper-triangle blocks and mirrored:
  // every L_g must be exactly symmetric (prepare_fp64 packs it only then)
  if (sys->G > 0 && !sys->chol_packed)
    return fail(AFQ_EUNSUPPORTED,
                "bf16x3 needs exactly symmetric Cholesky vectors (FP64 and FP32 take any)");
  Bf3System* b = new Bf3System();
  sys->chol_lp = b;
  return bf3_prepare(sys, b);
}

// the workspace is re-laid on the next step (its propagator scratch depends
// on the column blocking)
int lowp_invalidate(afq_system* sys) {
  LowpWork* L = (LowpWork*)sys->rchol_lp;
  if (L && L->buf) {
    AFQ_CHECK_HIP(hipStreamSynchronize(sys->ctx->stream));
    AFQ_CHECK_HIP(hipFree(L->buf));
    L->buf = nullptr;
    L->cap = 0;
  }
  return AFQ_OK;
}

void system_release_lowp(afq_system* sys) {
  if (sys->precision == AFQ_FP32) {
    f32_release(sys);
    return;
  }
  if (sys->chol_lp) {
    bf3_release((Bf3System*)sys->chol_lp);
    delete (Bf3System*)sys->chol_lp;
  }
  if (sys->rchol_lp) {
    LowpWork* L = (LowpWork*)sys->rchol_lp;
    if (L->buf) (void)hipFree(L->buf);
    delete L;
  }
  sys->chol_lp = sys->rchol_lp = nullptr;
}

int propagate_lowp(afq_system* sys, int W, double* phia, double* phib, double* weights,
                   double* e_hybrid, double eshift, const double* xi_in, uint64_t seed,
                   int64_t step, double* ovlp_new) {
  hipStream_t s = sys->ctx->stream;
  const Bf3System& b = *(const Bf3System*)sys->chol_lp;
  const int M = sys->M, Na = sys->Na, Nb = sys->Nb, G = sys->G;
  LowpWork* L = nullptr;
  int rc = ensure(sys, W, &L);
  if (rc) return rc;
  char* base = (char*)L->buf;
  bf16_t* gha = (bf16_t*)(base + L->gh_a);
  bf16_t* ghb = (bf16_t*)(base + L->gh_b);
  double* oo = (double*)(base + L->oo);
  float* part = (float*)(base + L->part);
  bf16_t* X = (bf16_t*)(base + L->X);
  double* cfb = (double*)(base + L->cfb);
  double* cmf = (double*)(base + L->cmf);
  bf16_t* vhs = (bf16_t*)(base + L->vhs);
  const double* xi = xi_in;
  if (!xi) {
    double* xg = (double*)(base + L->xi);
    phase_begin(sys, PH_RNG);
    rc = aux_fields_launch(s, seed, step, W, G, xg, sys->step_dev);
    phase_end(sys, PH_RNG);
    if (rc) return rc;
    xi = xg;
  }
  // 1: Ghalf (split bf16 planes only) and ovlp_old, FP64 LU
  GhalfBf16 gbf{Na ? gha : nullptr, Nb ? ghb : nullptr, b.Ka, b.Kb, b.ldq};
  phase_begin(sys, PH_GREENS);
  rc = greens_launch(s, 1 | sys->eye_mask, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn,
                     nullptr, nullptr, oo, &gbf, (double*)(base + L->inv));
  phase_end(sys, PH_GREENS);
  if (rc) return rc;
  // 3: force bias (split-K, FP32 partials), shift in FP64
  int nsplit = 0;
  phase_begin(sys, PH_FORCEBIAS);
  rc = bf3_force_bias(s, b, W, gha, ghb, part, &nsplit, sys->fb_nsplit / 2);
  phase_end(sys, PH_FORCEBIAS);
  if (rc) return rc;
  if (nsplit > FB_SPLITS_MAX) return fail(AFQ_EUNSUPPORTED, "force-bias splits %d", nsplit);
  if (nsplit == 0) {
    if ((rc = zero_async(s, part, sizeof(float2) * (size_t)W * G))) return rc;
    nsplit = 1;
  }
  phase_begin(sys, PH_SHIFT);
  rc = shift_launch(s, W, G, nullptr, sys->mf_shift, xi, sys->dt, sys->fbbound, nullptr, nullptr,
                    cfb, cmf, nsplit, part, X, b.Gp, sys->chol_sign);
  phase_end(sys, PH_SHIFT);
  if (rc) return rc;
  // 4
  phase_begin(sys, PH_VHS);
  rc = bf3_vhs(s, b, W, X, sys->dt, vhs);
  phase_end(sys, PH_VHS);
  if (rc) return rc;
  // 2, 5, 6
  phase_begin(sys, PH_TAYLOR);
  rc = bf3_propagate(s, b, W, vhs, (float*)(base + L->t0), phia, phib, sys->nmax);
  phase_end(sys, PH_TAYLOR);
  if (rc) return rc;
  // 7, 8 in FP64
  phase_begin(sys, PH_OVERLAP);
  rc = greens_launch(s, sys->eye_mask, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn,
                     nullptr, nullptr, ovlp_new);
  phase_end(sys, PH_OVERLAP);
  if (rc) return rc;
  phase_begin(sys, PH_WEIGHTS);
  rc = weights_launch(s, W, weights, oo, ovlp_new, cfb, cmf, e_hybrid, eshift, sys->dt, weights,
                      e_hybrid, sys->eshift_dev);
  phase_end(sys, PH_WEIGHTS);
  return rc;
}

// Block-end energy (afqmc_gto.py:1466-1473): Ghalf, overlap and E_1b in FP64;
// the Coulomb X_g (force-bias GEMM) and the exchange T_g products in split
// BF16 with FP32 accumulation, their squares and all sums over g in FP64.
int system_energy_lowp(afq_system* sys, int W, const double* phia, const double* phib,
                       const double* h1e, double enuc, double* e_tot, double* e_1b, double* e_2b) {
  const Bf3System& b = *(const Bf3System*)sys->chol_lp;
  if (!bf3_exchange_supported(b) || sys->chol_sign)
    return system_energy_fp64(sys, W, phia, phib, h1e, enuc, e_tot, e_1b, e_2b);
  hipStream_t s = sys->ctx->stream;
  LowpWork* L = nullptr;
  int rc = ensure(sys, W, &L);
  if (rc) return rc;
  char* base = (char*)L->buf;
  bf16_t* gha = (bf16_t*)(base + L->gh_a);
  bf16_t* ghb = (bf16_t*)(base + L->gh_b);
  float* part = (float*)(base + L->part);
  GhalfBf16 gbf{sys->Na ? gha : nullptr, sys->Nb ? ghb : nullptr, b.Ka, b.Kb, b.ldq};
  EnergySlots es;
  if ((rc = energy_prologue(sys, W, phia, phib, h1e, enuc, &gbf, e_1b, &es))) return rc;
  if (bf3_exchange_part_bytes(b, W) > es.part_bytes)
    return fail(AFQ_EUNSUPPORTED, "exchange partials exceed the workspace");
  int nsplit = 0;
  if ((rc = bf3_force_bias(s, b, W, gha, ghb, part, &nsplit, sys->fb_nsplit / 2))) return rc;
  if (nsplit > FB_SPLITS_MAX) return fail(AFQ_EUNSUPPORTED, "force-bias splits %d", nsplit);
  if ((rc = bf3_coulomb(s, W, sys->G, nsplit, part, es.ec))) return rc;
  if ((rc = bf3_exchange(s, b, W, gha, ghb, es.part, es.ex))) return rc;
  return combine_launch(s, W, e_1b, es.ec, es.ex, e_tot, e_2b);
}

}  // namespace afq

extern "C" int afq_system_set_propagator_block(afq_system* sys, int cols) {
  AFQ_REQUIRE(sys, "system is NULL");
  AFQ_REQUIRE(cols == 16 || cols == 32 || cols == 48, "propagator block %d (16, 32 or 48)", cols);
  if (sys->precision != AFQ_BF16X3 || !sys->chol_lp) return AFQ_OK;  // fixed 16-column blocks
  afq::Bf3System* b = (afq::Bf3System*)sys->chol_lp;
  if (b->pcols == cols) return AFQ_OK;
  int rc = afq::lowp_invalidate(sys);
  if (rc) return rc;
  b->pcols = cols;
  return AFQ_OK;
}
