// Prepared system and the fused phaseless step
// (OmegaQMC/afqmc_gto.py:630-725 propagate_walkers; 1466-1473 block energy).
//
// Step order, identical to the reference:
//   1. Ghalf_s, ovlp_old  from the current walkers        (greens.py:104-129)
//   2. phi_s <- expH1 phi_s                                (afqmc_gto.py:678-679)
//   3. vbias, xbar (capped), xshifted, cfb, cmf            (682-706)
//   4. VHS = i sqrt(dt) sum_g xshifted_g L_g               (472-482)
//   5. phi_s <- sum_{n<=nmax} VHS^n phi_s / n!             (451-469)
//   6. phi_s <- expH1 phi_s                                (714-715)
//   7. ovlp_new                                            (greens.py:132-154)
//   8. phaseless weight update                             (527-571)
// The system owns a workspace sized for the largest walker count seen; a
// steady-state step allocates nothing.
#include "afq_gemm.h"
#include "afq_kernels.h"
#include "afq_step.h"

namespace afq {

namespace {

struct Layout {
  size_t off[16];
  size_t total;
};

enum Slot { GHA, GHB, OVLP_OLD, VBIAS, XS, CFB, CMF, XI, VHS, A1, A2, B1, B2, T0, T1, NSLOT };

Layout layout(const afq_system* s, int W) {
  const long M = s->M, Na = s->Na, Nb = s->Nb, G = s->G;
  const long N = std::max(Na, Nb);
  size_t sz[NSLOT];
  sz[GHA] = 16L * W * Na * M;
  sz[GHB] = 16L * W * Nb * M;
  sz[OVLP_OLD] = 16L * W;
  sz[VBIAS] = 16L * W * G;
  sz[XS] = 16L * W * G;
  sz[CFB] = 16L * W;
  sz[CMF] = 16L * W;
  sz[XI] = 8L * W * G;
  sz[VHS] = 16L * W * M * M;
  sz[A1] = 16L * W * M * Na;
  sz[A2] = 16L * W * M * Na;
  sz[B1] = 16L * W * M * Nb;
  sz[B2] = 16L * W * M * Nb;
  sz[T0] = 16L * W * M * N;
  sz[T1] = 16L * W * M * N;
  Layout L;
  size_t o = 0;
  for (int i = 0; i < NSLOT; ++i) {
    L.off[i] = o;
    o += align256(sz[i]);
  }
  L.total = o;
  return L;
}

int ensure_work(afq_system* s, int W) {
  if (W <= s->cap_walkers && s->work) return AFQ_OK;
  Layout L = layout(s, W);
  if (s->work) {
    AFQ_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    AFQ_CHECK_HIP(hipFree(s->work));
    s->work = nullptr;
  }
  hipError_t e = hipMalloc(&s->work, L.total);
  if (e != hipSuccess) {
    s->work = nullptr;
    s->cap_walkers = 0;
    return fail(AFQ_ENOMEM, "system workspace of %zu bytes for %d walkers", L.total, W);
  }
  s->work_bytes = L.total;
  s->cap_walkers = W;
  return AFQ_OK;
}

}  // namespace

double* work_slot(afq_system* s, int W, int slot) {
  Layout L = layout(s, s->cap_walkers);
  (void)W;
  return (double*)((char*)s->work + L.off[slot]);
}

int propagate_fp64(afq_system* sys, int W, double* phia, double* phib, double* weights,
                   double* e_hybrid, double eshift, const double* xi_in, uint64_t seed,
                   int64_t step, double* ovlp_new) {
  hipStream_t s = sys->ctx->stream;
  const int M = sys->M, Na = sys->Na, Nb = sys->Nb, G = sys->G;
  int rc = ensure_work(sys, W);
  if (rc) return rc;
  double* gha = work_slot(sys, W, GHA);
  double* ghb = work_slot(sys, W, GHB);
  double* oo = work_slot(sys, W, OVLP_OLD);
  double* vb = work_slot(sys, W, VBIAS);
  double* xs = work_slot(sys, W, XS);
  double* cfb = work_slot(sys, W, CFB);
  double* cmf = work_slot(sys, W, CMF);
  double* vhs = work_slot(sys, W, VHS);
  double* a1 = work_slot(sys, W, A1);
  double* a2 = work_slot(sys, W, A2);
  double* b1 = work_slot(sys, W, B1);
  double* b2 = work_slot(sys, W, B2);
  double* t0 = work_slot(sys, W, T0);
  double* t1 = work_slot(sys, W, T1);
  const double* xi = xi_in;
  if (!xi) {
    double* xg = work_slot(sys, W, XI);
    rc = aux_fields_launch(s, seed, step, W, G, xg);
    if (rc) return rc;
    xi = xg;
  }
  // 1
  rc = greens_launch(s, 1, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn, gha, ghb, oo);
  if (rc) return rc;
  // 2
  if ((rc = one_body_fp64(s, W, M, Na, sys->expH1, phia, a1))) return rc;
  if ((rc = one_body_fp64(s, W, M, Nb, sys->expH1, phib, b1))) return rc;
  // 3
  if ((rc = force_bias_gemm_fp64(s, W, M, Na, Nb, G, sys->rchol_a, sys->rchol_b, gha, ghb, vb)))
    return rc;
  if ((rc = shift_launch(s, W, G, vb, sys->mf_shift, xi, sys->dt, sys->fbbound, nullptr, xs, cfb,
                         cmf)))
    return rc;
  // 4
  if ((rc = build_vhs_fp64(s, W, M, G, xs, sys->chol, sys->dt, vhs))) return rc;
  // 5
  if ((rc = apply_exp_vhs_fp64(s, W, M, Na, vhs, a1, sys->nmax, a2, t0, t1))) return rc;
  if ((rc = apply_exp_vhs_fp64(s, W, M, Nb, vhs, b1, sys->nmax, b2, t0, t1))) return rc;
  // 6
  if ((rc = one_body_fp64(s, W, M, Na, sys->expH1, a2, phia))) return rc;
  if ((rc = one_body_fp64(s, W, M, Nb, sys->expH1, b2, phib))) return rc;
  // 7
  if ((rc = greens_launch(s, 0, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn, nullptr,
                          nullptr, ovlp_new)))
    return rc;
  // 8
  return weights_launch(s, W, weights, oo, ovlp_new, cfb, cmf, e_hybrid, eshift, sys->dt, weights,
                        e_hybrid);
}

int system_energy_fp64(afq_system* sys, int W, const double* phia, const double* phib,
                       const double* h1e, double enuc, double* e_tot, double* e_1b, double* e_2b) {
  hipStream_t s = sys->ctx->stream;
  const int M = sys->M, Na = sys->Na, Nb = sys->Nb, G = sys->G;
  int rc = ensure_work(sys, W);
  if (rc) return rc;
  double* gha = work_slot(sys, W, GHA);
  double* ghb = work_slot(sys, W, GHB);
  double* oo = work_slot(sys, W, OVLP_OLD);
  double* X = work_slot(sys, W, VBIAS);
  double* hT = work_slot(sys, W, XS);  // (M, Na) + (M, Nb) complex: fits in W*G complex? see below
  double* ec = work_slot(sys, W, CFB);
  double* ex = work_slot(sys, W, CMF);
  double* part = work_slot(sys, W, VHS);
  if ((long)M * (Na + Nb) > (long)W * G) return fail(AFQ_EUNSUPPORTED, "energy workspace too small");
  rc = greens_launch(s, 1, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn, gha, ghb, oo);
  if (rc) return rc;
  // hT_s = h1e @ conj(trial_s)  (M x N_s)
  double* hTa = hT;
  double* hTb = hT + 2L * M * Na;
  const int Ns[2] = {Na, Nb};
  const double* tr[2] = {sys->trial_up, sys->trial_dn};
  double* hts[2] = {hTa, hTb};
  for (int sp = 0; sp < 2; ++sp) {
    if (Ns[sp] == 0) continue;
    GemmArgs g;
    g.m = M; g.n = Ns[sp]; g.k = M;
    g.A = h1e; g.lda_m = M; g.lda_k = 1; g.a_cplx = false;
    g.B = tr[sp]; g.ldb_k = Ns[sp]; g.ldb_n = 1; g.b_cplx = true; g.conjB = true;
    g.C = (cplx*)hts[sp]; g.ldc = Ns[sp];
    if ((rc = gemm_fp64(s, g))) return rc;
  }
  if ((rc = e1b_half_launch(s, W, M, Na, Nb, hTa, hTb, gha, ghb, enuc, e_1b))) return rc;
  if ((rc = coulomb_launch(s, W, M, Na, Nb, G, sys->rchol_a, sys->rchol_b, false, gha, ghb, X, ec)))
    return rc;
  // T slab in the VHS region (W*M*M complex): gchunk with gchunk*N*W*N <= W*M*M
  const int N = std::max(Na, Nb);
  int gc = N > 0 ? std::max(1, (M * M) / (N * N)) : G;
  gc = std::min(gc, std::max(G, 1));
  if ((rc = exchange_fp64(s, W, M, Na, G, gha, sys->rchol_a, false, gc, part, ex, false))) return rc;
  if ((rc = exchange_fp64(s, W, M, Nb, G, ghb, sys->rchol_b, false, gc, part, ex, true))) return rc;
  return combine_launch(s, W, e_1b, ec, ex, e_tot, e_2b);
}

}  // namespace afq

extern "C" {

int afq_system_create(afq_ctx* ctx, afq_system** out, int M, int Na, int Nb, int G,
                      const double* chol, const double* ra, const double* rb, const double* expH1,
                      const double* mf, const double* tu, const double* td, double dt,
                      double fbbound, int nmax, int precision) {
  AFQ_REQUIRE(ctx && out, "ctx/out is NULL");
  AFQ_REQUIRE(M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M && G >= 0, "bad shape");
  AFQ_REQUIRE(Na <= 64 && Nb <= 64, "nocc > 64 is not supported");
  AFQ_REQUIRE(dt > 0 && nmax >= 0, "bad dt / nmax");
  AFQ_REQUIRE(precision == AFQ_FP64 || precision == AFQ_FP32 || precision == AFQ_BF16X3,
              "unknown precision %d", precision);
  AFQ_REQUIRE(chol && expH1 && mf && tu && td && (Na == 0 || ra) && (Nb == 0 || rb),
              "NULL system array");
  afq_system* s = new afq_system();
  s->ctx = ctx;
  s->M = M; s->Na = Na; s->Nb = Nb; s->G = G;
  s->chol = chol; s->rchol_a = ra; s->rchol_b = rb; s->expH1 = expH1; s->mf_shift = mf;
  s->trial_up = tu; s->trial_dn = td;
  s->dt = dt; s->fbbound = fbbound; s->nmax = nmax; s->precision = precision;
  int rc = afq::system_prepare_lowp(s);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return AFQ_OK;
}

int afq_system_destroy(afq_system* s) {
  if (!s) return AFQ_OK;
  (void)hipStreamSynchronize(s->ctx->stream);
  if (s->work) (void)hipFree(s->work);
  afq::system_release_lowp(s);
  delete s;
  return AFQ_OK;
}

int afq_propagate_walkers(afq_system* sys, int W, double* phia, double* phib, double* weights,
                          double* e_hybrid, double eshift, const double* xi, uint64_t seed,
                          int64_t step, double* ovlp_new) {
  AFQ_REQUIRE(sys, "system is NULL");
  AFQ_REQUIRE(W >= 0, "bad walker count");
  if (W == 0) return AFQ_OK;
  AFQ_REQUIRE(phia && phib && weights && e_hybrid && ovlp_new, "NULL walker array");
  if (sys->precision == AFQ_FP64)
    return afq::propagate_fp64(sys, W, phia, phib, weights, e_hybrid, eshift, xi, seed, step,
                               ovlp_new);
  return afq::propagate_lowp(sys, W, phia, phib, weights, e_hybrid, eshift, xi, seed, step,
                             ovlp_new);
}

int afq_system_local_energy(afq_system* sys, int W, const double* phia, const double* phib,
                            const double* h1e, double enuc, double* e_tot, double* e_1b,
                            double* e_2b) {
  AFQ_REQUIRE(sys, "system is NULL");
  AFQ_REQUIRE(W >= 0, "bad walker count");
  AFQ_REQUIRE(e_tot && e_1b, "NULL output");
  if (W == 0) return AFQ_OK;
  if (sys->precision == AFQ_FP64)
    return afq::system_energy_fp64(sys, W, phia, phib, h1e, enuc, e_tot, e_1b, e_2b);
  return afq::system_energy_lowp(sys, W, phia, phib, h1e, enuc, e_tot, e_1b, e_2b);
}

}  // extern "C"
