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
#include <vector>

#include "afq_gemm.h"
#include "afq_kernels.h"
#include "afq_mfma.h"
#include "afq_multidet.h"
#include "afq_step.h"

struct Timing {
  bool on = false;
  unsigned mask = ~0u;  // phases timed while on
  std::vector<hipEvent_t> ev[PH_NPHASE];  // pairs (start, stop)
  int used[PH_NPHASE] = {};
  double total_ms[PH_NPHASE] = {};
  int64_t count[PH_NPHASE] = {};
};

namespace afq {

void phase_begin(afq_system* s, int ph) {
  Timing* t = s->timing;
  if (!t || !t->on || !((t->mask >> ph) & 1u)) return;
  if ((int)t->ev[ph].size() < 2 * (t->used[ph] + 1)) {
    for (int k = 0; k < 2; ++k) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      t->ev[ph].push_back(e);
    }
  }
  (void)hipEventRecord(t->ev[ph][2 * t->used[ph]], s->ctx->stream);
}

void phase_end(afq_system* s, int ph) {
  Timing* t = s->timing;
  if (!t || !t->on || !((t->mask >> ph) & 1u) || (int)t->ev[ph].size() < 2 * (t->used[ph] + 1))
    return;
  (void)hipEventRecord(t->ev[ph][2 * t->used[ph] + 1], s->ctx->stream);
  t->used[ph] += 1;
}

static int timing_collect(afq_system* s) {
  Timing* t = s->timing;
  if (!t) return AFQ_OK;
  AFQ_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
  for (int ph = 0; ph < PH_NPHASE; ++ph) {
    for (int i = 0; i < t->used[ph]; ++i) {
      float ms = 0.f;
      AFQ_CHECK_HIP(hipEventElapsedTime(&ms, t->ev[ph][2 * i], t->ev[ph][2 * i + 1]));
      t->total_ms[ph] += ms;
      t->count[ph] += 1;
    }
    t->used[ph] = 0;
  }
  return AFQ_OK;
}

namespace {

__global__ void k_sym_check(long n, int M, const double* __restrict__ L, int* __restrict__ bad) {
  const long MM = (long)M * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long g = i / MM, r = i % MM;
    const int p = (int)(r / M), q = (int)(r % M);
    if (q > p && L[g * MM + (long)p * M + q] != L[g * MM + (long)q * M + p]) atomicOr(bad, 1);
  }
}

__global__ void k_pack_upper(long n, int P, long MM, const double* __restrict__ L,
                             const int2* __restrict__ colmap, double* __restrict__ Lp) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long g = i / P;
    const int c = (int)(i % P);
    Lp[i] = L[g * MM + colmap[c].x];
  }
}

struct Layout {
  size_t off[16];
  size_t total;
};

enum Slot { GHA, GHB, OVLP_OLD, VBIAS, XS, CFB, CMF, XI, VHS, A1, A2, B1, B2, T0, T1, INV, NSLOT };

Layout layout(const afq_system* s, int W) {
  const long M = s->M, Na = s->Na, Nb = s->Nb, G = s->G;
  const long N = std::max(Na, Nb);
  size_t sz[NSLOT];
  sz[GHA] = 16L * W * Na * M;
  sz[GHB] = 16L * W * Nb * M;
  sz[OVLP_OLD] = 16L * W;
  sz[VBIAS] = 16L * W * G * std::max(1, s->fb_nsplit);
  sz[XS] = 16L * std::max((long)W * G, M * (Na + Nb));  // also h1e Psi* of the energy
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
  sz[INV] = greens_inv_bytes(W, (int)Na, (int)Nb);
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

// FP64 operands of the step: the upper triangle of every L_g (when every L_g
// is exactly symmetric, so that VHS_qp = VHS_pq bit for bit) and the number
// of K splits of the force bias.
int prepare_fp64(afq_system* s) {
  hipStream_t st = s->ctx->stream;
  const int M = s->M, G = s->G;
  s->fb_nsplit = 4;
  // identity trials make S = phi^T Psi* a transposed gather
  const int Ns[2] = {s->Na, s->Nb};
  const double* tr[2] = {s->trial_up, s->trial_dn};
  s->eye_mask = 0;
  for (int sp = 0; sp < 2; ++sp) {
    const size_t n = 2 * (size_t)M * Ns[sp];
    if (n == 0 || !tr[sp]) continue;  // no single-det trial (multi-det system)
    double* h = new double[n];
    hipError_t e = hipMemcpyAsync(h, tr[sp], n * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      delete[] h;
      return hip_fail(e, "trial copy");
    }
    bool eye = true;
    for (int p = 0; p < M && eye; ++p)
      for (int i = 0; i < Ns[sp]; ++i) {
        const double* z = h + 2 * ((size_t)p * Ns[sp] + i);
        if (z[0] != (p == i ? 1.0 : 0.0) || z[1] != 0.0) { eye = false; break; }
      }
    delete[] h;
    if (eye) s->eye_mask |= (sp == 0 ? 2 : 4);
  }
  if (G == 0) return AFQ_OK;
  int* bad = nullptr;
  AFQ_CHECK_HIP(hipMalloc(&bad, sizeof(int)));
  AFQ_CHECK_HIP(hipMemsetAsync(bad, 0, sizeof(int), st));
  const long n = (long)G * M * M;
  hipLaunchKernelGGL(k_sym_check, dim3(4096), dim3(256), 0, st, n, M, s->chol, bad);
  int hbad = 1;
  AFQ_CHECK_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
  AFQ_CHECK_HIP(hipStreamSynchronize(st));
  AFQ_CHECK_HIP(hipFree(bad));
  if (hbad) return AFQ_OK;  // not exactly symmetric: full M*M columns
  const int P = M * (M + 1) / 2;
  int2* hmap = new int2[P];
  int c = 0;
  for (int p = 0; p < M; ++p)
    for (int q = p; q < M; ++q) hmap[c++] = make_int2(p * M + q, q == p ? -1 : q * M + p);
  int2* dmap = nullptr;
  double* Lp = nullptr;
  hipError_t e1 = hipMalloc(&dmap, sizeof(int2) * P);
  hipError_t e2 = hipMalloc(&Lp, sizeof(double) * (size_t)G * P);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    delete[] hmap;
    if (dmap) (void)hipFree(dmap);
    if (Lp) (void)hipFree(Lp);
    return fail(AFQ_ENOMEM, "packed Cholesky operand (%zu bytes)", sizeof(double) * (size_t)G * P);
  }
  AFQ_CHECK_HIP(hipMemcpyAsync(dmap, hmap, sizeof(int2) * P, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_pack_upper, dim3(8192), dim3(256), 0, st, (long)G * P, P, (long)M * M,
                     s->chol, dmap, Lp);
  AFQ_CHECK_LAUNCH();
  AFQ_CHECK_HIP(hipStreamSynchronize(st));
  delete[] hmap;
  s->colmap = dmap;
  s->chol_packed = Lp;
  s->packed_P = P;
  return AFQ_OK;
}

double* work_slot(afq_system* s, int W, int slot) {
  Layout L = layout(s, s->cap_walkers);
  (void)W;
  return (double*)((char*)s->work + L.off[slot]);
}

int step_slots(afq_system* sys, int W, StepSlots* o) {
  int rc = ensure_work(sys, W);
  if (rc) return rc;
  o->gha = work_slot(sys, W, GHA);
  o->ghb = work_slot(sys, W, GHB);
  o->oo = work_slot(sys, W, OVLP_OLD);
  o->vb = work_slot(sys, W, VBIAS);
  o->xs = work_slot(sys, W, XS);
  o->cfb = work_slot(sys, W, CFB);
  o->cmf = work_slot(sys, W, CMF);
  o->xi = work_slot(sys, W, XI);
  o->vhs = work_slot(sys, W, VHS);
  o->inv = work_slot(sys, W, INV);
  return AFQ_OK;
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
    phase_begin(sys, PH_RNG);
    rc = aux_fields_launch(s, seed, step, W, G, xg, sys->step_dev);
    phase_end(sys, PH_RNG);
    if (rc) return rc;
    xi = xg;
  }
  // 1
  phase_begin(sys, PH_GREENS);
  rc = greens_launch(s, 1 | sys->eye_mask, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn,
                     gha, ghb, oo, nullptr, work_slot(sys, W, INV));
  phase_end(sys, PH_GREENS);
  if (rc) return rc;
  // 3 (the first one-body half step (2) is fused into the propagator below:
  //    it does not feed the force bias, which uses the pre-step Ghalf)
  phase_begin(sys, PH_FORCEBIAS);
  FbSplits sp = fb_splits(M, Na, Nb, gha, ghb, sys->rchol_a, sys->rchol_b,
                          std::max(1, sys->fb_nsplit / 2));
  if ((rc = fb_mfma_f64(s, W, G, sp, vb))) return rc;
  phase_end(sys, PH_FORCEBIAS);
  phase_begin(sys, PH_SHIFT);
  if ((rc = shift_launch(s, W, G, vb, sys->mf_shift, xi, sys->dt, sys->fbbound, nullptr, xs, cfb,
                         cmf, sp.n, nullptr, nullptr, 0, sys->chol_sign)))
    return rc;
  phase_end(sys, PH_SHIFT);
  // 4
  phase_begin(sys, PH_VHS);
  if (sys->chol_packed)
    rc = vhs_mfma_f64(s, W, M, G, xs, sys->chol_packed, sys->packed_P, sys->colmap, sys->dt, vhs);
  else
    rc = vhs_mfma_f64(s, W, M, G, xs, sys->chol, M * M, nullptr, sys->dt, vhs);
  phase_end(sys, PH_VHS);
  if (rc) return rc;
  // 2, 5, 6: expH1 . exp(VHS) . expH1, in place
  phase_begin(sys, PH_TAYLOR);
  if (prop_f64_fused_fits(M))
    rc = prop_mfma_f64(s, W, M, Na, Nb, sys->expH1, vhs, phia, phib, phia, phib, sys->nmax);
  else
    rc = prop_generic_f64(s, W, M, Na, Nb, sys->expH1, vhs, phia, phib, a1, a2, b1, b2, t0, t1,
                          sys->nmax);
  phase_end(sys, PH_TAYLOR);
  if (rc) return rc;
  // 7
  phase_begin(sys, PH_OVERLAP);
  if ((rc = greens_launch(s, sys->eye_mask, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn, nullptr,
                          nullptr, ovlp_new)))
    return rc;
  phase_end(sys, PH_OVERLAP);
  // 8
  phase_begin(sys, PH_WEIGHTS);
  rc = weights_launch(s, W, weights, oo, ovlp_new, cfb, cmf, e_hybrid, eshift, sys->dt, weights,
                      e_hybrid, sys->eshift_dev);
  phase_end(sys, PH_WEIGHTS);
  return rc;
}

int energy_prologue(afq_system* sys, int W, const double* phia, const double* phib,
                    const double* h1e, double enuc, const GhalfBf16* gbf, double* e_1b,
                    EnergySlots* es) {
  hipStream_t s = sys->ctx->stream;
  const int M = sys->M, Na = sys->Na, Nb = sys->Nb, G = sys->G;
  int rc = ensure_work(sys, W);
  if (rc) return rc;
  es->gha = work_slot(sys, W, GHA);
  es->ghb = work_slot(sys, W, GHB);
  double* oo = work_slot(sys, W, OVLP_OLD);
  es->X = work_slot(sys, W, VBIAS);
  double* hT = work_slot(sys, W, XS);  // (M, Na) + (M, Nb) complex
  es->ec = work_slot(sys, W, CFB);
  es->ex = work_slot(sys, W, CMF);
  es->part = work_slot(sys, W, VHS);
  es->part_bytes = (size_t)W * M * M * 16;
  rc = greens_launch(s, 1 | sys->eye_mask, W, M, Na, Nb, phia, phib, sys->trial_up, sys->trial_dn,
                     es->gha, es->ghb, oo, gbf, work_slot(sys, W, INV));
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
  return e1b_half_launch(s, W, M, Na, Nb, hTa, hTb, es->gha, es->ghb, enuc, e_1b);
}

int system_energy_fp64(afq_system* sys, int W, const double* phia, const double* phib,
                       const double* h1e, double enuc, double* e_tot, double* e_1b, double* e_2b) {
  hipStream_t s = sys->ctx->stream;
  const int M = sys->M, Na = sys->Na, Nb = sys->Nb, G = sys->G;
  EnergySlots es;
  int rc = energy_prologue(sys, W, phia, phib, h1e, enuc, nullptr, e_1b, &es);
  if (rc) return rc;
  // es.X: the force-bias partial slot, max(1, fb_nsplit) slices of (W,G)
  if ((rc = coulomb_launch(s, W, M, Na, Nb, G, sys->rchol_a, sys->rchol_b, false, es.gha, es.ghb,
                           es.X, std::max(1, sys->fb_nsplit), es.ec, sys->chol_sign)))
    return rc;
  // T slab in the VHS region (W*M*M complex): gchunk with gchunk*N*W*N <= W*M*M
  const int N = std::max(Na, Nb);
  int gc = N > 0 ? std::max(1, (M * M) / (N * N)) : G;
  gc = std::min(gc, std::max(G, 1));
  if ((rc = exchange_fp64(s, W, M, Na, G, es.gha, sys->rchol_a, false, gc, es.part, es.ex, false,
                          sys->chol_sign)))
    return rc;
  if ((rc = exchange_fp64(s, W, M, Nb, G, es.ghb, sys->rchol_b, false, gc, es.part, es.ex, true,
                          sys->chol_sign)))
    return rc;
  return combine_launch(s, W, e_1b, es.ec, es.ex, e_tot, e_2b);
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
  AFQ_REQUIRE(chol && expH1 && mf && (Na == 0 || (tu && ra)) && (Nb == 0 || (td && rb)),
              "NULL system array");
  afq_system* s = new afq_system();
  s->ctx = ctx;
  s->M = M; s->Na = Na; s->Nb = Nb; s->G = G;
  s->chol = chol; s->rchol_a = ra; s->rchol_b = rb; s->expH1 = expH1; s->mf_shift = mf;
  s->trial_up = tu; s->trial_dn = td;
  s->dt = dt; s->fbbound = fbbound; s->nmax = nmax; s->precision = precision;
  int rc = afq::prepare_fp64(s);
  if (!rc) rc = afq::system_prepare_lowp(s);
  if (rc) {
    afq_system_destroy(s);
    return rc;
  }
  *out = s;
  return AFQ_OK;
}

int afq_system_set_chol_sign(afq_system* s, const double* sign) {
  AFQ_REQUIRE(s, "system is NULL");
  if (!sign || s->G == 0) {
    s->chol_sign = nullptr;
    return AFQ_OK;
  }
  std::vector<double> h(s->G);
  AFQ_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
  AFQ_CHECK_HIP(hipMemcpy(h.data(), sign, sizeof(double) * s->G, hipMemcpyDeviceToHost));
  bool neg = false;
  for (double v : h) {
    AFQ_REQUIRE(v == 1.0 || v == -1.0, "chol_sign entries must be +1 or -1 (got %g)", v);
    neg = neg || v < 0.0;
  }
  // FP64 and FP32 single-determinant steps take signed vectors (the shift
  // kernel folds the sign into the VHS coefficients; the energy is the FP64
  // estimator with the signs); a BF16x3 system would need its vectors
  // symmetric, which the electron gas's antisymmetric (sign -1) ones are not
  if (neg && (s->precision == AFQ_BF16X3 || s->ndet > 0))
    return afq::fail(AFQ_EUNSUPPORTED,
                     "Cholesky vectors with sign -1 need the FP64 or FP32 single-determinant path");
  s->chol_sign = neg ? sign : nullptr;
  return AFQ_OK;
}

int afq_system_set_force_bias_splits(afq_system* s, int n) {
  AFQ_REQUIRE(s, "system is NULL");
  AFQ_REQUIRE(n == 2 || n == 4 || (n == 8 && s->precision != AFQ_BF16X3),
              "force-bias splits %d (2, 4 or 8; BF16X3: 2 or 4)", n);
  if (n == s->fb_nsplit) return AFQ_OK;
  // the FP64 workspace holds one partial force bias per split: re-sized on
  // the next step
  if (s->work) {
    AFQ_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    AFQ_CHECK_HIP(hipFree(s->work));
    s->work = nullptr;
    s->cap_walkers = 0;
  }
  s->fb_nsplit = n;
  return AFQ_OK;
}

int afq_system_timing_enable(afq_system* s, int on) {
  AFQ_REQUIRE(s, "system is NULL");
  if (!s->timing) s->timing = new Timing();
  if (!on) {
    int rc = afq::timing_collect(s);
    if (rc) return rc;
  }
  s->timing->on = (on != 0);
  s->timing->mask = on < 0 ? (unsigned)(-(long)on) : ~0u;
  return AFQ_OK;
}

int afq_system_timing_read(afq_system* s, double* total_ms, int64_t* launches, int nphase,
                           int reset) {
  AFQ_REQUIRE(s && total_ms && launches, "NULL argument");
  for (int i = 0; i < nphase; ++i) { total_ms[i] = 0.0; launches[i] = 0; }
  if (!s->timing) return AFQ_OK;
  int rc = afq::timing_collect(s);
  if (rc) return rc;
  for (int i = 0; i < nphase && i < PH_NPHASE; ++i) {
    total_ms[i] = s->timing->total_ms[i];
    launches[i] = s->timing->count[i];
    if (reset) { s->timing->total_ms[i] = 0.0; s->timing->count[i] = 0; }
  }
  return AFQ_OK;
}

int afq_system_destroy(afq_system* s) {
  if (!s) return AFQ_OK;
  (void)hipStreamSynchronize(s->ctx->stream);
  if (s->timing) {
    for (int ph = 0; ph < PH_NPHASE; ++ph)
      for (hipEvent_t e : s->timing->ev[ph]) (void)hipEventDestroy(e);
    delete s->timing;
  }
  if (s->work) (void)hipFree(s->work);
  if (s->md_work) (void)hipFree(s->md_work);
  if (s->occ_a) (void)hipFree(s->occ_a);
  if (s->occ_b) (void)hipFree(s->occ_b);
  if (s->ci) (void)hipFree(s->ci);
  if (s->chol_packed) (void)hipFree(s->chol_packed);
  if (s->colmap) (void)hipFree(s->colmap);
  if (s->md_lpk) (void)hipFree(s->md_lpk);
  afq::md_ratio_release(s->md_ratio);
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
  AFQ_REQUIRE((sys->Na == 0 || phia) && (sys->Nb == 0 || phib) && weights && e_hybrid && ovlp_new,
              "NULL walker array");
  if (sys->ndet > 0)
    return afq::propagate_multidet(sys, W, phia, phib, weights, e_hybrid, eshift, xi, seed, step,
                                   ovlp_new);
  if (sys->precision == AFQ_FP64)
    return afq::propagate_fp64(sys, W, phia, phib, weights, e_hybrid, eshift, xi, seed, step,
                               ovlp_new);
  if (sys->precision == AFQ_FP32)
    return afq::propagate_fp32(sys, W, phia, phib, weights, e_hybrid, eshift, xi, seed, step,
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
  if (sys->ndet > 0)
    return afq::system_energy_multidet(sys, W, phia, phib, h1e, enuc, e_tot, e_1b, e_2b);
  if (sys->precision == AFQ_FP64 || sys->precision == AFQ_FP32)
    return afq::system_energy_fp64(sys, W, phia, phib, h1e, enuc, e_tot, e_1b, e_2b);
  return afq::system_energy_lowp(sys, W, phia, phib, h1e, enuc, e_tot, e_1b, e_2b);
}

}  // extern "C"
