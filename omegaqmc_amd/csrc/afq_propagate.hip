// Propagation pieces of the phaseless step (OmegaQMC/afqmc_gto.py:451-725):
// force bias + shifted fields, VHS build, Taylor exp(VHS) phi, one-body half
// step and the phaseless weight update, each as a reference-shaped ABI
// entry point.  The FP64 contractions run on the generic tiled GEMM; the
// fused, precision-selectable step lives in afq_step.hip.
#include "afq_gemm.h"
#include "afq_kernels.h"
#include "afq_mfma.h"

namespace afq {

namespace {

constexpr int NT = 256;

// afqmc_gto.py:687-706 per walker (one workgroup): cap, shift, cfb, cmf.
// vbias comes as nsplit FP64 partials (vbias) or FP32 partials (vpart), summed
// in FP64 in split order.  Optional outputs: xbar, xshifted (FP64) and the
// split-bf16 planes of xshifted, Xp[2][2W][Gp] with rows (w, re/im), K-tiled:
// element (r, g) at ((r / 16) (Gp / 32) + g / 32) 512 + (r % 16) 32 + g % 32, each
// plane 2W rounded up to 16 rows.
// sign (optional, (G,)): the electron-gas form (afqmc_pw_heg.py:895-922), HS
// phase i for sign +1 and 1 for sign -1: xbar = -sqrt(dt) (phase vbias - mf)
// and VHS = sqrt(dt) sum_g xs_g phase_g L_g, written here as the coefficient
// xs_g phase_g / i of the i sqrt(dt) sum_g (.) L_g the VHS kernels form (the
// xs output then holds that coefficient; cmf and cfb use xs itself).
__global__ __launch_bounds__(NT) void k_shift(int G, int nsplit, long sstride,
                                              const cplx* __restrict__ vbias,  // nsplit x (W,G)
                                              const float2* __restrict__ vpart,
                                              const cplx* __restrict__ mf,            // (G,)
                                              const double* __restrict__ xi,          // (W,G)
                                              double sqdt, double fb, cplx* __restrict__ xbar_out,
                                              cplx* __restrict__ xs_out, cplx* __restrict__ cfb,
                                              cplx* __restrict__ cmf, bf16_t* __restrict__ Xp,
                                              int Gp, long Xplane,
                                              const double* __restrict__ sign) {
  __shared__ double red[32];
  const long w = blockIdx.x;
  cplx s_mf = mkc(0, 0), s_xixb = mkc(0, 0), s_xb2 = mkc(0, 0);
  for (int g = threadIdx.x; g < G; g += NT) {
    cplx v;
    if (vpart) {
      v = mkc(0.0, 0.0);
      for (int sp = 0; sp < nsplit; ++sp) {
        const float2 pv = vpart[sp * sstride + w * G + g];
        v = cadd(v, mkc((double)pv.x, (double)pv.y));
      }
    } else {
      v = vbias[w * G + g];
      for (int sp = 1; sp < nsplit; ++sp) v = cadd(v, vbias[sp * sstride + w * G + g]);
    }
    const cplx m = mf[g];
    const bool btype = sign && sign[g] < 0.0;
    cplx xb = btype ? mkc(-sqdt * (v.x - m.x), -sqdt * (v.y - m.y))
                    : mkc(sqdt * (v.y + m.x), -sqdt * (v.x - m.y));
    const double a = hypot(xb.x, xb.y);
    if (a > fb) xb = mkc((xb.x * fb) / a, (xb.y * fb) / a);
    const double x = xi[w * G + g];
    const cplx xs = mkc(x - xb.x, -xb.y);
    const cplx xv = btype ? mkc(xs.y, -xs.x) : xs;  // coefficient of i sqrt(dt) L_g
    if (xbar_out) xbar_out[w * G + g] = xb;
    if (xs_out) xs_out[w * G + g] = xv;
    if (Xp) {
      bf16_t h, l;
      const long r = 2 * w;  // even: rows r, r + 1 share a 16-row block
      const long o = ((r >> 4) * (Gp >> 5) + (g >> 5)) * 512 + (r & 15) * 32 + (g & 31);
      split2(xv.x, h, l);
      Xp[o] = h;
      Xp[Xplane + o] = l;
      split2(xv.y, h, l);
      Xp[o + 32] = h;
      Xp[Xplane + o + 32] = l;
    }
    s_mf = cadd(s_mf, cmul(xs, m));
    s_xixb = cadd(s_xixb, cscale(xb, x));
    s_xb2 = cadd(s_xb2, cmul(xb, xb));
  }
  s_mf = block_sum_c(s_mf, red);
  s_xixb = block_sum_c(s_xixb, red);
  s_xb2 = block_sum_c(s_xb2, red);
  if (threadIdx.x == 0) {
    cmf[w] = cscale(s_mf, -sqdt);
    cfb[w] = csub(s_xixb, cscale(s_xb2, 0.5));
  }
}

// afqmc_gto.py:527-571, one thread per walker.
__global__ void k_weights(int W, const double* __restrict__ w_in, const cplx* __restrict__ o_old,
                          const cplx* __restrict__ o_new, const cplx* __restrict__ cfb,
                          const cplx* __restrict__ cmf, const cplx* __restrict__ eh_old,
                          double eshift_arg, const double* __restrict__ eshift_dev, double dt,
                          double* __restrict__ w_out,
                          cplx* __restrict__ eh_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  const double eshift = eshift_dev ? *eshift_dev : eshift_arg;
  const cplx r = cdiv(o_new[i], o_old[i]);
  const cplx lr = mkc(log(hypot(r.x, r.y)), atan2(r.y, r.x));
  const cplx f = cfb[i], m = cmf[i];
  cplx e = mkc(-((lr.x + f.x) + m.x) / dt, -((lr.y + f.y) + m.y) / dt);
  const double eb = sqrt(2.0 / dt);
  e.x = jmin(jmax(e.x, eshift - eb), eshift + eb);
  const cplx eo = eh_old[i];
  const cplx avg = mkc(0.5 * (e.x + eo.x), 0.5 * (e.y + eo.y));
  const double zr = -dt * (avg.x - eshift), zi = -dt * avg.y;
  const double ez = exp(zr);
  const double magn = hypot(ez * cos(zi), ez * sin(zi));
  const double dth = -dt * e.y - f.y;
  const double cf = jmax(0.0, cos(dth));
  w_out[i] = w_in[i] * magn * cf;
  eh_out[i] = e;
}

__global__ void k_copy(long n, const double* __restrict__ a, double* __restrict__ b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_zero(long n, unsigned* __restrict__ p) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

}  // namespace

int copy_async(hipStream_t s, const double* src, double* dst, long n) {
  if (n <= 0 || src == dst) return AFQ_OK;
  int blocks = (int)std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, n, src, dst);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

// zero fill by a kernel rather than hipMemsetAsync: these run inside steps
// that a caller may capture into a hipGraph
int zero_async(hipStream_t s, void* p, size_t bytes) {
  const long n = (long)(bytes / 4);
  if (n <= 0) return AFQ_OK;
  int blocks = (int)std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, s, n, (unsigned*)p);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

// vbias^T (W,G) = sum_s Ghalf_s (W, N_s M) . rchol_s (G, N_s M)^T  (real rchol)
int force_bias_gemm_fp64(hipStream_t s, int W, int M, int Na, int Nb, int G, const double* ra,
                         const double* rb, const double* Gha, const double* Ghb, double* vbiasT) {
  const int Ns[2] = {Na, Nb};
  const double* R[2] = {ra, rb};
  const double* Gh[2] = {Gha, Ghb};
  bool first = true;
  for (int sp = 0; sp < 2; ++sp) {
    if (Ns[sp] == 0) continue;
    const long K = (long)Ns[sp] * M;
    GemmArgs g;
    g.m = W; g.n = G; g.k = (int)K; g.batch = 1;
    g.A = Gh[sp]; g.lda_m = K; g.lda_k = 1; g.a_cplx = true;
    g.B = R[sp]; g.ldb_k = 1; g.ldb_n = K; g.b_cplx = false;
    g.C = (cplx*)vbiasT; g.ldc = G;
    g.beta = first ? 0.0 : 1.0;
    int rc = gemm_fp64(s, g);
    if (rc) return rc;
    first = false;
  }
  if (first) return zero_async(s, vbiasT, sizeof(cplx) * (size_t)W * G);  // no electrons
  return AFQ_OK;
}

int shift_launch(hipStream_t s, int W, int G, const double* vbiasT, const double* mf,
                 const double* xi, double dt, double fb, double* xbar, double* xs, double* cfb,
                 double* cmf, int nsplit, const float* vpart, bf16_t* Xp, int Gp,
                 const double* sign) {
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_shift, dim3(W), dim3(NT), 0, s, G, nsplit, (long)W * G,
                     (const cplx*)vbiasT, (const float2*)vpart, (const cplx*)mf, xi, sqrt(dt), fb,
                     (cplx*)xbar, (cplx*)xs, (cplx*)cfb, (cplx*)cmf, Xp, Gp, (2L * W + 15) / 16 * 16 * Gp,
                     sign);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int build_vhs_fp64(hipStream_t s, int W, int M, int G, const double* xs, const double* chol,
                   double dt, double* VHS) {
  GemmArgs g;
  g.m = W; g.n = M * M; g.k = G; g.batch = 1;
  g.A = xs; g.lda_m = G; g.lda_k = 1; g.a_cplx = true;
  g.B = chol; g.ldb_k = (long)M * M; g.ldb_n = 1; g.b_cplx = false;
  g.C = (cplx*)VHS; g.ldc = (long)M * M;
  g.alpha = mkc(0.0, sqrt(dt));
  return gemm_fp64(s, g);
}

// phi_out = sum_{n<=nmax} VHS^n phi / n!  (tmp0, tmp1: W*M*N complex each)
int apply_exp_vhs_fp64(hipStream_t s, int W, int M, int N, const double* VHS, const double* phi,
                       int nmax, double* phi_out, double* tmp0, double* tmp1) {
  if (N == 0 || W == 0) return AFQ_OK;
  const long n = (long)W * M * N * 2;
  int rc = copy_async(s, phi, tmp0, n);
  if (rc) return rc;
  rc = copy_async(s, phi, phi_out, n);
  if (rc) return rc;
  double* tin = tmp0;
  double* tout = tmp1;
  for (int k = 1; k <= nmax; ++k) {
    GemmArgs g;
    g.m = M; g.n = N; g.k = M; g.batch = W;
    g.A = VHS; g.lda_m = M; g.lda_k = 1; g.sA = (long)M * M; g.a_cplx = true;
    g.B = tin; g.ldb_k = N; g.ldb_n = 1; g.sB = (long)M * N; g.b_cplx = true;
    g.C = (cplx*)tout; g.ldc = N; g.sC = (long)M * N;
    g.R = (cplx*)phi_out; g.ldr = N; g.sR = (long)M * N;
    g.alpha = mkc(1.0 / k, 0.0);
    rc = gemm_fp64(s, g);
    if (rc) return rc;
    double* t = tin; tin = tout; tout = t;
  }
  return AFQ_OK;
}

// expH1 . exp(VHS) . expH1 on both spins through the batched GEMMs, for
// orbital counts whose column block does not fit the fused kernel's LDS.
// phi_s is overwritten; a_s, b_s (W M N_s complex) and t0, t1 (W M max N_s
// complex) are scratch.
int prop_generic_f64(hipStream_t s, int W, int M, int Na, int Nb, const double* expH1,
                     const double* VHS, double* pa, double* pb, double* a1, double* a2,
                     double* b1, double* b2, double* t0, double* t1, int nmax) {
  for (int sp = 0; sp < 2; ++sp) {
    const int N = sp ? Nb : Na;
    double* phi = sp ? pb : pa;
    double* x = sp ? b1 : a1;
    double* y = sp ? b2 : a2;
    int rc = one_body_fp64(s, W, M, N, expH1, phi, x);
    if (!rc) rc = apply_exp_vhs_fp64(s, W, M, N, VHS, x, nmax, y, t0, t1);
    if (!rc) rc = one_body_fp64(s, W, M, N, expH1, y, phi);
    if (rc) return rc;
  }
  return AFQ_OK;
}

int one_body_fp64(hipStream_t s, int W, int M, int N, const double* expH1, const double* phi,
                  double* out) {
  if (N == 0 || W == 0) return AFQ_OK;
  GemmArgs g;
  g.m = M; g.n = N; g.k = M; g.batch = W;
  g.A = expH1; g.lda_m = M; g.lda_k = 1; g.sA = 0; g.a_cplx = false;
  g.B = phi; g.ldb_k = N; g.ldb_n = 1; g.sB = (long)M * N; g.b_cplx = true;
  g.C = (cplx*)out; g.ldc = N; g.sC = (long)M * N;
  return gemm_fp64(s, g);
}

int weights_launch(hipStream_t s, int W, const double* w, const double* oo, const double* on,
                   const double* cfb, const double* cmf, const double* eh, double eshift, double dt,
                   double* w_out, double* eh_out, const double* eshift_dev) {
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL(k_weights, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, w, (const cplx*)oo,
                     (const cplx*)on, (const cplx*)cfb, (const cplx*)cmf, (const cplx*)eh, eshift,
                     eshift_dev, dt, w_out, (cplx*)eh_out);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace afq

extern "C" {

int afq_build_vhs(afq_ctx* ctx, int W, int M, int G, const double* xs, const double* chol,
                  double dt, double* VHS) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && G >= 0, "bad shape");
  return afq::vhs_mfma_f64(ctx->stream, W, M, G, xs, chol, M * M, nullptr, dt, VHS);
}

int afq_apply_exp_vhs(afq_ctx* ctx, int W, int M, int N, const double* VHS, const double* phi,
                      int nmax, double* phi_out) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && N >= 0 && nmax >= 0, "bad shape");
  if (afq::prop_f64_fused_fits(M))
    return afq::prop_mfma_f64(ctx->stream, W, M, N, 0, nullptr, VHS, phi, nullptr, phi_out,
                              nullptr, nmax);
  size_t one = align256(sizeof(double) * 2 * (size_t)W * M * N);
  double* tmp = (double*)afq::scratch(ctx, 2 * one);
  if (!tmp && one) return AFQ_ENOMEM;
  return afq::apply_exp_vhs_fp64(ctx->stream, W, M, N, VHS, phi, nmax, phi_out, tmp,
                                 (double*)((char*)tmp + one));
}

int afq_apply_one_body(afq_ctx* ctx, int W, int M, int N, const double* expH1, const double* phi,
                       double* out) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && N >= 0, "bad shape");
  AFQ_REQUIRE(out != phi || N == 0 || W == 0, "afq_apply_one_body: out must not alias phi");
  return afq::one_body_fp64(ctx->stream, W, M, N, expH1, phi, out);
}

int afq_force_bias(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, const double* ra,
                   const double* rb, const double* Gha, const double* Ghb, const double* mf,
                   const double* xi, double dt, double fb, double* xbar, double* xs, double* cfb,
                   double* cmf) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && G >= 0, "bad shape");
  afq::FbSplits sp = afq::fb_splits(M, Na, Nb, Gha, Ghb, ra, rb, 2);
  const size_t nsp = std::max(1, sp.n);
  double* vb = (double*)afq::scratch(ctx, sizeof(double) * 2 * (size_t)W * G * nsp);
  if (!vb && W * (size_t)G) return AFQ_ENOMEM;
  int rc;
  if (sp.n == 0) {
    if ((rc = afq::zero_async(ctx->stream, vb, sizeof(double) * 2 * (size_t)W * G))) return rc;
  } else if ((rc = afq::fb_mfma_f64(ctx->stream, W, G, sp, vb))) {
    return rc;
  }
  return afq::shift_launch(ctx->stream, W, G, vb, mf, xi, dt, fb, xbar, xs, cfb, cmf,
                           (int)nsp);
}

int afq_update_weights_phaseless(afq_ctx* ctx, int W, const double* w, const double* oo,
                                 const double* on, const double* cfb, const double* cmf,
                                 const double* eh, double eshift, double dt, double* w_out,
                                 double* eh_out) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && dt > 0, "bad arguments");
  return afq::weights_launch(ctx->stream, W, w, oo, on, cfb, cmf, eh, eshift, dt, w_out, eh_out);
}

}  // extern "C"
