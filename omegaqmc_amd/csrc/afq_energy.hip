// Mixed-estimator local energy (OmegaQMC/observables/energy.py:32-252).
//
//   E_1b   = sum_s Tr(h G_s) + E_nuc          (or sum_pi (h Psi*)_pi Ghalf_ip)
//   X_s[g] = sum_iq Lhalf_s[g,i,q] Ghalf_s[i,q]
//   E_coul = 1/2 sum_g (X_a[g] + X_b[g])^2    (complex square)
//   T_s[g] = Lhalf_s[g] Ghalf_s^T  (N x N)
//   E_exch = 1/2 sum_s sum_g sum_ij T_ij T_ji  (= Tr(T T), not Tr(T T^T))
//
// FP64 exchange: with real Lhalf and N <= 40 the fused MFMA kernel of
// afq_exch.hip (T in registers and LDS only); otherwise g is processed in
// slabs, each slab's T block for all walkers one GEMM (Lhalf slab (gs N x M)
// times Ghalf^T (M x W N)) into scratch, contracted per walker by
// k_exch_contract in a fixed order.
#include "afq_gemm.h"
#include "afq_kernels.h"
#include "afq_mfma.h"

namespace afq {

namespace {

constexpr int NT = 256;

// e[w] (+)= 1/2 sum_{g<gs} sign_g sum_ij T[(g,i)][(w,j)] T[(g,j)][(w,i)],
// T row-major (gs*N) x (W*N); sign (optional) the slab's ERI signs
// (afqmc_pw_heg.py:414-457).  One workgroup per walker.
__global__ __launch_bounds__(NT) void k_exch_contract(int W, int N, int gs,
                                                      const cplx* __restrict__ T,
                                                      cplx* __restrict__ e, int accumulate,
                                                      const double* __restrict__ sign) {
  __shared__ double red[32];
  const long w = blockIdx.x;
  const long ld = (long)W * N;
  cplx acc = mkc(0, 0);
  const int per = N * N;
  for (long e2 = threadIdx.x; e2 < (long)gs * per; e2 += NT) {
    int g = (int)(e2 / per), r = (int)(e2 % per);
    int i = r / N, j = r % N;
    cplx a = T[((long)g * N + i) * ld + w * N + j];
    cplx b = T[((long)g * N + j) * ld + w * N + i];
    if (sign && sign[g] < 0.0) a = mkc(-a.x, -a.y);
    cfma(acc, a, b);
  }
  acc = block_sum_c(acc, red);
  if (threadIdx.x == 0) {
    cplx v = cscale(acc, 0.5);
    e[w] = accumulate ? cadd(e[w], v) : v;
  }
}

// e_coul[w] = 1/2 sum_g sign_g X[w][g]^2, X (W,G) complex, sign optional.
__global__ __launch_bounds__(NT) void k_coulomb(int G, const cplx* __restrict__ X,
                                                cplx* __restrict__ ec,
                                                const double* __restrict__ sign) {
  __shared__ double red[32];
  const long w = blockIdx.x;
  cplx acc = mkc(0, 0);
  for (int g = threadIdx.x; g < G; g += NT) {
    cplx x = X[w * G + g];
    if (sign && sign[g] < 0.0) cfms(acc, x, x);
    else cfma(acc, x, x);
  }
  acc = block_sum_c(acc, red);
  if (threadIdx.x == 0) ec[w] = cscale(acc, 0.5);
}

// e1b[w] = sum_{p,q} h[p][q] (Ga[w][q][p] + Gb[w][q][p]) + enuc
template <bool HC>
__global__ __launch_bounds__(NT) void k_e1b_fullg(int M, const double* __restrict__ h,
                                                  const cplx* __restrict__ Ga,
                                                  const cplx* __restrict__ Gb, double enuc,
                                                  cplx* __restrict__ e1b) {
  __shared__ double red[32];
  const long w = blockIdx.x;
  const long MM = (long)M * M;
  cplx sa = mkc(0, 0), sb = mkc(0, 0);
  for (long e2 = threadIdx.x; e2 < MM; e2 += NT) {
    int p = (int)(e2 / M), q = (int)(e2 % M);
    cplx hv = HC ? ((const cplx*)h)[e2] : mkc(h[e2], 0.0);
    cfma(sa, hv, Ga[w * MM + (long)q * M + p]);
    cfma(sb, hv, Gb[w * MM + (long)q * M + p]);
  }
  sa = block_sum_c(sa, red);
  sb = block_sum_c(sb, red);
  if (threadIdx.x == 0) e1b[w] = mkc(sa.x + sb.x + enuc, sa.y + sb.y);
}

// e1b[w] = sum_{p,i} hTa[p][i] Gha[w][i][p] + sum hTb Ghb + enuc
__global__ __launch_bounds__(NT) void k_e1b_half(int M, int Na, int Nb,
                                                 const cplx* __restrict__ hTa,
                                                 const cplx* __restrict__ hTb,
                                                 const cplx* __restrict__ Gha,
                                                 const cplx* __restrict__ Ghb, double enuc,
                                                 cplx* __restrict__ e1b) {
  __shared__ double red[32];
  const long w = blockIdx.x;
  cplx sa = mkc(0, 0), sb = mkc(0, 0);
  for (long e2 = threadIdx.x; e2 < (long)Na * M; e2 += NT) {
    int i = (int)(e2 / M), p = (int)(e2 % M);
    cfma(sa, hTa[(long)p * Na + i], Gha[w * Na * M + e2]);
  }
  for (long e2 = threadIdx.x; e2 < (long)Nb * M; e2 += NT) {
    int i = (int)(e2 / M), p = (int)(e2 % M);
    cfma(sb, hTb[(long)p * Nb + i], Ghb[w * Nb * M + e2]);
  }
  sa = block_sum_c(sa, red);
  sb = block_sum_c(sb, red);
  if (threadIdx.x == 0) e1b[w] = mkc(sa.x + sb.x + enuc, sa.y + sb.y);
}

// energy.py:363-393, one det at a time: w_I = conj(c_I) o_Ia o_Ib;
// acc1 (+)= w_I e1b_I, acc2 (+)= w_I (e_coul_I - e_exch_I)  (I == 0 writes)
__global__ void k_md_stream_acc(int W, int I, const cplx* __restrict__ ci,
                                const cplx* __restrict__ oa_all, const cplx* __restrict__ ob_all,
                                const cplx* __restrict__ e1, const cplx* __restrict__ ec,
                                const cplx* __restrict__ ex, cplx* __restrict__ acc1,
                                cplx* __restrict__ acc2) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const cplx wi = cmul(cmul(cconj(ci[I]), oa_all[(long)I * W + w]), ob_all[(long)I * W + w]);
  const cplx a = cmul(wi, e1[w]), b = cmul(wi, csub(ec[w], ex[w]));
  acc1[w] = I ? cadd(acc1[w], a) : a;
  acc2[w] = I ? cadd(acc2[w], b) : b;
}

// overlap = sum_I w_I; e_1b = acc1 / overlap + enuc; e_2b = acc2 / overlap
__global__ void k_md_stream_final(int W, int nd, const cplx* __restrict__ ci,
                                  const cplx* __restrict__ oa_all,
                                  const cplx* __restrict__ ob_all, double enuc,
                                  cplx* __restrict__ e_tot, cplx* __restrict__ e1b,
                                  cplx* __restrict__ e2b) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  cplx ov = mkc(0.0, 0.0);
  for (int I = 0; I < nd; ++I)
    ov = cadd(ov, cmul(cmul(cconj(ci[I]), oa_all[(long)I * W + w]), ob_all[(long)I * W + w]));
  const cplx a = cadd(cdiv(e1b[w], ov), mkc(enuc, 0.0));
  const cplx b = cdiv(e2b[w], ov);
  e1b[w] = a;
  e2b[w] = b;
  e_tot[w] = cadd(a, b);
}

__global__ void k_combine(int W, const cplx* __restrict__ e1b, const cplx* __restrict__ ec,
                          const cplx* __restrict__ ex, cplx* __restrict__ etot,
                          cplx* __restrict__ e2b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  cplx d = csub(ec[i], ex[i]);
  if (e2b) e2b[i] = d;
  etot[i] = cadd(e1b[i], d);
}

}  // namespace

// Exchange energy of one spin, FP64.  partial: scratch of >= gchunk*N*W*N complex.
int exchange_fp64(hipStream_t s, int W, int M, int N, int G, const double* Ghalf,
                  const double* rchol, bool rchol_cplx, int gchunk, double* partial,
                  double* e_exch, bool accumulate, const double* sign) {
  if (W <= 0) return AFQ_OK;
  bool acc = accumulate;
  if (N == 0) {
    if (!acc) AFQ_CHECK_HIP(hipMemsetAsync(e_exch, 0, sizeof(cplx) * W, s));
    return AFQ_OK;
  }
  if (!rchol_cplx && exchange_f64_fused_fits(N))
    return exchange_f64_fused(s, W, M, N, G, Ghalf, rchol, partial,
                              16UL * (size_t)gchunk * N * W * N, e_exch, accumulate, sign);
  for (int g0 = 0; g0 < G; g0 += gchunk) {
    int gs = std::min(gchunk, G - g0);
    GemmArgs ga;
    ga.m = gs * N; ga.n = W * N; ga.k = M; ga.batch = 1;
    ga.A = rchol_cplx ? (const void*)((const cplx*)rchol + (long)g0 * N * M)
                      : (const void*)(rchol + (long)g0 * N * M);
    ga.lda_m = M; ga.lda_k = 1; ga.a_cplx = rchol_cplx;
    ga.B = Ghalf; ga.ldb_k = 1; ga.ldb_n = M; ga.b_cplx = true;
    ga.C = (cplx*)partial; ga.ldc = (long)W * N;
    int rc = gemm_fp64(s, ga);
    if (rc) return rc;
    hipLaunchKernelGGL(k_exch_contract, dim3(W), dim3(NT), 0, s, W, N, gs, (const cplx*)partial,
                       (cplx*)e_exch, acc ? 1 : 0, sign ? sign + g0 : nullptr);
    AFQ_CHECK_LAUNCH();
    acc = true;
  }
  if (G == 0 && !accumulate) AFQ_CHECK_HIP(hipMemsetAsync(e_exch, 0, sizeof(cplx) * W, s));
  return AFQ_OK;
}

// x[0] += x[1] + ... + x[ns - 1], slices of n doubles, in slice order
__global__ void k_sum_slices(long n, int ns, double* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double v = x[i];
    for (int k = 1; k < ns; ++k) v += x[k * n + i];
    x[i] = v;
  }
}

// X (W,G) = sum_s Ghalf_s . rchol_s^T (rchol real or complex); X holds nx
// slices of (W,G) complex.  Real vectors: the force-bias contraction on FP64
// MFMA (k_fb, one split per spin into slices 0, 1, summed into slice 0);
// complex vectors (or fewer slices than spins): the generic FP64 GEMM.
static int coulomb_x(hipStream_t s, int W, int M, int Na, int Nb, int G, const double* ra,
                     const double* rb, bool rc_cplx, const double* Gha, const double* Ghb,
                     double* X, int nx) {
  if (G == 0) return AFQ_OK;  // no vectors: X is empty
  if (!rc_cplx) {
    const FbSplits sp = fb_splits(M, Na, Nb, Gha, Ghb, ra, rb, 1);
    if (sp.n == 0) return zero_async(s, X, sizeof(cplx) * (size_t)W * G);
    if (sp.n <= nx) {
      int rc = fb_mfma_f64(s, W, G, sp, X);
      if (rc || sp.n == 1) return rc;
      const long n = 2L * W * G;
      hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)std::min<long>(cdiv_int(n, 256), 8192)),
                         dim3(256), 0, s, n, sp.n, X);
      AFQ_CHECK_LAUNCH();
      return AFQ_OK;
    }
  }
  const int Ns[2] = {Na, Nb};
  const double* R[2] = {ra, rb};
  const double* Gh[2] = {Gha, Ghb};
  bool first = true;
  for (int sp = 0; sp < 2; ++sp) {
    if (Ns[sp] == 0) continue;
    const long K = (long)Ns[sp] * M;
    GemmArgs g;
    g.m = W; g.n = G; g.k = (int)K;
    g.A = Gh[sp]; g.lda_m = K; g.lda_k = 1; g.a_cplx = true;
    g.B = R[sp]; g.ldb_k = 1; g.ldb_n = K; g.b_cplx = rc_cplx;
    g.C = (cplx*)X; g.ldc = G;
    g.beta = first ? 0.0 : 1.0;
    int rc = gemm_fp64(s, g);
    if (rc) return rc;
    first = false;
  }
  if (first) AFQ_CHECK_HIP(hipMemsetAsync(X, 0, sizeof(cplx) * (size_t)W * G, s));
  return AFQ_OK;
}

// Scratch layout of the two-body energy: X (W,G) | 3 slots of W complex |
// T slab (gchunk*N x W*N complex, <= 256 MiB).
struct E2bScratch {
  double* X = nullptr;
  double* slot[3] = {nullptr, nullptr, nullptr};
  double* part = nullptr;
  int gchunk = 1;
};

static int e2b_scratch(afq_ctx* ctx, int W, int Na, int Nb, int G, int gchunk_req, E2bScratch* out) {
  const int N = std::max(Na, Nb);
  long per_g = (long)N * W * N * 16;
  int gc = per_g > 0 ? (int)std::max<long>(1, (256L << 20) / per_g) : std::max(G, 1);
  gc = std::min(gc, std::max(G, 1));
  if (gchunk_req > 0) gc = std::min(gc, gchunk_req);
  size_t sx = align256(2 * 16L * W * G), sw = align256(16L * W);  // X: two (W,G) slices
  size_t need = sx + 3 * sw + align256((long)gc * per_g);
  char* sc = (char*)scratch(ctx, need);
  if (!sc) return AFQ_ENOMEM;
  out->X = (double*)sc;
  for (int i = 0; i < 3; ++i) out->slot[i] = (double*)(sc + sx + i * sw);
  out->part = (double*)(sc + sx + 3 * sw);
  out->gchunk = gc;
  return AFQ_OK;
}

// sign (optional, (G,)): ERI signs of the vectors, Coulomb and exchange
// weighted by them (afqmc_pw_heg.py:414-457)
static int two_body(hipStream_t s, const E2bScratch& sc, int W, int M, int Na, int Nb, int G,
                    const double* Gha, const double* Ghb, const double* ra, const double* rb,
                    bool rc_cplx, double* e_coul, double* e_exch, const double* sign = nullptr) {
  int rc = coulomb_x(s, W, M, Na, Nb, G, ra, rb, rc_cplx, Gha, Ghb, sc.X, 2);
  if (rc) return rc;
  hipLaunchKernelGGL(k_coulomb, dim3(W), dim3(NT), 0, s, G, (const cplx*)sc.X, (cplx*)e_coul,
                     sign);
  AFQ_CHECK_LAUNCH();
  rc = exchange_fp64(s, W, M, Na, G, Gha, ra, rc_cplx, sc.gchunk, sc.part, e_exch, false, sign);
  if (rc) return rc;
  return exchange_fp64(s, W, M, Nb, G, Ghb, rb, rc_cplx, sc.gchunk, sc.part, e_exch, true, sign);
}

int coulomb_launch(hipStream_t s, int W, int M, int Na, int Nb, int G, const double* ra,
                   const double* rb, bool rc_cplx, const double* Gha, const double* Ghb,
                   double* X, int nx, double* e_coul, const double* sign) {
  int rc = coulomb_x(s, W, M, Na, Nb, G, ra, rb, rc_cplx, Gha, Ghb, X, nx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_coulomb, dim3(W), dim3(NT), 0, s, G, (const cplx*)X, (cplx*)e_coul, sign);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int e1b_half_launch(hipStream_t s, int W, int M, int Na, int Nb, const double* hTa,
                    const double* hTb, const double* Gha, const double* Ghb, double enuc,
                    double* e1b) {
  hipLaunchKernelGGL(k_e1b_half, dim3(W), dim3(NT), 0, s, M, Na, Nb, (const cplx*)hTa,
                     (const cplx*)hTb, (const cplx*)Gha, (const cplx*)Ghb, enuc, (cplx*)e1b);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int e1b_fullg_launch(hipStream_t s, int W, int M, const double* h1e, const double* Ga,
                     const double* Gb, double enuc, double* e1b) {
  if (W <= 0) return AFQ_OK;
  hipLaunchKernelGGL((k_e1b_fullg<false>), dim3(W), dim3(NT), 0, s, M, h1e, (const cplx*)Ga,
                     (const cplx*)Gb, enuc, (cplx*)e1b);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int combine_launch(hipStream_t s, int W, const double* e1b, const double* ec, const double* ex,
                   double* etot, double* e2b) {
  hipLaunchKernelGGL(k_combine, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, (const cplx*)e1b,
                     (const cplx*)ec, (const cplx*)ex, (cplx*)etot, (cplx*)e2b);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace afq

extern "C" {

int afq_local_energy_2body(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, const double* Gha,
                           const double* Ghb, const double* ra, const double* rb, int dtype,
                           double* e_coul, double* e_exch) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && G >= 0, "bad shape");
  AFQ_REQUIRE(dtype == AFQ_REAL || dtype == AFQ_COMPLEX, "bad dtype %d", dtype);
  if (W == 0) return AFQ_OK;
  afq::E2bScratch sc;
  int rc = afq::e2b_scratch(ctx, W, Na, Nb, G, 0, &sc);
  if (rc) return rc;
  return afq::two_body(ctx->stream, sc, W, M, Na, Nb, G, Gha, Ghb, ra, rb, dtype == AFQ_COMPLEX,
                       e_coul, e_exch);
}

int afq_local_energy_2body_pw(afq_ctx* ctx, int W, int M, int Na, int Nb, int G,
                              const double* Gha, const double* Ghb, const double* ra,
                              const double* rb, int dtype, const double* chol_sign,
                              double* e_coul, double* e_exch) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && G >= 0, "bad shape");
  AFQ_REQUIRE(dtype == AFQ_REAL || dtype == AFQ_COMPLEX, "bad dtype %d", dtype);
  AFQ_REQUIRE(chol_sign || G == 0, "chol_sign is NULL");
  if (W == 0) return AFQ_OK;
  afq::E2bScratch sc;
  int rc = afq::e2b_scratch(ctx, W, Na, Nb, G, 0, &sc);
  if (rc) return rc;
  return afq::two_body(ctx->stream, sc, W, M, Na, Nb, G, Gha, Ghb, ra, rb, dtype == AFQ_COMPLEX,
                       e_coul, e_exch, chol_sign);
}

static int local_energy_impl(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, const double* h1e,
                             const double* Ga, const double* Gb, const double* Gha,
                             const double* Ghb, const double* ra, const double* rb, int dtype,
                             double enuc, const double* sign, double* e_tot, double* e_1b,
                             double* e_2b);

int afq_local_energy_pw(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, const double* h1e,
                        const double* Ga, const double* Gb, const double* Gha, const double* Ghb,
                        const double* ra, const double* rb, int dtype, double enuc,
                        const double* chol_sign, double* e_tot, double* e_1b, double* e_2b) {
  AFQ_REQUIRE(chol_sign || G == 0, "chol_sign is NULL");
  return local_energy_impl(ctx, W, M, Na, Nb, G, h1e, Ga, Gb, Gha, Ghb, ra, rb, dtype, enuc,
                           chol_sign, e_tot, e_1b, e_2b);
}

int afq_local_energy(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, const double* h1e,
                     const double* Ga, const double* Gb, const double* Gha, const double* Ghb,
                     const double* ra, const double* rb, int dtype, double enuc, double* e_tot,
                     double* e_1b, double* e_2b) {
  return local_energy_impl(ctx, W, M, Na, Nb, G, h1e, Ga, Gb, Gha, Ghb, ra, rb, dtype, enuc,
                           nullptr, e_tot, e_1b, e_2b);
}

static int local_energy_impl(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, const double* h1e,
                             const double* Ga, const double* Gb, const double* Gha,
                             const double* Ghb, const double* ra, const double* rb, int dtype,
                             double enuc, const double* sign, double* e_tot, double* e_1b,
                             double* e_2b) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && G >= 0, "bad shape");
  AFQ_REQUIRE(dtype == AFQ_REAL || dtype == AFQ_COMPLEX, "bad dtype %d", dtype);
  AFQ_REQUIRE(e_1b != nullptr, "e_1b is required");
  if (W == 0) return AFQ_OK;
  hipStream_t s = ctx->stream;
  if (dtype == AFQ_COMPLEX)
    hipLaunchKernelGGL((afq::k_e1b_fullg<true>), dim3(W), dim3(afq::NT), 0, s, M, h1e,
                       (const cplx*)Ga, (const cplx*)Gb, enuc, (cplx*)e_1b);
  else
    hipLaunchKernelGGL((afq::k_e1b_fullg<false>), dim3(W), dim3(afq::NT), 0, s, M, h1e,
                       (const cplx*)Ga, (const cplx*)Gb, enuc, (cplx*)e_1b);
  AFQ_CHECK_LAUNCH();
  afq::E2bScratch sc;
  int rc = afq::e2b_scratch(ctx, W, Na, Nb, G, 0, &sc);
  if (rc) return rc;
  rc = afq::two_body(s, sc, W, M, Na, Nb, G, Gha, Ghb, ra, rb, dtype == AFQ_COMPLEX, sc.slot[0],
                     sc.slot[1], sign);
  if (rc) return rc;
  return afq::combine_launch(s, W, e_1b, sc.slot[0], sc.slot[1], e_tot, e_2b);
}

int afq_local_energy_streamed(afq_ctx* ctx, int W, int M, int Na, int Nb, int G,
                              const double* hTa, const double* hTb, const double* Gha,
                              const double* Ghb, const double* ra, const double* rb, int dtype,
                              double enuc, int e_chunk_g, double* e_tot, double* e_1b,
                              double* e_2b) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && G >= 0 && e_chunk_g >= 1, "bad shape");
  AFQ_REQUIRE(dtype == AFQ_REAL || dtype == AFQ_COMPLEX, "bad dtype %d", dtype);
  AFQ_REQUIRE(e_1b != nullptr, "e_1b is required");
  if (W == 0) return AFQ_OK;
  hipStream_t s = ctx->stream;
  int rc = afq::e1b_half_launch(s, W, M, Na, Nb, hTa, hTb, Gha, Ghb, enuc, e_1b);
  if (rc) return rc;
  afq::E2bScratch sc;
  rc = afq::e2b_scratch(ctx, W, Na, Nb, G, e_chunk_g, &sc);
  if (rc) return rc;
  rc = afq::two_body(s, sc, W, M, Na, Nb, G, Gha, Ghb, ra, rb, dtype == AFQ_COMPLEX, sc.slot[0],
                     sc.slot[1]);
  if (rc) return rc;
  return afq::combine_launch(s, W, e_1b, sc.slot[0], sc.slot[1], e_tot, e_2b);
}

int afq_local_energy_multidet_streamed(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, int nd,
                                       const double* hTa, const double* hTb, const double* Gha,
                                       const double* Ghb, const double* ra, const double* rb,
                                       int dtype, const double* ci, const double* oa_all,
                                       const double* ob_all, double enuc, int e_chunk_g,
                                       double* e_tot, double* e_1b, double* e_2b) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && G >= 0 && nd > 0 && e_chunk_g >= 1,
              "bad shape");
  AFQ_REQUIRE(dtype == AFQ_REAL || dtype == AFQ_COMPLEX, "bad dtype %d", dtype);
  AFQ_REQUIRE(e_tot && e_1b && e_2b && ci && oa_all && ob_all, "NULL argument");
  AFQ_REQUIRE((Na == 0 || (hTa && Gha && (ra || G == 0))) && (Nb == 0 || (hTb && Ghb && (rb || G == 0))),
              "NULL per-det array");
  if (W == 0) return AFQ_OK;
  hipStream_t s = ctx->stream;
  afq::E2bScratch sc;
  int rc = afq::e2b_scratch(ctx, W, Na, Nb, G, e_chunk_g, &sc);
  if (rc) return rc;
  const int rw = dtype == AFQ_COMPLEX ? 2 : 1;
  for (int I = 0; I < nd; ++I) {
    const double* ga = Gha ? Gha + 2L * I * W * Na * M : nullptr;
    const double* gb = Ghb ? Ghb + 2L * I * W * Nb * M : nullptr;
    // per-det one-body trace (enuc added once, at the end)
    if ((rc = afq::e1b_half_launch(s, W, M, Na, Nb, hTa ? hTa + 2L * I * M * Na : nullptr,
                                   hTb ? hTb + 2L * I * M * Nb : nullptr, ga, gb, 0.0,
                                   sc.slot[2])))
      return rc;
    if ((rc = afq::two_body(s, sc, W, M, Na, Nb, G, ga, gb,
                            ra ? ra + (long)rw * I * G * Na * M : nullptr,
                            rb ? rb + (long)rw * I * G * Nb * M : nullptr, dtype == AFQ_COMPLEX,
                            sc.slot[0], sc.slot[1])))
      return rc;
    hipLaunchKernelGGL(afq::k_md_stream_acc, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, I,
                       (const cplx*)ci, (const cplx*)oa_all, (const cplx*)ob_all,
                       (const cplx*)sc.slot[2], (const cplx*)sc.slot[0], (const cplx*)sc.slot[1],
                       (cplx*)e_1b, (cplx*)e_2b);
    AFQ_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(afq::k_md_stream_final, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, nd,
                     (const cplx*)ci, (const cplx*)oa_all, (const cplx*)ob_all, enuc,
                     (cplx*)e_tot, (cplx*)e_1b, (cplx*)e_2b);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // extern "C"
