// Cavity-QED phaseless AFQMC (OmegaQMC/qed_afqmc_gto.py): the pieces the QED
// step adds to the electronic one.  Each walker carries a photon coordinate
// q; the Pauli-Fierz Hamiltonian in the dipole gauge adds
//   * the photon half steps of the exact harmonic-oscillator kernel with the
//     Gaussian trial f(q) = (s/pi)^(1/4) exp(-s (q - q0)^2 / 2) as importance
//     function (propagate_photon, qed_afqmc_gto.py:262-311),
//   * the one-body coupling exp(-dt/2 sqrt(Omega) q d) applied to the
//     walkers (afq_apply_exp_scaled, the scaled-matrix mode of the FP64
//     propagator; qed_afqmc_gto.py:314-332),
//   * the photon weight factor exp(clip(log w1 + log w2, -10, 10))
//     (qed_afqmc_gto.py:436-438),
//   * the local energy terms sqrt(Omega) q Tr(d G) and
//     Omega/2 (s - s^2 (q - q0)^2 + q^2 - 1) (qed_afqmc_gto.py:156-201),
//   * the photon coordinate's share of the comb (qed_afqmc_gto.py:106-153).
// The dipole self-energy enters as one more Cholesky vector (the caller's
// augmented chol), so the two-body part is the plain step's.
#include "afq_kernels.h"
#include "afq_mfma.h"

namespace afq {

namespace {

__global__ void k_qed_photon(int W, const double* __restrict__ q, const double* __restrict__ eta,
                             double omega, double dt, double s, double q0,
                             double* __restrict__ q_out, double* __restrict__ log_w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  const double delta = omega * dt / 2.0;
  const double ch = cosh(delta), th = tanh(delta);
  const double qi = q[i];
  const double drift = -s * (qi - q0);
  const double qn = qi / ch + th * drift + sqrt(th) * eta[i];
  const double ltr = -s / 2.0 * ((qn - q0) * (qn - q0) - (qi - q0) * (qi - q0));
  const double corr = (qi / ch - qn) * drift;
  const double add = -th / 2.0 * (-(drift * drift) + qi * qi) + delta / 2.0;
  log_w[i] = ltr + corr + add - 0.5 * log(ch);
  q_out[i] = qn;
}

__global__ void k_qed_weights(int W, double* __restrict__ w, const double* __restrict__ l1,
                              const double* __restrict__ l2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  w[i] *= exp(fmin(fmax(l1[i] + l2[i], -10.0), 10.0));
}

__global__ void k_gather_values(int W, const int32_t* __restrict__ idx,
                                const double* __restrict__ src, double* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  const int j = idx[i];
  dst[i] = (j >= 0 && j < W) ? src[j] : __builtin_nan("");
}

// e_1b += sqrt(Omega) q trdip; e_ph = Omega/2 (s - s^2 (q - q0)^2 + q^2 - 1);
// e_tot = e_1b + e_2b + e_ph.  trdip arrives in e_ph.
__global__ void k_qed_energy(int W, const double* __restrict__ q, double omega, double s, double q0,
                             cplx* __restrict__ e_tot, cplx* __restrict__ e_1b,
                             const cplx* __restrict__ e_2b, cplx* __restrict__ e_ph) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  const double qi = q[i];
  const cplx td = e_ph[i];
  const double sq = sqrt(omega) * qi;
  const cplx e1 = cadd(e_1b[i], mkc(sq * td.x, sq * td.y));
  const double d = qi - q0;
  const double ph = omega / 2.0 * (s - s * s * d * d + qi * qi - 1.0);
  e_1b[i] = e1;
  e_ph[i] = mkc(ph, 0.0);
  e_tot[i] = cadd(cadd(e1, e_2b[i]), mkc(ph, 0.0));
}

// V[w] = coef[w] * mat as complex (the batched-GEMM form of the scaled
// exponential, for M beyond the fused propagator)
__global__ void k_scaled_matrix(long n, long MM, const double* __restrict__ coef,
                                const double* __restrict__ mat, cplx* __restrict__ V) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x)
    V[e] = mkc(coef[e / MM] * mat[e % MM], 0.0);
}

}  // namespace

}  // namespace afq

extern "C" {

int afq_qed_propagate_photon(afq_ctx* ctx, int W, const double* q, const double* eta,
                             double omega, double dt, double s, double q0, double* q_out,
                             double* log_w) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && omega >= 0 && dt > 0 && s > 0, "bad argument");
  AFQ_REQUIRE(W == 0 || (q && eta && q_out && log_w), "NULL array");
  if (W == 0) return AFQ_OK;
  hipLaunchKernelGGL(afq::k_qed_photon, dim3(cdiv_int(W, 256)), dim3(256), 0, ctx->stream, W, q,
                     eta, omega, dt, s, q0, q_out, log_w);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int afq_apply_exp_scaled(afq_ctx* ctx, int W, int M, int N, const double* coef, const double* mat,
                         const double* phi, int nmax, double* phi_out) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && N >= 0 && nmax >= 0, "bad shape");
  AFQ_REQUIRE(W == 0 || N == 0 || (coef && mat && phi && phi_out), "NULL array");
  if (W == 0 || N == 0) return AFQ_OK;
  if (afq::prop_f64_fused_fits(M))
    return afq::prop_scaled_f64(ctx->stream, W, M, N, mat, coef, phi, phi_out, nmax);
  // batched GEMMs on the explicit (W, M, M) matrices; an aliased phi is
  // copied first (the Taylor sum accumulates into phi_out while reading phi)
  const size_t nv = align256(16UL * W * M * M), nt = align256(16UL * W * M * N);
  char* scr = (char*)afq::scratch(ctx, nv + 3 * nt);
  if (!scr) return AFQ_ENOMEM;
  const long n = (long)W * M * M;
  hipLaunchKernelGGL(afq::k_scaled_matrix, dim3((unsigned)std::min<long>((n + 255) / 256, 16384)),
                     dim3(256), 0, ctx->stream, n, (long)M * M, coef, mat, (cplx*)scr);
  AFQ_CHECK_LAUNCH();
  const double* src = phi;
  if (phi_out == phi) {
    double* c = (double*)(scr + nv);
    int rc = afq::copy_async(ctx->stream, phi, c, 2L * W * M * N);
    if (rc) return rc;
    src = c;
  }
  return afq::apply_exp_vhs_fp64(ctx->stream, W, M, N, (const double*)scr, src, nmax, phi_out,
                                 (double*)(scr + nv + nt), (double*)(scr + nv + 2 * nt));
}

int afq_qed_photon_weights(afq_ctx* ctx, int W, double* weights, const double* log_w1,
                           const double* log_w2) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0, "bad shape");
  if (W == 0) return AFQ_OK;
  AFQ_REQUIRE(weights && log_w1 && log_w2, "NULL array");
  hipLaunchKernelGGL(afq::k_qed_weights, dim3(cdiv_int(W, 256)), dim3(256), 0, ctx->stream, W,
                     weights, log_w1, log_w2);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int afq_gather_walker_values(afq_ctx* ctx, int W, const int32_t* idx, const double* src,
                             double* dst) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0, "bad shape");
  if (W == 0) return AFQ_OK;
  AFQ_REQUIRE(idx && src && dst && src != dst, "NULL or aliased array");
  hipLaunchKernelGGL(afq::k_gather_values, dim3(cdiv_int(W, 256)), dim3(256), 0, ctx->stream, W,
                     idx, src, dst);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int afq_qed_local_energy(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, const double* h1e,
                         const double* dip, const double* Ga, const double* Gb,
                         const double* Ghalfa, const double* Ghalfb, const double* rchol_a,
                         const double* rchol_b, double enuc, const double* q, double omega,
                         double s, double q0, double* e_tot, double* e_1b, double* e_2b,
                         double* e_ph) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0, "bad shape");
  if (W == 0) return AFQ_OK;
  AFQ_REQUIRE(dip && q && e_tot && e_1b && e_2b && e_ph, "NULL array");
  int rc = afq_local_energy(ctx, W, M, Na, Nb, G, h1e, Ga, Gb, Ghalfa, Ghalfb, rchol_a, rchol_b,
                            AFQ_REAL, enuc, e_tot, e_1b, e_2b);
  if (rc) return rc;
  // Tr(d Ga) + Tr(d Gb) into e_ph, then the photon terms
  if ((rc = afq::e1b_fullg_launch(ctx->stream, W, M, dip, Ga, Gb, 0.0, e_ph))) return rc;
  hipLaunchKernelGGL(afq::k_qed_energy, dim3(cdiv_int(W, 256)), dim3(256), 0, ctx->stream, W, q,
                     omega, s, q0, (cplx*)e_tot, (cplx*)e_1b, (const cplx*)e_2b, (cplx*)e_ph);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // extern "C"
