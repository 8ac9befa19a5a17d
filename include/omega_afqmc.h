/*
 * omega_afqmc.h -- C ABI of the MI355X (gfx950) phaseless-AFQMC walker step.
 *
 * Every entry point replaces one function of the reference's Python/JAX
 * operator API (myung-group/OmegaQMC); the citation above each declaration
 * names the function it stands in for.  Conventions:
 *
 *   - All array arguments are DEVICE pointers on the context's device,
 *     caller-owned, row-major (C order) with the reference's shapes.
 *   - "complex" arrays are interleaved complex128: (re, im) double pairs, so a
 *     complex (W, M, N) array is W*M*N*2 doubles.  "real" arrays are float64.
 *   - Calls are asynchronous on the context's stream; nothing synchronises
 *     unless the name says so.  Scratch memory is owned by the context and
 *     grows on demand (afq_ctx_reserve pre-sizes it so steady-state calls
 *     never allocate).
 *   - Return value: AFQ_OK (0) or a negative afq_status; afq_last_error()
 *     gives a message (thread-local).  Invalid shapes are rejected before any
 *     launch.  Numerical failures propagate as NaN/inf exactly as in the
 *     reference (no error code).
 *   - The reference draws auxiliary fields and comb offsets with jax.random
 *     inside the step; here they are inputs (xi, r) or come from the
 *     counter-based Philox stream of afq_aux_fields / afq_comb_offset.
 */
#ifndef OMEGA_AFQMC_H
#define OMEGA_AFQMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afq_ctx afq_ctx;
typedef struct afq_system afq_system;

typedef enum {
  AFQ_OK = 0,
  AFQ_EINVAL = -1,       /* bad argument / shape */
  AFQ_ENOMEM = -2,       /* device allocation failed */
  AFQ_EUNSUPPORTED = -3, /* shape outside what the kernels handle */
  AFQ_EHIP = -4          /* HIP runtime error */
} afq_status;

typedef enum {
  AFQ_FP64 = 0,   /* every contraction in FP64 */
  AFQ_FP32 = 1,   /* FP32 MFMA contractions, FP64 det/inv/QR/weights */
  AFQ_BF16X3 = 2  /* split-BF16 (hi*hi + hi*lo + lo*hi) MFMA, FP32 accumulate */
} afq_precision;

typedef enum { AFQ_REAL = 0, AFQ_COMPLEX = 1 } afq_dtype;

int afq_version(void);
const char* afq_last_error(void);

/* ---- context ------------------------------------------------------------ */
int afq_ctx_create(afq_ctx** out, int device, void* hip_stream);
int afq_ctx_destroy(afq_ctx* ctx);
int afq_ctx_set_stream(afq_ctx* ctx, void* hip_stream);
int afq_ctx_synchronize(afq_ctx* ctx);
int afq_ctx_reserve(afq_ctx* ctx, size_t scratch_bytes);

/* ---- Green's functions and overlaps (OmegaQMC/observables/greens.py) ---- */

/* greens.py:104-129 greens_function_force_bias(phia, phib, trial_up, trial_dn)
 * -> (Ghalfa (W,Na,M), Ghalfb (W,Nb,M), overlap (W,)), all complex.
 * phia (W,M,Na), phib (W,M,Nb), trial_up (M,Na), trial_dn (M,Nb): complex. */
int afq_greens_function_force_bias(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                                   const double* phia, const double* phib,
                                   const double* trial_up, const double* trial_dn,
                                   double* Ghalfa, double* Ghalfb, double* overlap);

/* greens.py:132-154 greens_function_overlap -> overlap (W,) complex. */
int afq_greens_function_overlap(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                                const double* phia, const double* phib,
                                const double* trial_up, const double* trial_dn,
                                double* overlap);

/* greens.py:157-198 greens_function -> (Ga (W,M,M), Gb (W,M,M), Ghalfa, Ghalfb, overlap). */
int afq_greens_function(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                        const double* phia, const double* phib,
                        const double* trial_up, const double* trial_dn,
                        double* Ga, double* Gb, double* Ghalfa, double* Ghalfb,
                        double* overlap);

/* ---- local energy (OmegaQMC/observables/energy.py) ----------------------- */

/* energy.py:51-91 local_energy_2body(Ghalfa, Ghalfb, rchol_a, rchol_b)
 * -> (e_coul (W,), e_exch (W,)) complex.  rchol_{a,b} (G,N,M) are real or
 * complex according to rchol_dtype. */
int afq_local_energy_2body(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown, int naux,
                           const double* Ghalfa, const double* Ghalfb,
                           const double* rchol_a, const double* rchol_b, int rchol_dtype,
                           double* e_coul, double* e_exch);

/* energy.py:94-130 local_energy(h1e, Ga, Gb, Ghalfa, Ghalfb, rchol_a, rchol_b, enuc)
 * -> (e_tot, e_1b, e_2b) (W,) complex.  h1e (M,M) and rchol share dtype. */
int afq_local_energy(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown, int naux,
                     const double* h1e, const double* Ga, const double* Gb,
                     const double* Ghalfa, const double* Ghalfb,
                     const double* rchol_a, const double* rchol_b, int dtype, double enuc,
                     double* e_tot, double* e_1b, double* e_2b);

/* afqmc_pw_heg.py:414-457 local_energy_2body_pw: the two-body energy with the
 * ERI sign of each vector, (pq|rs) = sum_g sign_g L_g,pq L_g,rs: Coulomb
 * 1/2 sum_g sign_g X_g^2 and exchange 1/2 sum_g sign_g sum_ij T_ij T_ji.
 * chol_sign: device (G,) of +1 / -1. */
int afq_local_energy_2body_pw(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                              int naux, const double* Ghalfa, const double* Ghalfb,
                              const double* rchol_a, const double* rchol_b, int rchol_dtype,
                              const double* chol_sign, double* e_coul, double* e_exch);

/* afqmc_pw_heg.py:460-471 local_energy_pw: afq_local_energy with the signed
 * two-body part above. */
int afq_local_energy_pw(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown, int naux,
                        const double* h1e, const double* Ga, const double* Gb,
                        const double* Ghalfa, const double* Ghalfb, const double* rchol_a,
                        const double* rchol_b, int dtype, double enuc, const double* chol_sign,
                        double* e_tot, double* e_1b, double* e_2b);

/* energy.py:218-252 local_energy_streamed(h1e_trial_a (M,Na), h1e_trial_b (M,Nb),
 * Ghalfa, Ghalfb, rchol_a, rchol_b, enuc, e_chunk_g).  h1e_trial_* complex.
 * e_chunk_g only fixes the g-slab order of the exchange sum (any value >= 1). */
int afq_local_energy_streamed(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown, int naux,
                              const double* h1e_trial_a, const double* h1e_trial_b,
                              const double* Ghalfa, const double* Ghalfb,
                              const double* rchol_a, const double* rchol_b, int rchol_dtype,
                              double enuc, int e_chunk_g,
                              double* e_tot, double* e_1b, double* e_2b);

/* ---- propagation pieces (OmegaQMC/afqmc_gto.py) ------------------------- */

/* afqmc_gto.py:472-482 _build_vhs(xshifted, chol, dt, chol_chunk_g)
 * xshifted (W,G) complex, chol (G,M,M) real -> VHS (W,M,M) complex
 * = i sqrt(dt) sum_g xshifted[w,g] chol[g]. */
int afq_build_vhs(afq_ctx* ctx, int nwalkers, int nbasis, int naux,
                  const double* xshifted, const double* chol, double dt, double* VHS);

/* afqmc_gto.py:451-469 _apply_exp_vhs(VHS, phi, nmax): phi_out = sum_n VHS^n phi / n!.
 * VHS (W,M,M), phi (W,M,N) complex.  phi_out may alias phi. */
int afq_apply_exp_vhs(afq_ctx* ctx, int nwalkers, int nbasis, int nocc,
                      const double* VHS, const double* phi, int nmax, double* phi_out);

/* afqmc_gto.py:678-679 one-body half step phi_out[w] = expH1 @ phi[w];
 * expH1 (M,M) real, phi (W,M,N) complex.  phi_out must not alias phi. */
int afq_apply_one_body(afq_ctx* ctx, int nwalkers, int nbasis, int nocc,
                       const double* expH1, const double* phi, double* phi_out);

/* afqmc_gto.py:682-706 force bias and shifted fields:
 * vbias[g,w] = sum_s sum_iq rchol_s[g,i,q] Ghalf_s[w,i,q];
 * xbar = -sqrt(dt) (i vbias - mf_shift)^T capped at |xbar| <= fbbound;
 * xshifted = xi - xbar; cmf = -sqrt(dt) sum_g xshifted mf_shift;
 * cfb = sum_g xi xbar - 1/2 sum_g xbar^2.
 * rchol real (G,N,M); Ghalf complex; mf_shift (G,) complex; xi (W,G) real.
 * Outputs complex: xbar, xshifted (W,G); cfb, cmf (W,).  xbar may be NULL. */
int afq_force_bias(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown, int naux,
                   const double* rchol_a, const double* rchol_b,
                   const double* Ghalfa, const double* Ghalfb,
                   const double* mf_shift, const double* xi, double dt, double fbbound,
                   double* xbar, double* xshifted, double* cfb, double* cmf);

/* afqmc_gto.py:527-571 _update_weights_phaseless(weights, ovlp_old, ovlp_new, cfb,
 * cmf, e_hybrid_old, eshift, dt) -> (weights_new, e_hybrid_new).
 * weights real (W,); the rest complex (W,).  Outputs may alias inputs. */
int afq_update_weights_phaseless(afq_ctx* ctx, int nwalkers, const double* weights,
                                 const double* ovlp_old, const double* ovlp_new,
                                 const double* cfb, const double* cmf,
                                 const double* e_hybrid_old, double eshift, double dt,
                                 double* weights_new, double* e_hybrid_new);

/* afqmc_gto.py:160-181 orthogonalize_walkers(phia, phib) -> (phia, phib, log_detR,
 * sign_detR).  In place on phia/phib; log_detR (W,) real, sign_detR (W,)
 * complex (may be NULL).  Thin QR with positive real diagonal R (the unique
 * one, as the reference's sign-fixed Householder QR): CholeskyQR2, with a
 * shifted CholeskyQR3 when the Gram matrix is ill-conditioned. */
int afq_orthogonalize_walkers(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                              double* phia, double* phib, double* log_detR, double* sign_detR);

/* afqmc_gto.py:184-238 population_control_comb(weights, phia, phib, rng_key, e_hybrid)
 * with the comb offset r in [0,1) given.  Writes weights_new (= 1), the
 * resampled walkers phia_new/phib_new (must not alias the inputs),
 * e_hybrid_new (if e_hybrid != NULL) and the survivor indices idx (W,) int32.
 * The running weight sum is sequential FP64, left to right. */
int afq_population_control_comb(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                                const double* weights, const double* phia, const double* phib,
                                double r, const double* e_hybrid,
                                double* weights_new, double* phia_new, double* phib_new,
                                double* e_hybrid_new, int32_t* idx);

/* No counterpart in the reference: afq_population_control_comb for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results; the running sums of W > ~20000 walkers (beyond LDS) go to a
 * fixed device array of AFQ_COMB_NOALLOC_MAX_WALKERS doubles (8 MiB, held
 * by every device the library runs on until the process ends) instead of
 * the context scratch.  W above the limit returns AFQ_EINVAL.  Calls that
 * overlap on one device share that array and race. */
#define AFQ_COMB_NOALLOC_MAX_WALKERS (1 << 20)
int afq_population_control_comb_noalloc(afq_ctx* ctx, int nwalkers, int nbasis, int nup,
                                        int ndown, const double* weights, const double* phia,
                                        const double* phib, double r, const double* e_hybrid,
                                        double* weights_new, double* phia_new,
                                        double* phib_new, double* e_hybrid_new, int32_t* idx);

/* afqmc_gto.py:229-238 (second half of population_control_comb): gather the
 * walkers and e_hybrid by given survivor indices idx (W,) int32 and reset the
 * weights to 1.  Lets a caller replay the population-control decisions of
 * another run (or of a global comb over several devices).  An index outside
 * [0, nwalkers) is not read: that output row (and its e_hybrid) becomes NaN. */
int afq_population_control_apply(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                                 const int32_t* idx, const double* phia, const double* phib,
                                 const double* e_hybrid, double* weights_new, double* phia_new,
                                 double* phib_new, double* e_hybrid_new);

/* afqmc_gto.py:241-284 (population_control_pair_branch), in place: walkers
 * ordered by weight (ties in index order, a stable argsort), the k-th
 * lightest paired with the k-th heaviest; when the heavy weight exceeds
 * 4 <w> and the light one is below 0.1 <w> (<w> = sum w / W, summed left to
 * right) the light walker becomes a copy of the heavy one and both weights
 * become their mean.  src (W,) int32, optional: the walker each walker was
 * copied from (itself when unchanged).  W <= AFQ_PAIR_BRANCH_MAX_WALKERS (one
 * workgroup sorts the weights in LDS). */
#define AFQ_PAIR_BRANCH_MAX_WALKERS 8192
int afq_population_control_pair_branch(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                                       double* weights, double* phia, double* phib, int32_t* src);

/* afqmc_gto.py:1428-1432 weight clip w <- clip(w, 0, frac * sum|w|) in place. */
int afq_clip_weights(afq_ctx* ctx, int nwalkers, double* weights, double frac);

/* afqmc_gto.py:1448-1451 per-step accumulators: acc[0] += sum|w|,
 * acc[1] += sum |w| Re(e_hybrid).  acc is a device array of 2 doubles. */
int afq_accumulate_hybrid(afq_ctx* ctx, int nwalkers, const double* weights,
                          const double* e_hybrid, double* acc);

/* afqmc_gto.py:1488-1495 block statistics: out[0] = sum|w| Re(e)/sum|w|,
 * out[1] = sum|w|, out[2] = sum w^2, out[3] = max w.  out: 4 device doubles. */
int afq_block_stats(afq_ctx* ctx, int nwalkers, const double* weights, const double* e_tot,
                    double* out);

/* ---- per-walker numerical status ---------------------------------------- */
/* The reference lets NaN/inf propagate through the step (numerical failures
 * surface as NaN energies) and masks only near-zero multi-det overlaps
 * (observables/greens.py:264-276, greens_th4.py:17-44).  afq_walker_status
 * writes one flag word per walker (bits below) from phi_a, phi_b (W,M,N_s)
 * complex, overlap (W,) complex, weights (W,) real and e_hybrid (W,)
 * complex; overlap, weights and e_hybrid may be NULL (not checked).  An
 * overlap with |o| <= zero_overlap_tol counts as zero.  summary (2 device
 * uint32, may be NULL): [0] = OR of every walker's flags, [1] = number of
 * walkers with a flag other than AFQ_WALKER_ZERO_WEIGHT (a walker removed
 * by the phaseless projection is not a numerical failure). */
enum afq_walker_flag {
  AFQ_WALKER_NONFINITE_PHI = 1,      /* NaN or inf in phi_a or phi_b */
  AFQ_WALKER_ZERO_OVERLAP = 2,       /* |overlap| <= zero_overlap_tol */
  AFQ_WALKER_NONFINITE_OVERLAP = 4,  /* overlap NaN or inf */
  AFQ_WALKER_BAD_WEIGHT = 8,         /* weight NaN, inf or negative */
  AFQ_WALKER_NONFINITE_EHYBRID = 16, /* hybrid energy NaN or inf */
  AFQ_WALKER_ZERO_WEIGHT = 32        /* weight exactly 0 (informational) */
};
int afq_walker_status(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                      const double* phia, const double* phib, const double* overlap,
                      const double* weights, const double* e_hybrid, double zero_overlap_tol,
                      uint32_t* flags, uint32_t* summary);

/* ---- counter-based auxiliary fields ------------------------------------ */

/* xi[w,g] ~ N(0,1): Philox4x32-10 counter (g/2, w, step, 0), key = seed,
 * Box-Muller.  xi (W,G) real. */
int afq_aux_fields(afq_ctx* ctx, uint64_t seed, int64_t step, int nwalkers, int naux, double* xi);
/* Comb offset r in [0,1) for a step (host computation, no device work). */
double afq_comb_offset(uint64_t seed, int64_t step);

/* ---- multi-determinant trials ------------------------------------------ */
/* The determinants are occupation strings, trial_s[I] = eye(M)[:, occ_s[I]]
 * (the driver's construction, afqmc_gto.py:1163-1174): occ_up (ndet, nup) and
 * occ_dn (ndet, ndown) are device int32 arrays of distinct orbital indices in
 * [0, nbasis) (checked on the host: these calls synchronise), ci (ndet,)
 * complex.  Dets whose |o_Ia o_Ib| is below max(1e-14 max_J |o_J|, 1e-300)
 * are masked as in greens_th4.py:34-46; det chunks of det_chunk_size are
 * summed in the order of the reference's scan.  N_s <= 40. */

/* greens_th4.py:75-148 greens_function_multidet -> Ga, Gb (W,M,M), overlap (W,) complex */
int afq_greens_function_multidet(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                                 int ndet, const double* phia, const double* phib,
                                 const int32_t* occ_up, const int32_t* occ_dn,
                                 const double* ci_coeffs, int det_chunk_size,
                                 double* Ga, double* Gb, double* overlap);

/* greens_th4.py:186-228 greens_function_multidet_overlap_only -> overlap (W,) */
int afq_greens_function_multidet_overlap_only(afq_ctx* ctx, int nwalkers, int nbasis, int nup,
                                              int ndown, int ndet, const double* phia,
                                              const double* phib, const int32_t* occ_up,
                                              const int32_t* occ_dn, const double* ci_coeffs,
                                              int det_chunk_size, double* overlap);

/* energy_th4.py:26-119 local_energy_multidet: h1e (M,M) real, chol (G,M,M)
 * real -> e_tot, e_1b, e_2b (W,) complex (e_2b may be NULL); FP64. */
int afq_local_energy_multidet(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                              int naux, int ndet, const double* h1e, const double* chol,
                              const double* phia, const double* phib, const int32_t* occ_up,
                              const int32_t* occ_dn, const double* ci_coeffs, double enuc,
                              int det_chunk_size, double* e_tot, double* e_1b, double* e_2b);

/* greens.py:314-369 greens_function_multidet_force_bias: every det's Ghalf
 * -> Ghalfa_all (ndet,W,Na,M), Ghalfb_all (ndet,W,Nb,M) complex, non-finite
 * entries set to 0 and no small-overlap mask; overlap (W,) = sum_I
 * conj(c_I) o_Ia o_Ib over all dets; per-det spin overlaps ovlp_a_all,
 * ovlp_b_all (ndet,W) complex (may be NULL).  Dets as occupation strings. */
int afq_greens_function_multidet_force_bias(afq_ctx* ctx, int nwalkers, int nbasis, int nup,
                                            int ndown, int ndet, const double* phia,
                                            const double* phib, const int32_t* occ_up,
                                            const int32_t* occ_dn, const double* ci_coeffs,
                                            double* Ghalfa_all, double* Ghalfb_all,
                                            double* overlap, double* ovlp_a_all,
                                            double* ovlp_b_all);

/* energy.py:318-395 local_energy_multidet_streamed (the memory-lean
 * estimator of afqmc_gto_estream.py): per det, the one-body trace against
 * h1e_trials (ndet,M,N_s) complex (h1e @ T_I^*) and the two-body energy of
 * Ghalf_I with rchols (ndet,G,N_s,M) (dtype AFQ_REAL or AFQ_COMPLEX) in
 * slabs of e_chunk_g vectors; weighted by conj(c_I) o_Ia o_Ib (per-det
 * overlaps (ndet,W) complex) over their sum.  No (W,M,M) Green's function is
 * formed.  Outputs (W,) complex. */
int afq_local_energy_multidet_streamed(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown,
                                       int naux, int ndet, const double* h1e_trials_a,
                                       const double* h1e_trials_b, const double* Ghalfa_all,
                                       const double* Ghalfb_all, const double* rchols_a,
                                       const double* rchols_b, int dtype, const double* ci_coeffs,
                                       const double* ovlp_a_all, const double* ovlp_b_all,
                                       double enuc, int e_chunk_g, double* e_tot, double* e_1b,
                                       double* e_2b);

/* System with a multi-determinant trial (afqmc_gto.py:1163-1225): the
 * propagator (expH1, mf_shift from the |c|^2-weighted charge density) is the
 * caller's; occ/ci are copied.  afq_propagate_walkers then runs
 * propagate_walkers_multidet (afqmc_gto.py:728-839: force bias from the
 * multi-det G, overlap_new from the multi-det overlap) and
 * afq_system_local_energy runs local_energy_multidet.  precision AFQ_FP64
 * or AFQ_BF16X3; BF16X3 applies to the force bias, the VHS build, the
 * propagator and each determinant's exchange energy (the determinants,
 * Green's functions, Coulomb and one-body energy stay FP64). */
int afq_system_create_multidet(afq_ctx* ctx, afq_system** out, int nbasis, int nup, int ndown,
                               int naux, const double* chol, const double* expH1,
                               const double* mf_shift, int ndet, const int32_t* occ_up,
                               const int32_t* occ_dn, const double* ci_coeffs, double dt,
                               double fbbound, int exp_nmax, int det_chunk_size, int precision);

/* How the step of a multi-det system forms the multi-det overlaps and G:
 * 0 one Gauss-Jordan inverse per (walker, det) (greens_th4.py:75-148);
 * 1 ratio updates relative to det 0 -- one inverse per walker, each det's
 * overlap ratio det Q_I by the matrix determinant lemma and its G by a
 * Woodbury correction (the evaluator of wick.py:239-284, 425-485 used by
 * propagate_walkers_wick, afqmc_gto.py:842-935).  Method 1 is the default
 * when every det is within four orbitals of det 0; setting it otherwise
 * returns AFQ_EUNSUPPORTED.  The block energy always takes every det's
 * Green's function explicitly.  get returns the method in use. */
int afq_system_set_multidet_method(afq_system* sys, int method);
int afq_system_get_multidet_method(afq_system* sys);

/* ---- prepared system and the fused step -------------------------------- */

/* Setup counterpart of build_propagator (afqmc_gto.py:574-627) + the half
 * rotation (cholesky.py:470-510): keeps references to the caller's device
 * arrays chol (G,M,M) real, rchol_a (G,Na,M) / rchol_b (G,Nb,M) real,
 * expH1 (M,M) real, mf_shift (G,) complex, trial_up (M,Na) / trial_dn
 * (M,Nb) complex, and builds the reduced-precision copies the chosen
 * precision needs.  The caller keeps those arrays alive and unchanged.
 * AFQ_BF16X3 needs every L_g exactly symmetric (AFQ_EUNSUPPORTED otherwise). */
int afq_system_create(afq_ctx* ctx, afq_system** out, int nbasis, int nup, int ndown, int naux,
                      const double* chol, const double* rchol_a, const double* rchol_b,
                      const double* expH1, const double* mf_shift,
                      const double* trial_up, const double* trial_dn,
                      double dt, double fbbound, int exp_nmax, int precision);
int afq_system_destroy(afq_system* sys);

/* Optional ERI sign of each Cholesky vector, (pq|rs) = sum_g sign_g L_g,pq
 * L_g,rs: the symmetrised plane-wave vectors of the electron gas
 * (afqmc_pw_heg.py:161-254, chol_sign).  The step then uses the HS phase i
 * for sign +1 and 1 for sign -1 (afqmc_pw_heg.py:895-922, chol_phase) and the
 * local energy weights Coulomb and exchange by the sign
 * (afqmc_pw_heg.py:414-457).  chol_sign: device (G,) of +1 / -1, kept alive
 * and unchanged by the caller; NULL restores all +1.  Negative entries need
 * an FP64 or FP32 single-determinant system (AFQ_EUNSUPPORTED otherwise: the
 * split-BF16 VHS assumes symmetric L_g, the electron gas's sign -1 vectors are
 * antisymmetric); the block energy of a signed FP32 system is the FP64
 * estimator's. */
int afq_system_set_chol_sign(afq_system* sys, const double* chol_sign);

/* K-splits of the force-bias contraction (no counterpart in the reference,
 * which contracts it in one einsum): the tile choice of afq_batch's
 * autotuner.  nsplit in {2, 4, 8} (BF16X3: {2, 4}); default 4. */
int afq_system_set_force_bias_splits(afq_system* sys, int nsplit);

/* Widest column block of the BF16X3 propagator (no counterpart in the
 * reference): a walker's Na + Nb columns are cut into ceil((Na + Nb) / cols)
 * equal workgroup blocks; wider blocks reuse the walker's VHS over more
 * columns, narrower ones give more workgroups to small systems.  cols in
 * {16, 32, 48}, default 48; the autotuner's second knob.  Other precisions
 * accept and ignore it (their propagators use 16-column blocks).  Set it
 * before a batch captures the system's steps (the workspace is re-laid). */
int afq_system_set_propagator_block(afq_system* sys, int cols);

/* afqmc_gto.py:630-725 propagate_walkers(...) for one step, in place on
 * phia/phib/weights/e_hybrid; overlap_new (W,) complex is written.  The
 * reference ignores its `overlap` argument, so there is none here.  xi (W,G)
 * real; NULL draws it from the Philox stream (seed, step). */
int afq_propagate_walkers(afq_system* sys, int nwalkers, double* phia, double* phib,
                          double* weights, double* e_hybrid, double eshift,
                          const double* xi, uint64_t seed, int64_t step,
                          double* overlap_new);

/* Block-end mixed-estimator energy of the system's walkers
 * (afqmc_gto.py:1466-1473: greens_function + local_energy) without full G:
 * E_1b via h1e Psi* (energy.py:193-215), exchange with the system's
 * precision.  h1e (M,M) real.  Outputs (W,) complex. */
int afq_system_local_energy(afq_system* sys, int nwalkers, const double* phia,
                            const double* phib, const double* h1e, double enuc,
                            double* e_tot, double* e_1b, double* e_2b);

/* Per-phase timing of afq_propagate_walkers with HIP events on the system's
 * stream (no counterpart in the reference, which times whole blocks with
 * perf_counter at afqmc_gto.py:1380-1382).  Phases, in order: greens,
 * one-body, force-bias GEMM, shift, VHS build, Taylor, overlap, weights,
 * aux-field RNG.  on: 0 stops timing, 1 times every phase, a negative value
 * -m times only the phases p with bit p of m set (two event records per timed
 * phase and step on the stream).  read() synchronises the stream and returns
 * the summed milliseconds and the number of timed intervals per phase. */
int afq_system_timing_enable(afq_system* sys, int on);
int afq_system_timing_read(afq_system* sys, double* total_ms, int64_t* launches, int nphase,
                           int reset);

/* ---- cavity QED (OmegaQMC/qed_afqmc_gto.py) ----------------------------- */
/* Walkers carry a photon coordinate q (W,) real.  The QED step is the
 * electronic step of afq_force_bias / afq_build_vhs / afq_apply_exp_vhs /
 * afq_apply_one_body with the augmented Cholesky vectors (dipole self-energy
 * as one more vector, integrals/qed.py:130-190), wrapped by the calls below
 * (order of propagate_qed_walkers, qed_afqmc_gto.py:335-440). */

/* qed_afqmc_gto.py:262-311 propagate_photon: one half step of the exact
 * harmonic-oscillator kernel (delta = omega dt / 2), importance-sampled with
 * the Gaussian trial (s, q0); eta (W,) the N(0,1) draws.  Writes q_out and
 * the log weight factor log_w (W,). */
int afq_qed_propagate_photon(afq_ctx* ctx, int nwalkers, const double* q, const double* eta,
                             double omega, double dt, double s, double q0, double* q_out,
                             double* log_w);

/* qed_afqmc_gto.py:314-332 _apply_exp_vhs_photon: phi_out[w] =
 * sum_{n<=nmax} (coef[w] mat)^n phi[w] / n!; coef (W,) and mat (M,M) real,
 * phi (W,M,N) complex; phi_out may alias phi.  (The reference applies it
 * with nmax = 4 and coef = -dt/2 sqrt(omega) q.) */
int afq_apply_exp_scaled(afq_ctx* ctx, int nwalkers, int nbasis, int nocc, const double* coef,
                         const double* mat, const double* phi, int nmax, double* phi_out);

/* qed_afqmc_gto.py:436-438: weights *= exp(clip(log_w1 + log_w2, -10, 10)). */
int afq_qed_photon_weights(afq_ctx* ctx, int nwalkers, double* weights, const double* log_w1,
                           const double* log_w2);

/* qed_afqmc_gto.py:106-153 (the photon share of population_control_comb_qed):
 * dst[i] = src[idx[i]] for the survivor indices of afq_population_control_comb;
 * an index outside [0, W) gives NaN. */
int afq_gather_walker_values(afq_ctx* ctx, int nwalkers, const int32_t* idx, const double* src,
                             double* dst);

/* qed_afqmc_gto.py:156-201 qed_local_energy: the local_energy of energy.py:94-130
 * (h1e, augmented rchol, real) plus e_ep = sqrt(omega) q Tr(dip (Ga + Gb))
 * in e_1b and e_ph = omega/2 (s - s^2 (q - q0)^2 + q^2 - 1);
 * e_tot = e_1b + e_2b + e_ph.  dip (M,M) real; outputs (W,) complex. */
int afq_qed_local_energy(afq_ctx* ctx, int nwalkers, int nbasis, int nup, int ndown, int naux,
                         const double* h1e, const double* dip, const double* Ga,
                         const double* Gb, const double* Ghalfa, const double* Ghalfb,
                         const double* rchol_a, const double* rchol_b, double enuc,
                         const double* q, double omega, double s, double q0, double* e_tot,
                         double* e_1b, double* e_2b, double* e_ph);

/* ---- uneven batches of independent systems ----------------------------- */
/* The driver step of afqmc_gto.py:1383-1451 (QR every stabilize_freq steps
 * before propagating, the fused step, weight clip from step 2 on, comb every
 * pop_control_freq steps, hybrid-energy accumulators), run for many
 * single-determinant systems of any sizes, precisions and walker counts at
 * once.  The reference runs one system per driver (afqmc_gto.py:1257-1546);
 * here every system keeps its own walkers (allocated by the batch and
 * initialised to its trial, afqmc_gto.py:76-123), RNG seed (fields from the
 * Philox stream, comb offsets from afq_comb_offset) and energy shift, and
 * the systems share HIP streams: system i runs on stream stream_of[i] (NULL:
 * i mod nstreams) of the batch's nstreams (1..16).  While a batch lives, its
 * systems must not be stepped elsewhere with more walkers (the captured
 * steps point into their workspaces). */
typedef struct afq_batch afq_batch;
int afq_batch_create(afq_ctx* ctx, afq_batch** out, int nsys, afq_system* const* systems,
                     const int* nwalkers, const uint64_t* seeds, int stabilize_freq,
                     int pop_control_freq, int nstreams, const int* stream_of);
/* Energy shift of each system (nsys host doubles, the hybrid-energy level of
 * afqmc_gto.py:1357-1359, 1497-1502); read by the steps from device memory. */
int afq_batch_set_eshift(afq_batch* b, const double* eshift);
/* Advance every system by nsteps driver steps, ordered after the work
 * already on the context's stream (which waits for the batch in turn).
 * mode 0: the systems one after another on one stream (one driver at a
 * time, as the reference); 1: eager launches on the batch's streams; 2: one
 * hipGraph per driver cycle (lcm of the two frequencies) of all systems,
 * captured after one eager cycle, partial cycles eager.  All modes give the
 * same results. */
int afq_batch_run(afq_batch* b, int nsteps, int mode);
/* Device pointers of system i's current walker state (any may be NULL):
 * phia (W,M,Na), phib (W,M,Nb), weights (W,), e_hybrid (W,), overlap (W,)
 * of the last step, acc[2] the accumulators of afq_accumulate_hybrid; and
 * the number of steps run. */
int afq_batch_state(afq_batch* b, int i, double** phia, double** phib, double** weights,
                    double** e_hybrid, double** overlap, double** acc, int64_t* step);
/* Copy system i's current state into caller buffers (any may be NULL), on
 * the context's stream. */
int afq_batch_copy_state(afq_batch* b, int i, double* phia, double* phib, double* weights,
                         double* e_hybrid, double* overlap, double* acc);
/* Zero every system's hybrid-energy accumulators (acc[2]), on the context's
 * stream: the per-block reset of acc_weight / acc_ehybrid
 * (afqmc_gto.py:1384-1385) between afq_batch_run calls. */
int afq_batch_zero_accumulators(afq_batch* b);
int afq_batch_destroy(afq_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* OMEGA_AFQMC_H */
