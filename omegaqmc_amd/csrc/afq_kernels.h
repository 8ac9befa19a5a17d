// Internal launchers shared between the kernel files (not part of the ABI).
#pragma once
#include "afq_common.h"

namespace afq {

int copy_async(hipStream_t s, const double* src, double* dst, long n);
size_t pair_branch_lds(int W);
int zero_async(hipStream_t s, void* p, size_t bytes);  // bytes a multiple of 4

// greens (afq_greens.hip)
int greens_launch(hipStream_t s, int mode, int W, int M, int Na, int Nb, const double* phia,
                  const double* phib, const double* tu, const double* td, double* Gha,
                  double* Ghb, double* ovlp, const GhalfBf16* gbf = nullptr,
                  double* inv = nullptr);
// S^-1 buffer of greens_launch (identity trials, N <= 40: Ghalf by a second kernel)
size_t greens_inv_bytes(int W, int Na, int Nb);
int full_g_launch(hipStream_t s, int W, int M, int N, const double* trial, const double* Ghalf,
                  double* G);

// propagation (afq_propagate.hip)
int force_bias_gemm_fp64(hipStream_t s, int W, int M, int Na, int Nb, int G, const double* ra,
                         const double* rb, const double* Gha, const double* Ghb, double* vbiasT);
int shift_launch(hipStream_t s, int W, int G, const double* vbiasT, const double* mf,
                 const double* xi, double dt, double fb, double* xbar, double* xs, double* cfb,
                 double* cmf, int nsplit, const float* vpart = nullptr, bf16_t* Xp = nullptr,
                 int Gp = 0, const double* sign = nullptr);
int build_vhs_fp64(hipStream_t s, int W, int M, int G, const double* xs, const double* chol,
                   double dt, double* VHS);
int apply_exp_vhs_fp64(hipStream_t s, int W, int M, int N, const double* VHS, const double* phi,
                       int nmax, double* phi_out, double* tmp0, double* tmp1);
int one_body_fp64(hipStream_t s, int W, int M, int N, const double* expH1, const double* phi,
                  double* out);
// expH1 . exp(VHS) . expH1 by batched GEMMs (any M; scratch: a_s, b_s of
// W M N_s complex, t0, t1 of W M max(N_a, N_b) complex)
int prop_generic_f64(hipStream_t s, int W, int M, int Na, int Nb, const double* expH1,
                     const double* VHS, double* pa, double* pb, double* a1, double* a2,
                     double* b1, double* b2, double* t0, double* t1, int nmax);
// eshift_dev (optional): device energy shift used instead of eshift
int weights_launch(hipStream_t s, int W, const double* w, const double* oo, const double* on,
                   const double* cfb, const double* cmf, const double* eh, double eshift, double dt,
                   double* w_out, double* eh_out, const double* eshift_dev = nullptr);

// rng (afq_rng.hip)
// step_dev (optional): device step number used instead of step
int aux_fields_launch(hipStream_t s, uint64_t seed, int64_t step, int W, int G, double* xi,
                      const int64_t* step_dev = nullptr);
// r[i] = comb offset of seeds[i] at the step *step_dev
int comb_offsets_dev_launch(hipStream_t s, int n, const uint64_t* seeds, const int64_t* step_dev,
                            double* r);

// stabilisation / population control (afq_ortho.hip, afq_comb.hip)
int ortho_launch(hipStream_t s, int W, int M, int Na, int Nb, double* phia, double* phib,
                 double* log_detR, double* sign_detR, void* scratch);
size_t ortho_scratch_bytes(int W, int Na, int Nb);
// r_dev (optional): device comb offset used instead of r
int comb_launch(hipStream_t s, int W, int M, int Na, int Nb, const double* weights,
                const double* phia, const double* phib, double r, const double* eh,
                double* w_new, double* phia_new, double* phib_new, double* eh_new, int32_t* idx,
                double* cum_scratch, const double* r_dev = nullptr);
int gather_launch(hipStream_t s, int W, int M, int Na, int Nb, const int32_t* idx,
                  const double* phia, const double* phib, const double* eh, double* w_new,
                  double* phia_new, double* phib_new, double* eh_new);
// gate (optional): no clip while *gate <= 1 (the first driver step)
int clip_weights_launch(hipStream_t s, int W, double* weights, double frac, double* tmp,
                        const int64_t* gate = nullptr);
int block_stats_launch(hipStream_t s, int W, const double* weights, const double* e, double* out);
int accumulate_launch(hipStream_t s, int W, const double* weights, const double* eh, double* acc);

// energy (afq_energy.hip)
// fused FP64 MFMA exchange (real Lhalf, N <= EXCH_F64_NMAX): T stays in
// registers and LDS; partial: per (row tile, walker) sums, partial_bytes of
// them (the row tiles run in as many launches as they need)
constexpr int EXCH_F64_NMAX = 40;
bool exchange_f64_fused_fits(int N);
int exchange_f64_fused(hipStream_t s, int W, int M, int N, int G, const double* Ghalf,
                       const double* rchol, double* partial, size_t partial_bytes,
                       double* e_exch, bool accumulate, const double* sign);
int exchange_fp64(hipStream_t s, int W, int M, int N, int G, const double* Ghalf,
                  const double* rchol, bool rchol_cplx, int gchunk, double* partial,
                  double* e_exch, bool accumulate, const double* sign = nullptr);
int e1b_half_launch(hipStream_t s, int W, int M, int Na, int Nb, const double* hTa,
                    const double* hTb, const double* Gha, const double* Ghb, double enuc,
                    double* e1b);
int combine_launch(hipStream_t s, int W, const double* e1b, const double* ec, const double* ex,
                   double* etot, double* e2b);
// e1b[w] = sum_pq h[p][q] (Ga + Gb)[w][q][p] + enuc, h real
int e1b_fullg_launch(hipStream_t s, int W, int M, const double* h1e, const double* Ga,
                     const double* Gb, double enuc, double* e1b);
int coulomb_launch(hipStream_t s, int W, int M, int Na, int Nb, int G, const double* ra,
                   const double* rb, bool rc_cplx, const double* Gha, const double* Ghb,
                   double* X, int nx, double* e_coul, const double* sign = nullptr);

}  // namespace afq
