// Split-BF16 ("BF16x3": a*b ~ ah*bh + ah*bl + al*bh, FP32 accumulate) operands
// and kernels of the reduced-precision step.
#pragma once
#include "afq_common.h"

namespace afq {

// Storage of the reduced-precision system (owned by afq_system::rchol_lp).
struct Bf3System {
  int M, Na, Nb, G;
  int RT;        // 16-row tiles of M
  int Gp;        // G padded to a multiple of 32
  int nblk;      // upper-triangle 16x16 blocks of an M x M matrix (RT(RT+1)/2)
  int Kp;        // 16*RT padded to a multiple of 32 (K of the propagator products)
  int ldq;       // M padded to a multiple of 32: row stride of (i, q) in Lhalf / Ghalf planes
  int Ka, Kb;    // N_s * ldq padded to multiples of 32 (K of the force bias)
  // L_g block-packed: column c = blk*256 + i*16 + j  <->  (16P+i, 16Q+j), P <= Q;
  // stored column-major in g: Lb[plane][c][Gp]
  bf16_t* Lb;
  int2* blk_pq;  // (P, Q) of every block
  // Lhalf_s[g][i*ldq+q] as [plane][G][K_s], zero in the padding
  bf16_t* Lha;
  bf16_t* Lhb;
  // expH1 as [plane][16*RT][Kp]
  bf16_t* E;
  // widest column block of the propagator (16, 32 or 48): a walker's
  // concatenated spin columns go to ceil(Na + Nb / pcols) workgroups
  int pcols = 48;
};

int bf3_prepare(afq_system* s, Bf3System* b);
void bf3_release(Bf3System* b);

// x (W,G) complex -> Xh/Xl [2W][Gp] planes with rows (w, re/im).
// VHS in block-major bf16 planes: Vb[w][P][Q][plane][16][16], planes
// (re_hi, re_lo, im_hi, im_lo), both (P,Q) and (Q,P) blocks written.
int bf3_vhs(hipStream_t s, const Bf3System& b, int W, const bf16_t* X, double dt, bf16_t* Vb);
// Ghalf planes [2][2W][K_s] for both spins -> vbias partials (float2) part[s][W][G].
// per_spin K splits per spin (1 or 2: at most four FP32 partial sets)
int bf3_force_bias(hipStream_t s, const Bf3System& b, int W, const bf16_t* Gha, const bf16_t* Ghb,
                   float* part, int* nsplit_out, int per_spin = 2);
// phi <- expH1 . Taylor(VHS) . expH1 . phi in place, FP64 walkers; t0 is a
// scratch of bf3_prop_scratch_bytes(b, W).
size_t bf3_prop_scratch_bytes(const Bf3System& b, int W);
int bf3_propagate(hipStream_t s, const Bf3System& b, int W, const bf16_t* Vb, float* t0,
                  double* phia, double* phib, int nmax);
// Exchange energy e_x[w] = 1/2 sum_s sum_g sum_ij T_ij T_ji, T = Lhalf_s[g] Ghalf_s^T,
// from the split-bf16 Ghalf planes (N_s <= 40); part: bf3_exchange_part_bytes.
// rows x cols FP64 -> split planes [2][rows][cols_p] (zero padding)
int bf3_split_rows(hipStream_t s, const double* A, int rows, int cols, int cols_p, bf16_t* out);
// vbias partials [nsplit][W][G] (FP32 re, im) = sum_c L[g][c] Gpk[w][c] over
// split planes of the packed triangle: gpl [2][2W][Kp] (rows (w, re/im)),
// lpl [2][G][Kp]; K split in up to four chunks
int bf3_packed_force_bias(hipStream_t s, int W, int G, int Kp, const bf16_t* gpl,
                          const bf16_t* lpl, float* part, int* nsplit_out);
bool bf3_exchange_supported(const Bf3System& b);
size_t bf3_exchange_part_bytes(const Bf3System& b, int W);
int bf3_exchange(hipStream_t s, const Bf3System& b, int W, const bf16_t* gha, const bf16_t* ghb,
                 double* part, double* ex);
// Two-body energy of one determinant of an occupation-string trial: Lhalf =
// rows occ_s of every L_g (chol (G, M, M) FP64), Ghalf_s (W, N_s, M) complex
// FP64, both split into scratch planes (bf3_md_two_body_bytes) once; the
// Coulomb X_g on the split-BF16 force-bias GEMM (ec, as bf3_coulomb), the
// exchange on k_exch_bf3 (ex).
size_t bf3_md_two_body_bytes(const Bf3System& b, int W);
int bf3_md_two_body(hipStream_t s, const Bf3System& b, int W, const int* occa, const int* occb,
                    const double* chol, const double* gha, const double* ghb, void* scratch,
                    double* ec, double* ex);
// Coulomb energy 1/2 sum_g X_g^2 from the force-bias partials part[nsplit][W][G].
int bf3_coulomb(hipStream_t s, int W, int G, int nsplit, const float* part, double* ec);
size_t bf3_vhs_bytes(const Bf3System& b, int W);
size_t bf3_x_bytes(const Bf3System& b, int W);
size_t bf3_gh_bytes(const Bf3System& b, int W);

}  // namespace afq
