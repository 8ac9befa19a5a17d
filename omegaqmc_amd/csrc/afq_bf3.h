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
  int Ka, Kb;    // N_s * M padded to multiples of 32 (K of the force bias)
  // L_g block-packed: column c = blk*256 + i*16 + j  <->  (16P+i, 16Q+j), P <= Q;
  // stored column-major in g: Lb[plane][c][Gp]
  bf16_t* Lb;
  int2* blk_pq;  // (P, Q) of every block
  // Lhalf_s[g][i*M+q] as [plane][G][K_s]
  bf16_t* Lha;
  bf16_t* Lhb;
  // expH1 as [plane][16*RT][Kp]
  bf16_t* E;
};

int bf3_prepare(afq_system* s, Bf3System* b);
void bf3_release(Bf3System* b);

// x (W,G) complex -> Xh/Xl [2W][Gp] planes with rows (w, re/im).
// VHS in block-major bf16 planes: Vb[w][P][Q][plane][16][16], planes
// (re_hi, re_lo, im_hi, im_lo), both (P,Q) and (Q,P) blocks written.
int bf3_vhs(hipStream_t s, const Bf3System& b, int W, const bf16_t* X, double dt, bf16_t* Vb);
// Ghalf planes [2][2W][K_s] for both spins -> vbias partials (float2) part[s][W][G].
int bf3_force_bias(hipStream_t s, const Bf3System& b, int W, const bf16_t* Gha, const bf16_t* Ghb,
                   float* part, int* nsplit_out);
// phi <- expH1 . Taylor(VHS) . expH1 . phi in place, FP64 walkers.
int bf3_propagate(hipStream_t s, const Bf3System& b, int W, const bf16_t* Vb, double* phia,
                  double* phib, int nmax);
size_t bf3_vhs_bytes(const Bf3System& b, int W);
size_t bf3_x_bytes(const Bf3System& b, int W);
size_t bf3_gh_bytes(const Bf3System& b, int W);

}  // namespace afq
