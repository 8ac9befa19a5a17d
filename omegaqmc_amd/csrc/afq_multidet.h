// Multi-determinant trials built from occupation strings (internal).
#pragma once
#include "afq_common.h"

namespace afq {

// Scratch of one multi-determinant evaluation for W walkers (bytes), and the
// pieces carved out of it.
struct MdBufs {
  double* gha = nullptr;   // Ghalf_a of every det: [I][w][Na][M] complex
  double* ghb = nullptr;   // Ghalf_b: [I][w][Nb][M]
  double* oda = nullptr;   // per-det overlaps [w][I] complex
  double* odb = nullptr;
  double* z = nullptr;     // masked weights c_I^* o_Ia o_Ib: [w][I] complex (0 if masked)
  int* keep = nullptr;     // [w][I]
  double* ovlp = nullptr;  // [w] complex
  double* Ga = nullptr;    // (W, M, M) complex
  double* Gb = nullptr;
};
size_t md_bytes(int W, int M, int Na, int Nb, int nd, bool ghalf, bool gfull);
MdBufs md_carve(void* base, int W, int M, int Na, int Nb, int nd, bool ghalf, bool gfull);

// Per-det overlaps (and with ghalf, every Ghalf_I), the masked weights and
// the multi-det overlap; with gfull also Ga, Gb (greens_th4.py:75-148).
// occ_* are device int32 (nd, N_s) with entries in [0, M).
int md_greens(hipStream_t s, int W, int M, int Na, int Nb, int nd, int chunk, const double* phia,
              const double* phib, const int* occa, const int* occb, const double* ci,
              bool ghalf, bool gfull, const MdBufs& b);

// Host check of occupation strings (device arrays): entries in [0, M), distinct per det.
int md_check_occ(hipStream_t s, int nd, int N, int M, const int32_t* occ);

int propagate_multidet(afq_system* sys, int W, double* phia, double* phib, double* weights,
                       double* e_hybrid, double eshift, const double* xi, uint64_t seed,
                       int64_t step, double* ovlp_new);
int system_energy_multidet(afq_system* sys, int W, const double* phia, const double* phib,
                           const double* h1e, double enuc, double* e_tot, double* e_1b,
                           double* e_2b);

}  // namespace afq
