// Multi-determinant trials built from occupation strings (internal).
#pragma once
#include "afq_common.h"

namespace afq {

// What one multi-determinant evaluation forms besides the overlaps:
//   MD_OVERLAP  overlaps only;
//   MD_INVERSE  every S_I^-1 and the scattered sum B_s = sum_I z_I T_I^* S_I^-1
//               / overlap, from which G_s = B_s phi_s^T (one GEMM per spin);
//   MD_GHALF    every Ghalf_I (the local energy needs them) and G.
// MD_GHALF_RAW: every det's Ghalf and overlaps as MD_GHALF, without zeroing
// the Ghalf of masked dets (greens.py:314-369 keeps them)
enum MdKind { MD_OVERLAP = 0, MD_INVERSE = 1, MD_GHALF = 2, MD_GHALF_RAW = 3 };

// Scratch of one multi-determinant evaluation for W walkers (bytes), and the
// pieces carved out of it.
struct MdBufs {
  double* gha = nullptr;   // Ghalf_a of every det: [I][w][Na][M] complex
  double* ghb = nullptr;   // Ghalf_b: [I][w][Nb][M]
  double* inva = nullptr;  // S_Ia^-1 of every det: [I][w][Na][Na] complex
  double* invb = nullptr;  // [I][w][Nb][Nb]
  double* Ba = nullptr;    // B_a: [w][M][Na] complex
  double* Bb = nullptr;    // [w][M][Nb]
  int* ioa = nullptr;      // dets on orbital p: [M][nd] (I << 6 | row), then counts [M]
  int* iob = nullptr;
  double* oda = nullptr;   // per-det overlaps [w][I] complex
  double* odb = nullptr;
  double* z = nullptr;     // masked weights c_I^* o_Ia o_Ib: [w][I] complex (0 if masked)
  int* keep = nullptr;     // [w][I]
  double* ovlp = nullptr;  // [w] complex
  double* Ga = nullptr;    // (W, M, M) complex
  double* Gb = nullptr;
};
size_t md_bytes(int W, int M, int Na, int Nb, int nd, int kind, bool gfull);
MdBufs md_carve(void* base, int W, int M, int Na, int Nb, int nd, int kind, bool gfull);

// Per-det overlaps, the masked weights and the multi-det overlap, plus what
// `kind` names (greens_th4.py:75-148).  With kind != MD_OVERLAP and b.Ga set,
// G is written: Ga and Gb, or with gsum their sum into Ga alone.  occ_* are
// device int32 (nd, N_s) with entries in [0, M).
int md_greens(hipStream_t s, int W, int M, int Na, int Nb, int nd, int chunk, const double* phia,
              const double* phib, const int* occa, const int* occb, const double* ci, int kind,
              bool gsum, const MdBufs& b);

// Host check of occupation strings (device arrays): entries in [0, M), distinct per det.
int md_check_occ(hipStream_t s, int nd, int N, int M, const int32_t* occ);

// Ratio updates relative to det 0 (afq_mdratio.hip): excitation tables
// (ok = false when a det is more than four orbitals from det 0), whether the
// kernel's LDS fits, and the launch: mode 0 the multi-det overlap, mode 1
// also B_s (W, M, N_s), so that G_s = B_s phi_s^T.
struct MdRatio;
int md_ratio_prepare(afq_system* sys, bool* ok);
void md_ratio_release(MdRatio* r);
bool md_ratio_fits(const afq_system* sys);
int md_ratio_launch(afq_system* sys, int mode, int W, const double* phia, const double* phib,
                    double* ovlp, double* Ba, double* Bb);
// G_s = B_s phi_s^T (one batched GEMM per spin), or their sum into Ga (gsum)
int md_g_from_b(hipStream_t s, int W, int M, int Na, int Nb, const double* phia,
                const double* phib, bool gsum, const MdBufs& b);

int propagate_multidet(afq_system* sys, int W, double* phia, double* phib, double* weights,
                       double* e_hybrid, double eshift, const double* xi, uint64_t seed,
                       int64_t step, double* ovlp_new);
int system_energy_multidet(afq_system* sys, int W, const double* phia, const double* phib,
                           const double* h1e, double enuc, double* e_tot, double* e_1b,
                           double* e_2b);

}  // namespace afq
