// Step internals shared between the FP64 and reduced-precision paths.
#pragma once
#include "afq_common.h"

namespace afq {

// FP64 operands of a system (packed L, identity-trial detection)
int prepare_fp64(afq_system* s);
int propagate_fp64(afq_system* sys, int W, double* phia, double* phib, double* weights,
                   double* e_hybrid, double eshift, const double* xi, uint64_t seed, int64_t step,
                   double* ovlp_new);
int system_energy_fp64(afq_system* sys, int W, const double* phia, const double* phib,
                       const double* h1e, double enuc, double* e_tot, double* e_1b, double* e_2b);

// FP64 workspace pieces of the block-end energy (inside the step workspace).
struct EnergySlots {
  double* gha;    // Ghalf_a (W, Na, M) complex
  double* ghb;    // Ghalf_b (W, Nb, M) complex
  double* X;      // (W, G) complex
  double* ec;     // (W,) complex
  double* ex;     // (W,) complex
  double* part;   // W M M complex of scratch
  size_t part_bytes;
};
// Ghalf (FP64, plus split-bf16 planes when gbf != NULL), the overlap, and
// E_1b = Tr(h G) + E_nuc through h1e Psi* (energy.py:193-215).
int energy_prologue(afq_system* sys, int W, const double* phia, const double* phib,
                    const double* h1e, double enuc, const GhalfBf16* gbf, double* e_1b,
                    EnergySlots* es);

// The FP64 step workspace's slots (sized for W walkers; the VHS slot holds
// W M M complex128 values, enough for any precision's VHS).
struct StepSlots {
  double *gha, *ghb, *oo, *vb, *xs, *cfb, *cmf, *xi, *vhs, *inv;
};
int step_slots(afq_system* sys, int W, StepSlots* out);

// FP32-MFMA path (afq_fp32.hip)
int f32_prepare(afq_system* sys);
void f32_release(afq_system* sys);
int propagate_fp32(afq_system* sys, int W, double* phia, double* phib, double* weights,
                   double* e_hybrid, double eshift, const double* xi, uint64_t seed, int64_t step,
                   double* ovlp_new);

// reduced-precision path (afq_lowp.hip)
int system_prepare_lowp(afq_system* sys);
void system_release_lowp(afq_system* sys);
int propagate_lowp(afq_system* sys, int W, double* phia, double* phib, double* weights,
                   double* e_hybrid, double eshift, const double* xi, uint64_t seed, int64_t step,
                   double* ovlp_new);
int system_energy_lowp(afq_system* sys, int W, const double* phia, const double* phib,
                       const double* h1e, double enuc, double* e_tot, double* e_1b, double* e_2b);

}  // namespace afq
