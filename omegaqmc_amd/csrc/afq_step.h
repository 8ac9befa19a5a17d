// Step internals shared between the FP64 and reduced-precision paths.
#pragma once
#include "afq_common.h"

namespace afq {

int propagate_fp64(afq_system* sys, int W, double* phia, double* phib, double* weights,
                   double* e_hybrid, double eshift, const double* xi, uint64_t seed, int64_t step,
                   double* ovlp_new);
int system_energy_fp64(afq_system* sys, int W, const double* phia, const double* phib,
                       const double* h1e, double enuc, double* e_tot, double* e_1b, double* e_2b);

// reduced-precision path (afq_lowp.hip)
int system_prepare_lowp(afq_system* sys);
void system_release_lowp(afq_system* sys);
int propagate_lowp(afq_system* sys, int W, double* phia, double* phib, double* weights,
                   double* e_hybrid, double eshift, const double* xi, uint64_t seed, int64_t step,
                   double* ovlp_new);
int system_energy_lowp(afq_system* sys, int W, const double* phia, const double* phib,
                       const double* h1e, double enuc, double* e_tot, double* e_1b, double* e_2b);

}  // namespace afq
