// Reduced-precision step (FP32 MFMA / split-BF16 MFMA contractions).
#include "afq_kernels.h"
#include "afq_step.h"

namespace afq {

int system_prepare_lowp(afq_system* sys) {
  if (sys->precision == AFQ_FP64) return AFQ_OK;
  return fail(AFQ_EUNSUPPORTED, "precision %d is not built yet", sys->precision);
}

void system_release_lowp(afq_system* sys) {
  if (sys->chol_lp) (void)hipFree(sys->chol_lp);
  if (sys->rchol_lp) (void)hipFree(sys->rchol_lp);
  sys->chol_lp = sys->rchol_lp = nullptr;
}

int propagate_lowp(afq_system*, int, double*, double*, double*, double*, double, const double*,
                   uint64_t, int64_t, double*) {
  return fail(AFQ_EUNSUPPORTED, "reduced-precision step is not built yet");
}

int system_energy_lowp(afq_system*, int, const double*, const double*, const double*, double,
                       double*, double*, double*) {
  return fail(AFQ_EUNSUPPORTED, "reduced-precision energy is not built yet");
}

}  // namespace afq
