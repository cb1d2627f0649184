// Fused FP32 engine entry (filled in by the fused kernels).
#include "plan.h"

namespace ib {

int collide_fused(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                  int with_loss, double hb, hipStream_t st, int* handled) {
  (void)P; (void)f; (void)q; (void)batch; (void)pidx; (void)with_loss; (void)hb; (void)st;
  *handled = 0;
  return IB_OK;
}

}  // namespace ib
