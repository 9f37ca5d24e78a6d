// Helmholtz FMM driver (helmholtz/hfmm2d.f).  Placeholder until the
// Helmholtz translation operators land: reports "unsupported".
#include "helm_ops.h"

namespace fmm2d {

int helmholtz_fmm(const HelmArgs& a, hipStream_t st) {
  (void)a;
  (void)st;
  return FMM2D_ERR_UNSUPPORTED;
}

}  // namespace fmm2d
