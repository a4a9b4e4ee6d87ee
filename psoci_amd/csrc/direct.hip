// Direct sigma and the adjoint (RDM) kernels: elements regenerated on the fly.
#include "ctx.hpp"

namespace psoci {

void sigma_direct(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy) {
  (void)c; (void)prec; (void)k; (void)X; (void)ldx; (void)Y; (void)ldy;
  throw Error(PSOCI_ERR_UNSUPPORTED, "direct sigma is not available in this build");
}

void rdm_grad(Ctx& c, int k, const double* dC, const double* w, double* d_g1, double* d_g2) {
  (void)c; (void)k; (void)dC; (void)w; (void)d_g1; (void)d_g2;
  throw Error(PSOCI_ERR_UNSUPPORTED, "the adjoint is not available in this build");
}

}  // namespace psoci
