/* placeholder, replaced below */
#include "../../../include/feyntrop_amd.h"
extern "C" int ftr_integrate_batch(ftr_ctx* ctx, int n, ftr_problem* const* probs, const uint64_t* N,
                                   const uint64_t* seeds, ftr_moments* out) {
  for (int i = 0; i < n; i++) {
    int rc = ftr_integrate(ctx, probs[i], N[i], seeds[i], nullptr, &out[i]);
    if (rc) return rc;
  }
  return FTR_OK;
}
