// Interstitial (off-lattice atom) passes.  Placeholder: the list can be set
// only empty until the device passes are in place.
#include "ctx.h"
#include "host_helpers.h"

namespace mmd {
int itl_rho_df(mmd_ctx *, double, bool) { return fail(MMD_E_STATE, "interstitial passes not available"); }
int itl_df(mmd_ctx *) { return fail(MMD_E_STATE, "interstitial passes not available"); }
int itl_force(mmd_ctx *, double, bool) { return fail(MMD_E_STATE, "interstitial passes not available"); }
int itl_kick(mmd_ctx *, double, bool, bool) { return fail(MMD_E_STATE, "interstitial passes not available"); }
void itl_free(mmd_ctx *) {}
}  // namespace mmd

extern "C" int mmd_itl_set(mmd_ctx *c, int32_t n, const double *, const double *, const double *, const double *,
                           const double *, const double *, const int8_t *, const int32_t *, int32_t, int32_t) {
  if (!c) return mmd::fail(MMD_E_INVALID, "null context");
  if (n != 0) return mmd::fail(MMD_E_STATE, "mmd_itl_set: interstitial passes not available");
  c->n_itl = 0;
  return MMD_OK;
}

extern "C" int mmd_itl_get(mmd_ctx *c, int32_t *n, double *, double *, double *, double *, double *, double *,
                           double *, double *, double *, double *, double *, int8_t *) {
  if (!c) return mmd::fail(MMD_E_INVALID, "null context");
  if (n) *n = c->n_itl;
  return MMD_OK;
}
