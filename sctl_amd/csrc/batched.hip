// Batched near-field blocks (plan API).  Implementation follows.
#include "internal.hpp"

extern "C" {
int sctl_amd_plan_create(sctl_amd_plan*, int, int, int, int64_t, const int64_t*, const int64_t*, const int64_t*,
                         const int64_t*, int64_t, int64_t) {
  sctl_amd::set_error("plan API not built yet");
  return SCTL_AMD_ERROR_UNSUPPORTED;
}
int sctl_amd_plan_set_geometry(sctl_amd_plan, const double*, const double*, const double*, uint32_t, void*) {
  return SCTL_AMD_ERROR_UNSUPPORTED;
}
int sctl_amd_plan_apply(sctl_amd_plan, const double*, double*, uint32_t, void*) { return SCTL_AMD_ERROR_UNSUPPORTED; }
int sctl_amd_plan_interactions(sctl_amd_plan, int64_t*) { return SCTL_AMD_ERROR_UNSUPPORTED; }
int sctl_amd_plan_destroy(sctl_amd_plan) { return SCTL_AMD_SUCCESS; }
}
