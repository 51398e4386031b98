// PatchMatch repair (placeholder until the kernel lands).
#include "om_common.h"

using namespace om;

extern "C" size_t om_pm_workspace_bytes(int32_t H, int32_t W) { return (size_t)(H > 0 ? H : 0) * (W > 0 ? W : 0) * 8; }

extern "C" int om_pm_refine(const float *, int32_t, int32_t, const float *, const double *, int32_t,
                            const float *const *, const float *const *, const int32_t *, const int32_t *,
                            const double *, float *, float *, float *, const uint8_t *, int32_t, int32_t, int32_t,
                            double, double, int32_t, double, double, uint32_t, void *, size_t, int64_t *, void *)
{
    return fail(OM_EUNSUP, "pm_refine not built yet");
}
