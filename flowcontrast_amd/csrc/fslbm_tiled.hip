// Tiled z-marching FSLBM step kernel (placeholder until written).
#include "fslbm_internal.h"

namespace fslbm {

bool tiled_supported(const Geo &, int, int) { return false; }

hipError_t launch_step_tiled(const Geo &, int, const void *, void *,
                             const uint8_t *, const StepArgs &, hipStream_t)
{
    return hipErrorNotSupported;
}

}  // namespace fslbm
