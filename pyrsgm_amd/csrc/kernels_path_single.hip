// kernels_path_single.hip -- path-recurrence instantiations for the single mode
// (split per mode so the instantiations compile in parallel).
#include "path_impl.h"

namespace pyr {

hipError_t launchPathsSingle(const PathArgs& a, int adaptive, hipStream_t s)
{
    return path::launchPathsMode<kSingle>(a, adaptive, s);
}

} // namespace pyr
