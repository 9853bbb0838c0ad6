// kernels_path_refine.hip -- path-recurrence instantiations for the refine mode
// (split per mode so the instantiations compile in parallel).
#include "path_impl.h"

namespace pyr {

hipError_t launchPathsRefine(const PathArgs& a, int adaptive, hipStream_t s)
{
    return path::launchPathsMode<kRefine>(a, adaptive, s);
}

} // namespace pyr
