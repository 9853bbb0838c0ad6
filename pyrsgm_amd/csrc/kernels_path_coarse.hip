// kernels_path_coarse.hip -- path-recurrence instantiations for the coarse mode
// (split per mode so the instantiations compile in parallel).
#include "path_impl.h"

namespace pyr {

hipError_t launchPathsCoarse(const PathArgs& a, int adaptive, hipStream_t s)
{
    return path::launchPathsMode<kCoarse>(a, adaptive, s);
}

} // namespace pyr
