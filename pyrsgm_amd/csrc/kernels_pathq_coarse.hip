// kernels_pathq_coarse.hip -- queued (persistent, many-pair) path-recurrence
// instantiations for the coarse mode, in their own file so that they compile in
// parallel with the per-launch ones.
#include "path_impl.h"

namespace pyr {

hipError_t launchPathsQueuedCoarse(const QueueArgs& q, int Dp, int adaptive, int exact,
                                   int narrow, int paired, int grid, hipStream_t s)
{
    return path::launchQueuedMode<kCoarse>(q, Dp, adaptive, exact, narrow, paired, grid, s);
}

int queuedOccupancyCoarse(int Dp, int adaptive, int exact, int narrow, int paired)
{
    return path::queuedOccupancy<kCoarse>(Dp, adaptive, exact, narrow, paired);
}

} // namespace pyr
