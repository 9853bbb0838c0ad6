// kernels_pathq_refine.hip -- queued (persistent, many-pair) path-recurrence
// instantiations for the refine mode, in their own file so that they compile in
// parallel with the per-launch ones.
#include "path_impl.h"

namespace pyr {

hipError_t launchPathsQueuedRefine(const QueueArgs& q, int Dp, int adaptive, int exact,
                                   int narrow, int paired, int grid, hipStream_t s)
{
    return path::launchQueuedMode<kRefine>(q, Dp, adaptive, exact, narrow, paired, grid, s);
}

int queuedOccupancyRefine(int Dp, int adaptive, int exact, int narrow, int paired)
{
    return path::queuedOccupancy<kRefine>(Dp, adaptive, exact, narrow, paired);
}

} // namespace pyr
