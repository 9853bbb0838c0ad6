// kernels_pathq_single.hip -- queued (persistent, many-pair) path-recurrence
// instantiations for the single mode, in their own file so that they compile in
// parallel with the per-launch ones.
#include "path_impl.h"

namespace pyr {

hipError_t launchPathsQueuedSingle(const QueueArgs& q, int Dp, int adaptive, int narrow, int grid,
                                   hipStream_t s)
{
    return path::launchQueuedMode<kSingle>(q, Dp, adaptive, 0, narrow, 0, grid, s);
}

int queuedOccupancySingle(int Dp, int adaptive, int exact, int narrow, int paired)
{
    return path::queuedOccupancy<kSingle>(Dp, adaptive, exact, narrow, paired);
}

} // namespace pyr
