// stage_impl.h -- the queued form of the stage kernels (pyrDown, descriptor
// pass, cost volume, post): one launch runs every task of one kernel
// specialisation, e.g. the same stage of all pairs of a mixed batch at one
// pyramid step.
#pragma once

#include "device_common.h"
#include "internal.h"

namespace pyr {

// Tile (x, y, z) in the hardware's block order.
__device__ __forceinline__ Tile3 hwTile()
{
    return Tile3{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
}

// Runs body(args, tile) for this block's task of a queued launch.  Blocks are
// handed to the XCDs in contiguous ranges (neighbouring tiles of one task share
// an L2); block b belongs to the last task t with start[t] <= b and runs tile
// local = b - start[t] of it.  Per-call fields (the caller's images and level-0
// outputs) come from the pair's I/O entry.
template <class Body>
__device__ __forceinline__ void runQueuedStage(const StageQueue& q, Body body)
{
    const int b = xcdRemap((int)blockIdx.x, (int)gridDim.x);
    int lo = 0, hi = q.nTasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (q.start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const int t = __builtin_amdgcn_readfirstlane(lo);
    const StageTask& task = q.tasks[t];
    const int local = b - q.start[t];
    const int gx = task.gx, gy = task.gy;
    Tile3 tl;
    tl.x = local % gx;
    const int rest = local / gx;
    tl.y = rest % gy;
    tl.z = rest / gy;
    StageArgs a = task.a;
    if (task.io & kIoIn) {
        const StageIo& io = q.io[task.pair];
        a.left = io.left;
        a.right = io.right;
        a.lstep = (int)io.lstep;
        a.rstep = (int)io.rstep;
    }
    if (task.io & kIoOut) {
        const StageIo& io = q.io[task.pair];
        a.out16 = io.out16;
        a.out32 = io.out32;
        a.outc = io.outc;
        a.s16 = io.s16;
        a.s32 = io.s32;
        a.sc = io.sc;
    }
    body(a, tl);
}

} // namespace pyr
