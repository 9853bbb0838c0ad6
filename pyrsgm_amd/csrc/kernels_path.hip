// kernels_path.hip -- the path-recurrence kernels of one (mode, adaptive P2,
// volume form) specialisation, registered in the launcher table at load time.
//
// The build compiles this file once per specialisation so that the roughly
// 170 kernel instantiations compile in parallel:
//   PYRSGM_PATH_MODE    0 single scale, 1 coarse level, 2 refinement level
//   PYRSGM_PATH_ADAPT   0 fixed P2, 1 adaptive P2
//   PYRSGM_PATH_FORM    0 exact (pyramid levels whose path costs may reach
//                       SHRT_MAX), 1 byte volumes, 2 int16 volumes
//   PYRSGM_PATH_QUEUED  0 per-launch kernels (kPaths), 1 queued kernels (kPathsQ)
#include "path_impl.h"

#if !defined(PYRSGM_PATH_MODE) || !defined(PYRSGM_PATH_ADAPT) || !defined(PYRSGM_PATH_FORM) || \
    !defined(PYRSGM_PATH_QUEUED)
#error "kernels_path.hip is compiled with PYRSGM_PATH_MODE/ADAPT/FORM/QUEUED defined"
#endif

namespace pyr {
namespace {

constexpr int kMode = PYRSGM_PATH_MODE;
constexpr bool kAdapt = PYRSGM_PATH_ADAPT != 0;
constexpr bool kExact = PYRSGM_PATH_FORM == kFormExact;
constexpr bool kNarrow = PYRSGM_PATH_FORM == kFormNarrow;
static_assert(kMode != kSingle || !kExact, "single-scale paths have no exact form");

#if PYRSGM_PATH_QUEUED
const void* queuedKernelPart(int Dp, int paired)
{
    return path::queuedKernel<kMode, kAdapt, kExact, kNarrow>(Dp, paired);
}
#else
hipError_t launchPart(const PathArgs& a, hipStream_t s)
{
    return path::dispatchPaths<kMode, kAdapt, kExact, kNarrow>(a, s);
}
#endif

struct Register {
    Register()
    {
        PathPart& p = pathPart(kMode, kAdapt ? 1 : 0, PYRSGM_PATH_FORM);
#if PYRSGM_PATH_QUEUED
        p.queued = &queuedKernelPart;
#else
        p.launch = &launchPart;
#endif
    }
} registerPart;

} // namespace
} // namespace pyr
