// noalloc.hip -- pyrsgm_compute_single_scale_noalloc: the single-scale match
// of pyrsgm_compute_single_scale with its workspace in a fixed device array
// instead of memory the library allocates, for callers that must not call
// hipMalloc (code under stream capture, serving loops, harnesses that forbid
// allocation).
//
// The array `scratch` holds one single-scale workspace (costs, three path
// volumes, descriptors, per-pixel maps) of at most PYRSGM_NOALLOC_SCRATCH_BYTES.
// Each device the library runs on holds its own copy for the life of the
// process.  Its device address is looked up once per device with
// hipGetSymbolAddress (in host code the array's name is not a device
// address), and the kernels receive that address as the workspace base, as
// they receive a library-allocated workspace from pyrsgm_compute_single_scale.
//
// Nothing in the workspace has to start at zero: every buffer a call reads is
// written earlier in the same call (the projection keys are reset by the
// descriptor pass), so a call does not depend on what a previous call left.
// Calls that overlap on one device share the array and race; callers
// serialise them (one stream, or events between streams).
#include <stdio.h>

#include <mutex>

#include "internal.h"
#include "pyrsgm/pyrsgm_c.h"

namespace {

static __device__ __align__(256) unsigned char scratch[PYRSGM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device, once looked up

// The current device's address of `scratch`; null (with a message) on failure.
void* scratchOfCurrentDevice()
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        fprintf(stderr, "pyrsgm: pyrsgm_compute_single_scale_noalloc: no usable device (%s)\n",
                e != hipSuccess ? hipGetErrorString(e) : "device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "pyrsgm: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

} // namespace

extern "C" int pyrsgm_compute_single_scale_noalloc(
    const unsigned char* d_left_gray, const unsigned char* d_right_gray, short* d_disp,
    float* d_disp_float, float* d_conf_float, size_t left_step_bytes, size_t right_step_bytes,
    size_t disp_step_bytes, size_t disp_float_step_bytes, size_t conf_float_step_bytes,
    int width, int height, int numDisparities, int minDisparity, int blockSize, int P1, int P2,
    int uniquenessRatio, int disp12MaxDiff, int preFilterCap, int costMode, int postFlags,
    int numPaths, int adaptiveP2, float p2Alpha, float p2Beta, float p2LowerMult,
    float p2UpperMult, hipStream_t stream)
{
    if (width <= 0 || height <= 0 || numDisparities < 16 || numDisparities > 1024 ||
        numDisparities % 16 != 0 || width > pyr::kMaxWidth)
        return -1;
    const size_t need = pyr::workspaceBytes(width, height, numDisparities);
    if (need > (size_t)PYRSGM_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr,
                "pyrsgm: pyrsgm_compute_single_scale_noalloc: %dx%d with %d disparities needs "
                "%zu bytes of scratch, above the fixed %zu\n",
                width, height, numDisparities, need, (size_t)PYRSGM_NOALLOC_SCRATCH_BYTES);
        return -1;
    }
    void* base = scratchOfCurrentDevice();
    if (!base) return -1;
    pyr::Workspace ws;
    if (pyr::borrowWorkspace(&ws, base, PYRSGM_NOALLOC_SCRATCH_BYTES, width, height,
                             numDisparities) != 0)
        return -1;
    return pyr::computeSingleScale(&ws, d_left_gray, d_right_gray, d_disp, d_disp_float,
                                   d_conf_float, left_step_bytes, right_step_bytes,
                                   disp_step_bytes, disp_float_step_bytes, conf_float_step_bytes,
                                   width, height, numDisparities, minDisparity, blockSize, P1, P2,
                                   uniquenessRatio, disp12MaxDiff, preFilterCap, costMode,
                                   postFlags, numPaths, adaptiveP2, p2Alpha, p2Beta, p2LowerMult,
                                   p2UpperMult, pyr::Batch{}, stream);
}
