// kernels_io.hip -- batched 2-D copies between caller buffers and a serving
// lane's stacked staging buffers (one launch for any number of copies), and
// the per-call image patch of a queued path launch's task table.
#include "device_common.h"
#include "internal.h"

namespace pyr {
namespace {

// Block: 256 threads over 1024 bytes of one row of copy blockIdx.z.  Rows are
// copied in 4-byte words when both ends, both pitches and the row length are
// 4-byte aligned, else byte by byte.
__global__ __launch_bounds__(256) void kCopy2D(const CopyDesc* descs)
{
    const CopyDesc d = descs[blockIdx.z];
    const int y = blockIdx.y;
    const int x = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (y >= d.rows || x >= d.rowBytes) return;
    const uint8_t* src = d.src + (size_t)y * d.spitch;
    uint8_t* dst = d.dst + (size_t)y * d.dpitch;
    const bool words = (((uintptr_t)d.src | (uintptr_t)d.dst | d.spitch | d.dpitch |
                         (size_t)d.rowBytes) & 3) == 0;
    if (words) {
        *reinterpret_cast<uint32_t*>(dst + x) = *reinterpret_cast<const uint32_t*>(src + x);
    } else {
        const int end = imin(x + 4, d.rowBytes);
        for (int i = x; i < end; ++i) dst[i] = src[i];
    }
}

// Thread i < p.n points task p.idx[i] at its call's left image.
__global__ __launch_bounds__(64) void kImgPatch(ImgPatch p)
{
    const int i = threadIdx.x;
    if (i >= p.n) return;
    PathArgs* t = p.tasks + p.idx[i];
    t->img = p.img[i];
    t->istep = p.istep[i];
}

__global__ __launch_bounds__(64) void kIoPatch(IoPatch p)
{
    const int i = threadIdx.x;
    if (i >= p.n) return;
    p.table[p.idx[i]] = p.io[i];
}

} // namespace

hipError_t launchIoPatch(const IoPatch& p, hipStream_t s)
{
    if (p.n <= 0) return hipSuccess;
    if (p.n > kMaxIoPatch) return hipErrorInvalidValue;
    kIoPatch<<<1, 64, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t launchImgPatch(const ImgPatch& p, hipStream_t s)
{
    if (p.n <= 0) return hipSuccess;
    if (p.n > kMaxImgPatches) return hipErrorInvalidValue;
    kImgPatch<<<1, 64, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t launchCopy2D(const CopyDesc* descs, int count, int maxRowBytes, int maxRows,
                        hipStream_t stream)
{
    if (count <= 0) return hipSuccess;
    dim3 grid((unsigned)((maxRowBytes + 1023) / 1024), (unsigned)maxRows, (unsigned)count);
    kCopy2D<<<grid, 256, 0, stream>>>(descs);
    return hipGetLastError();
}

} // namespace pyr
