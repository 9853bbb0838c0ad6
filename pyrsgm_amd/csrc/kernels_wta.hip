// kernels_wta.hip -- path sums + winner-take-all for one level.
//
// Q = Dp/16 lanes per pixel, 16 labels per lane, pixels consecutive across
// lanes, so every volume read is a run of full 16-byte loads.  The per-pixel
// search and rules (the reference's sums, argmin, uniqueness, sub-pixel
// parabola, confidence, validity and right-view projection) are in
// wta_impl.h, which the fused R->L pass shares.
#include "wta_impl.h"

namespace pyr {
namespace {

// One block of pixels (block bx of the level, pair pr).
template <int Q, int MODE, bool NARROW>
__device__ __forceinline__ void wtaBlock(const PathArgs& a, size_t pr, size_t bx)
{
    constexpr bool PYR = MODE != kSingle;
    const int W = a.W, D = a.D;
    const size_t P = (size_t)W * a.H;
    const size_t wsS = a.batch.ws;
    const int16_t* const Arl = pairAt(a.Arl, pr, wsS);
    const int16_t* const Alr = pairAt(a.Alr, pr, wsS);
    const int16_t* const Atb = pairAt(a.Atb, pr, wsS);
    const int32_t* const labels = pairAt(a.labels, pr, wsS);
    const size_t t = bx * 256 + threadIdx.x;
    const size_t p = t / Q;
    const int sub = (int)(t - p * Q);
    const bool ok = p < P;
    const size_t pc = ok ? p : P - 1;
    const size_t off = pc * a.Dp + (size_t)sub * 16;
    const int np = a.numPaths;
    // The pixel's label word is loaded with the volumes (pc is in bounds), so
    // that the wave waits on memory once, not again after the label search.
    const int lwPre = MODE == kRefine ? wta::ldGlobal(labels + pc) : 0;

    const int d0 = sub * 16;
    wta::Pick k;
    if constexpr (NARROW) {
        typedef uint32_t v4 __attribute__((ext_vector_type(4)));
        auto vec = [&](const int16_t* v) {
            return wta::ldGlobal(reinterpret_cast<const v4*>(reinterpret_cast<const uint8_t*>(v) + off));
        };
        const v4 none = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        const v4 vr = vec(Arl);
        const v4 vl = np >= 2 ? vec(Alr) : none;
        const v4 vt = np >= 3 ? vec(Atb) : none;
        const uint32_t wr[4] = {vr[0], vr[1], vr[2], vr[3]};
        const uint32_t wl[4] = {vl[0], vl[1], vl[2], vl[3]};
        const uint32_t wt[4] = {vt[0], vt[1], vt[2], vt[3]};
        k = wta::searchNarrow<Q>(wr, wl, wt, d0, D);
    } else {
        k = wta::searchWide<Q, PYR>(Arl, Alr, Atb, off, np, d0, D);
    }
    if (sub != 0 || !ok) return;

    int y, x;
    if (P <= 0xFFFFFFFFull) {
        y = (int)wta::rowOf((uint32_t)p, a.invW);
        x = (int)((uint32_t)p - (uint32_t)y * (uint32_t)W);
    } else {
        y = (int)(p / W);
        x = (int)(p - (size_t)y * W);
    }
    wta::resolve<MODE>(a, k, x, y, p, lwPre, pairAt(a.disp, pr, wsS), pairAt(a.conf, pr, wsS),
                       pairAt(a.keys, pr, wsS));
}

template <int Q, int MODE, bool NARROW>
__global__ __launch_bounds__(256) void kWTA(PathArgs a)
{
    wtaBlock<Q, MODE, NARROW>(a, blockIdx.y, blockIdx.x);
}

// Many pairs' levels of one specialisation in one launch: block b belongs to
// task t with start[t] <= b < start[t + 1] (the tasks' own block counts).
template <int Q, int MODE, bool NARROW>
__global__ __launch_bounds__(256) void kWTAQ(WtaQueueArgs q)
{
    const int b = blockIdx.x;
    int lo = 0, hi = q.nTasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (q.start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    wtaBlock<Q, MODE, NARROW>(q.tasks[lo], 0, (size_t)(b - q.start[lo]));
}

template <int MODE, bool NARROW>
hipError_t dispatchWTAQ(const WtaQueueArgs& q, int Dp, hipStream_t s)
{
    const dim3 grid((unsigned)q.total);
    if constexpr (NARROW) {
        switch (Dp / 16) {
        case 1: kWTAQ<1, MODE, true><<<grid, 256, 0, s>>>(q); break;
        case 2: kWTAQ<2, MODE, true><<<grid, 256, 0, s>>>(q); break;
        case 4: kWTAQ<4, MODE, true><<<grid, 256, 0, s>>>(q); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (Dp / 16) {
    case 1: kWTAQ<1, MODE, false><<<grid, 256, 0, s>>>(q); break;
    case 2: kWTAQ<2, MODE, false><<<grid, 256, 0, s>>>(q); break;
    case 4: kWTAQ<4, MODE, false><<<grid, 256, 0, s>>>(q); break;
    case 8: kWTAQ<8, MODE, false><<<grid, 256, 0, s>>>(q); break;
    case 16: kWTAQ<16, MODE, false><<<grid, 256, 0, s>>>(q); break;
    case 32: kWTAQ<32, MODE, false><<<grid, 256, 0, s>>>(q); break;
    case 64: kWTAQ<64, MODE, false><<<grid, 256, 0, s>>>(q); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE, bool NARROW>
hipError_t dispatchWTA(const PathArgs& a, hipStream_t s)
{
    const int Q = a.Dp / 16;
    const size_t threads = (size_t)a.W * a.H * Q;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    const dim3 grid(blocks, a.batch.n);
    if constexpr (NARROW) {  // at most 64 labels
        switch (Q) {
        case 1: kWTA<1, MODE, true><<<grid, 256, 0, s>>>(a); break;
        case 2: kWTA<2, MODE, true><<<grid, 256, 0, s>>>(a); break;
        case 4: kWTA<4, MODE, true><<<grid, 256, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (Q) {
    case 1: kWTA<1, MODE, false><<<grid, 256, 0, s>>>(a); break;
    case 2: kWTA<2, MODE, false><<<grid, 256, 0, s>>>(a); break;
    case 4: kWTA<4, MODE, false><<<grid, 256, 0, s>>>(a); break;
    case 8: kWTA<8, MODE, false><<<grid, 256, 0, s>>>(a); break;
    case 16: kWTA<16, MODE, false><<<grid, 256, 0, s>>>(a); break;
    case 32: kWTA<32, MODE, false><<<grid, 256, 0, s>>>(a); break;
    case 64: kWTA<64, MODE, false><<<grid, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t dispatchWTAMode(const PathArgs& a, hipStream_t s)
{
    return a.narrow ? dispatchWTA<MODE, true>(a, s) : dispatchWTA<MODE, false>(a, s);
}

} // namespace

int wtaBlocks(const PathArgs& a)
{
    return (int)(((size_t)a.W * a.H * (size_t)(a.Dp / 16) + 255) / 256);
}

hipError_t launchWTAQueued(const WtaQueueArgs& q, int mode, int Dp, int narrow, hipStream_t s)
{
    if (q.total <= 0) return hipSuccess;
    if (mode == kRefine) return narrow ? dispatchWTAQ<kRefine, true>(q, Dp, s) : dispatchWTAQ<kRefine, false>(q, Dp, s);
    if (mode == kCoarse) return narrow ? dispatchWTAQ<kCoarse, true>(q, Dp, s) : dispatchWTAQ<kCoarse, false>(q, Dp, s);
    return narrow ? dispatchWTAQ<kSingle, true>(q, Dp, s) : dispatchWTAQ<kSingle, false>(q, Dp, s);
}

hipError_t launchWTA(const PathArgs& a, int mode, hipStream_t s)
{
    if (mode == kRefine) return dispatchWTAMode<kRefine>(a, s);
    if (mode == kCoarse) return dispatchWTAMode<kCoarse>(a, s);
    return dispatchWTAMode<kSingle>(a, s);
}

} // namespace pyr
