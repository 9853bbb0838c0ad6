// kernels_post.hip -- left-right check, 3x3 median, confidence clear and the
// output conversions of one level, fused into one pass over the level.
//
// The right-view disparity of column x2 is decoded from the projection key
// written by the R->L kernel: key = ((cost + 32768) << 16) | (0xFFFF - x) of the
// winning left pixel x (smallest cost, ties to the largest x -- the order in
// which the reference's sequential R->L update keeps a value,
// sgm_kernels.cu:794-797 / :1462-1465), so disp2(x2) = x - x2.
//
// Restated semantics: pseudo LR check sgm_kernels.cu:825-851 (single scale),
// strict LR check :855-889 (pyramid), median :891-920 (in-bounds neighbours,
// sorted[n/2]), confidence clear :922-931, float conversion :1530-1551, output
// routing :1837-1858 / :2053-2069.
#include <string.h>

#include <type_traits>

#include "device_common.h"
#include "internal.h"
#include "stage_impl.h"

namespace pyr {
namespace {

// Block: 64 x 16 pixels, four rows per thread (row segments of 64 pixels give
// whole-line stores for every output).
constexpr int kPostW = 64, kPostH = 16, kPostR = kPostH / 4;

struct PostArgs {
    const int16_t* disp;
    const float* conf;
    const uint32_t* keys;
    int W, H, minD, maxDiff;
    int lrcheck, median, strict, pyramid;
    int16_t* out16;
    float* out32;
    float* outc;
    size_t s16, s32, sc;  // bytes
    Batch batch;
};

// The two right-view keys the LR check of disparity d at (x, y) reads (columns
// x - (d >> 4) and x - ((d + 15) >> 4)).  Loads only, so that a thread can have
// the keys of several pixels in flight at once: both keys are read from
// clamped in-image cells whatever d is, with no branch between the loads (a
// load under a branch is waited for where the branch joins).  lrDecide uses a
// key only when d is valid and its column is inside the image.
struct LrProbe {
    uint32_t ka, kb;
};

__device__ __forceinline__ LrProbe lrLoad(const PostArgs& a, int x, int y, int d)
{
    const int xa = x - (d >> 4), xb = x - ((d + 15) >> 4);
    const uint32_t* row = a.keys + (size_t)imin(imax(y, 0), a.H - 1) * a.W;
    return {row[imin(imax(xa, 0), a.W - 1)], row[imin(imax(xb, 0), a.W - 1)]};
}

__device__ __forceinline__ int disp2Of(uint32_t k, int x2, int invalid)
{
    if (k == 0xFFFFFFFFu) return invalid;
    const int xw = 0xFFFF - (int)(k & 0xFFFFu);
    return (int)(int16_t)(xw - x2);
}

// d after the LR check, from the keys lrLoad read.
__device__ __forceinline__ int lrDecide(const PostArgs& a, int x, int d, LrProbe pr, int invalid)
{
    if (!a.lrcheck || d == invalid) return d;
    const int d0 = d >> 4, d1 = (d + 15) >> 4;
    const int xa = x - d0, xb = x - d1;
    const bool ina = xa >= 0 && xa < a.W, inb = xb >= 0 && xb < a.W;
    const int ra = ina ? disp2Of(pr.ka, xa, invalid) : 0;
    const int rb = inb ? disp2Of(pr.kb, xb, invalid) : 0;
    if (a.strict) {
        const bool oka = ina && ra >= a.minD && abs(ra - d0) <= a.maxDiff;
        const bool okb = inb && rb >= a.minD && abs(rb - d1) <= a.maxDiff;
        return (oka || okb) ? d : invalid;
    }
    const bool bada = ina && ra >= a.minD && abs(ra - d0) > a.maxDiff;
    const bool badb = inb && rb >= a.minD && abs(rb - d1) > a.maxDiff;
    return (bada && badb) ? invalid : d;
}

__device__ __forceinline__ void sort2(int& a, int& b)
{
    const int lo = imin(a, b), hi = imax(a, b);
    a = lo;
    b = hi;
}

// Median of 9 with a 19-exchange network.  Out-of-image neighbours are replaced
// by 4 - n/2 copies of -inf and the rest by +inf, so that element 4 of the
// sorted nine equals element n/2 of the n in-image values.
__device__ __forceinline__ int median9(int p[9])
{
    sort2(p[1], p[2]); sort2(p[4], p[5]); sort2(p[7], p[8]);
    sort2(p[0], p[1]); sort2(p[3], p[4]); sort2(p[6], p[7]);
    sort2(p[1], p[2]); sort2(p[4], p[5]); sort2(p[7], p[8]);
    sort2(p[0], p[3]); sort2(p[5], p[8]); sort2(p[4], p[7]);
    sort2(p[3], p[6]); sort2(p[1], p[4]); sort2(p[2], p[5]);
    sort2(p[4], p[7]); sort2(p[4], p[2]); sort2(p[6], p[4]);
    sort2(p[4], p[2]);
    return p[4];
}

__device__ __forceinline__ void kPostBody(const PostArgs& args, Tile3 tl)
{
    constexpr int HW = kPostW + 2, HALO = HW * (kPostH + 2), NH = (HALO + 255) / 256;
    // This pair's buffers (pair = tl.z).
    PostArgs a = args;
    const size_t pr = tl.z;
    a.disp = pairAt(a.disp, pr, a.batch.ws);
    a.conf = pairAt(a.conf, pr, a.batch.ws);
    a.keys = pairAt(a.keys, pr, a.batch.ws);
    a.out16 = pairAt(a.out16, pr, a.batch.out16);
    a.out32 = pairAt(a.out32, pr, a.batch.out32);
    a.outc = pairAt(a.outc, pr, a.batch.outc);
    __shared__ int16_t s_d[kPostH + 2][HW];
    const int invalid = (int)(int16_t)((a.minD - 1) * 16);
    const int lx = threadIdx.x & 63, ly0 = threadIdx.x >> 6;
    const int bx = tl.x * kPostW, by = tl.y * kPostH;
    const int x = bx + lx;
    const int xc = imin(x, a.W - 1);
    // Own pixels: raw disparity and confidence.
    int raw[kPostR];
    float cf0[kPostR];
#pragma unroll
    for (int r = 0; r < kPostR; ++r) {
        const int yc = imin(by + ly0 + 4 * r, a.H - 1);
        raw[r] = a.disp[(size_t)yc * a.W + xc];
        cf0[r] = a.conf[(size_t)yc * a.W + xc];
    }
    int chk[kPostR];
    if (a.median) {
        // Checked disparities of the tile and its one-pixel ring (in-image cells only).
        int hd[NH];
        int hx[NH], hy[NH];
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            const int i = threadIdx.x + k * 256;
            const int ty = i / HW, tx = i - ty * HW;
            hx[k] = bx + tx - 1;
            hy[k] = by + ty - 1;
            const bool in = i < HALO && hx[k] >= 0 && hx[k] < a.W && hy[k] >= 0 && hy[k] < a.H;
            // Loaded from a clamped cell unconditionally (see lrLoad).
            const int v = a.disp[(size_t)imin(imax(hy[k], 0), a.H - 1) * a.W + imin(imax(hx[k], 0), a.W - 1)];
            hd[k] = in ? v : invalid;
            if (!in) hx[k] = -1;
        }
        LrProbe pb[NH] = {};
        if (a.lrcheck) {
#pragma unroll
            for (int k = 0; k < NH; ++k) pb[k] = lrLoad(a, hx[k], hy[k], hd[k]);
        }
        // Out-of-image cells hold `invalid` (d = invalid passes lrDecide
        // unchanged); the median reads in-image cells only.
        int ck[NH];
#pragma unroll
        for (int k = 0; k < NH; ++k) ck[k] = lrDecide(a, hx[k], hd[k], pb[k], invalid);
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            const int i = threadIdx.x + k * 256;
            const int ty = i / HW, tx = i - ty * HW;
            if (i < HALO) s_d[ty][tx] = (int16_t)ck[k];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kPostR; ++r) chk[r] = s_d[ly0 + 4 * r + 1][lx + 1];
    } else {
        LrProbe pb[kPostR] = {};
        if (a.lrcheck) {
#pragma unroll
            for (int r = 0; r < kPostR; ++r) pb[r] = lrLoad(a, xc, imin(by + ly0 + 4 * r, a.H - 1), raw[r]);
        }
#pragma unroll
        for (int r = 0; r < kPostR; ++r) chk[r] = lrDecide(a, xc, raw[r], pb[r], invalid);
    }
    if (x >= a.W) return;
    // IN (compile time): the block and its one-pixel ring lie inside the image,
    // so every pixel has nine in-image neighbours and no bound is tested.
    auto finish = [&](auto inC) {
    constexpr bool IN = decltype(inC)::value;
#pragma unroll
    for (int r = 0; r < kPostR; ++r) {
        const int ly = ly0 + 4 * r;
        const int y = by + ly;
        if (!IN && y >= a.H) break;
        float cf = (chk[r] == raw[r]) ? cf0[r] : 0.0f;
        int fin = chk[r];
        if (a.median) {
            int v[9];
            int n = 0;
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx)
                    n += IN || (x + dx >= 0 && x + dx < a.W && y + dy >= 0 && y + dy < a.H);
            const int nLow = 4 - n / 2;
            int seenOut = 0;
            int idx = 0;
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    const bool in = IN || (x + dx >= 0 && x + dx < a.W && y + dy >= 0 && y + dy < a.H);
                    int val;
                    if (in) {
                        val = s_d[ly + 1 + dy][lx + 1 + dx];
                    } else {
                        val = seenOut < nLow ? -100000 : 100000;
                        ++seenOut;
                    }
                    v[idx++] = val;
                }
            fin = median9(v);
            if (fin != chk[r]) cf = 0.0f;
        }
        if (a.out16) *reinterpret_cast<int16_t*>((char*)a.out16 + (size_t)y * a.s16 + (size_t)x * 2) = (int16_t)fin;
        if (a.out32 && a.outc) {
            const float d = (float)fin / 16.0f;
            const bool ok = fin != invalid && d > 0.0f;
            *reinterpret_cast<float*>((char*)a.out32 + (size_t)y * a.s32 + (size_t)x * 4) = ok ? d : 0.0f;
            *reinterpret_cast<float*>((char*)a.outc + (size_t)y * a.sc + (size_t)x * 4) = ok ? cf : 0.0f;
        } else if (a.outc && a.pyramid) {
            *reinterpret_cast<float*>((char*)a.outc + (size_t)y * a.sc + (size_t)x * 4) = cf;
        }
    }
    };
    if (bx >= 1 && bx + kPostW <= a.W - 1 && by >= 1 && by + kPostH <= a.H - 1)
        finish(std::true_type{});
    else
        finish(std::false_type{});
}

__device__ __forceinline__ PostArgs postArgs(const StageArgs& s)
{
    PostArgs a;
    a.disp = s.disp;
    a.conf = s.conf;
    a.keys = s.keys;
    a.W = s.W;
    a.H = s.H;
    a.minD = s.minD;
    a.maxDiff = s.maxDiff;
    a.lrcheck = s.lrcheck;
    a.median = s.median;
    a.strict = s.strict;
    a.pyramid = s.strict;
    a.out16 = s.out16;
    a.out32 = s.out32;
    a.outc = s.outc;
    a.s16 = s.s16;
    a.s32 = s.s32;
    a.sc = s.sc;
    a.batch = s.batch;
    return a;
}

__global__ __launch_bounds__(256) void kPost(StageArgs a) { kPostBody(postArgs(a), xcdTile()); }

__global__ __launch_bounds__(256) void kPostQ(StageQueue q)
{
    runQueuedStage(q, [](const StageArgs& a, Tile3 tl) { kPostBody(postArgs(a), tl); });
}

} // namespace

hipError_t launchPostStage(const StageLaunch& l, hipStream_t s)
{
    kPost<<<l.grid, 256, 0, s>>>(l.a);
    return hipGetLastError();
}

hipError_t launchPostQueued(const StageQueue& q, int blocks, hipStream_t s)
{
    kPostQ<<<blocks, 256, 0, s>>>(q);
    return hipGetLastError();
}

hipError_t launchPost(const LevelCall& c, Workspace& ws, hipStream_t s)
{
    StageLaunch l;
    memset(&l, 0, sizeof(l));
    l.kind = kStPost;
    l.grid = dim3((c.W + kPostW - 1) / kPostW, (c.H + kPostH - 1) / kPostH, c.batch.n);
    StageArgs& a = l.a;
    a.disp = ws.disp;
    a.conf = ws.conf;
    a.keys = ws.keys;
    a.W = c.W;
    a.H = c.H;
    a.minD = c.minD;
    a.maxDiff = c.disp12;
    a.lrcheck = (c.post & 1) != 0;
    a.median = (c.post & 2) != 0;
    a.strict = c.mode != kSingle;
    a.out16 = c.out16;
    a.out32 = c.out32;
    a.outc = c.outc;
    a.s16 = c.s16;
    a.s32 = c.s32;
    a.sc = c.sc;
    a.batch = c.batch;
    // A recorded launch is kept even without outputs: its queued form takes
    // the outputs of each call from the pair's I/O entry.
    if (c.stageSink) {
        c.stageSink->launches.push_back(l);
        return hipSuccess;
    }
    if (!a.out16 && !a.outc) return hipSuccess;
    return launchStage(l, s);
}

} // namespace pyr
