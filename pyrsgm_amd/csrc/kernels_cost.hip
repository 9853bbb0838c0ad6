// kernels_cost.hip -- per-pixel descriptors, matching-cost volumes and pyrDown.
//
// Descriptors (kPrepCensus / kPrepBT), per pixel and view.  Census: one 32-bit
// 5x5 census word (sgm_kernels.cu:251-298).  BT: the reference's per-channel BT interval
// (u, min, max of the two half-pixel neighbours, sgm_kernels.cu:1002-1028) is
// a property of one pixel, so it is computed once per pixel and packed as
// x = gradient channel, y = intensity channel, each (v | v0<<8 | v1<<16).
// A BT cost is then two interval distances (sgm_kernels.cu:1025-1027, :1042).
//
// Volumes are int16 [y][x][Dp], labels innermost, stored as SHRT_MAX - cost;
// every thread produces 8 consecutive labels of one pixel and writes them with
// one 16-byte store.
#include <string.h>

#include <climits>
#include <type_traits>

#include "device_common.h"
#include "internal.h"
#include "stage_impl.h"

namespace pyr {
namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return imin(imax(v, lo), hi); }

// ---- BT helpers ----------------------------------------------------------
__device__ __forceinline__ uint32_t btInterval(int v, int l, int r)
{
    const int hl = (v + l) / 2, hr = (v + r) / 2;
    const int v0 = imin(imin(hl, hr), v), v1 = imax(imax(hl, hr), v);
    return (uint32_t)v | ((uint32_t)v0 << 8) | ((uint32_t)v1 << 16);
}

__device__ __forceinline__ int btChannel(uint32_t a, uint32_t b)
{
    const int u = a & 0xFF, u0 = (a >> 8) & 0xFF, u1 = (a >> 16) & 0xFF;
    const int v = b & 0xFF, v0 = (b >> 8) & 0xFF, v1 = (b >> 16) & 0xFF;
    const int c0 = imax(imax(0, u - v1), v0 - u);
    const int c1 = imax(imax(0, v - u1), u0 - v);
    return imin(c0, c1);
}

__device__ __forceinline__ int btCost(uint2 l, uint2 r)
{
    return imin(255, btChannel(l.x, r.x) + (btChannel(l.y, r.y) >> 2));
}

// sgm_kernels.cu:94-118
__device__ __forceinline__ int btGrad(const uint8_t* img, int step, int W, int H, int x, int y,
                                      int ftzero)
{
    const int xp = imax(x - 1, 0), xn = imin(x + 1, W - 1);
    const int yp = imax(y - 1, 0), yn = imin(y + 1, H - 1);
    const uint8_t* r = img + (size_t)y * step;
    const uint8_t* u = img + (size_t)yp * step;
    const uint8_t* d = img + (size_t)yn * step;
    int g = ((int)r[xn] - (int)r[xp]) * 2 + ((int)u[xn] - (int)u[xp]) + ((int)d[xn] - (int)d[xp]);
    g = clampi(g, -ftzero, ftzero);
    // Stored as a byte (sgm_kernels.cu:117): wraps when ftzero exceeds 127.
    return (int)(uint8_t)(g + ftzero);
}

// ---- descriptors + projection-key reset ----------------------------------
// Census (sgm_kernels.cu:251-298): one 32-bit word per pixel and view, view 0
// (left) then view 1 (right).  Block: 64 x 32 pixels, four horizontally
// adjacent pixels in two rows per thread; the 68 x 36 window (replicate-clamped) is loaded
// with every global load in flight at once and staged in LDS.  A thread reads
// each window row of its four pixels as two dwords; each census bit is the
// sign of (neighbour - centre), shifted in with one v_alignbit.
constexpr int kCenRows = 32;

__device__ __forceinline__ int byteOf(uint32_t lo, uint32_t hi, int b)
{
    return (int)(((b < 4 ? lo : hi) >> (8 * (b & 3))) & 0xFFu);
}

__device__ __forceinline__ void kPrepCensusBody(Tile3 tl, const uint8_t* left, int lstep,
                                                const uint8_t* right, int rstep, uint32_t* cen,
                                                uint32_t* keys, int W, int H, Batch b)
{
    constexpr int TW = 68, TS = 72, TH = kCenRows + 4, TILE = TW * TH, NL = (TILE + 255) / 256;
    __shared__ __attribute__((aligned(16))) uint8_t s_t[TH][TS];
    const int z = tl.z & 1;          // view
    const size_t pr = tl.z >> 1;     // pair
    const uint8_t* img = z ? pairAt(right, pr, b.right) : pairAt(left, pr, b.left);
    const int step = z ? rstep : lstep;
    cen = pairAt(cen, pr, b.ws);
    keys = pairAt(keys, pr, b.ws);
    const int x0 = tl.x * 64, y0 = tl.y * kCenRows;
    uint32_t v[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int i = threadIdx.x + k * 256;
        const int ty = i / TW, tx = i - ty * TW;
        if (i < TILE)
            v[k] = img[(size_t)clampi(y0 - 2 + ty, 0, H - 1) * step + clampi(x0 - 2 + tx, 0, W - 1)];
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int i = threadIdx.x + k * 256;
        const int ty = i / TW, tx = i - ty * TW;
        if (i < TILE) s_t[ty][tx] = (uint8_t)v[k];
    }
    __syncthreads();
    const int q = threadIdx.x & 15;
    const int xg = x0 + 4 * q;
    if (xg >= W) return;
#pragma unroll
    for (int rg = 0; rg < kCenRows / 16; ++rg) {
        const int ly = (threadIdx.x >> 4) + 16 * rg;
        const int y = y0 + ly;
        if (y >= H) break;
        uint32_t lo[5], hi[5];  // window columns 4q .. 4q+7 of rows ly .. ly+4
#pragma unroll
        for (int dy = 0; dy < 5; ++dy) {
            const uint32_t* r = reinterpret_cast<const uint32_t*>(&s_t[ly + dy][4 * q]);
            lo[dy] = r[0];
            hi[dy] = r[1];
        }
        uint32_t cv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = byteOf(lo[2], hi[2], j + 2);
            uint32_t w = 0;
#pragma unroll
            for (int dy = 0; dy < 5; ++dy) {
#pragma unroll
                for (int dx = 0; dx < 5; ++dx)
                    w = __builtin_amdgcn_alignbit(w, (uint32_t)(byteOf(lo[dy], hi[dy], j + dx) - c), 31);
            }
            cv[j] = w;
        }
        const size_t p = (size_t)y * W + xg;
        uint32_t* const out = cen + (size_t)z * W * H + p;
        if (xg + 3 < W && (((size_t)z * W * H + p) & 3) == 0) {  // whole, 16-byte aligned group
            *reinterpret_cast<uint4*>(out) = make_uint4(cv[0], cv[1], cv[2], cv[3]);
            if (z == 0) *reinterpret_cast<uint4*>(keys + p) = make_uint4(~0u, ~0u, ~0u, ~0u);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (xg + j < W) {
                    out[j] = cv[j];
                    if (z == 0) keys[p + j] = 0xFFFFFFFFu;
                }
            }
        }
    }
}

// BT: the packed intervals of a pixel (x: gradient channel, y: intensity).
__device__ __forceinline__ void kPrepBTBody(Tile3 tl, const uint8_t* left, int lstep,
                                            const uint8_t* right, int rstep, uint2* desc,
                                            uint32_t* keys, int W, int H, int cap, Batch b)
{
    const int x = tl.x * 64 + (threadIdx.x & 63);
    const int y = tl.y * 4 + (threadIdx.x >> 6);
    const int z = tl.z & 1;          // view
    const size_t pr = tl.z >> 1;     // pair
    if (x >= W || y >= H) return;
    const uint8_t* img = z ? pairAt(right, pr, b.right) : pairAt(left, pr, b.left);
    const int step = z ? rstep : lstep;
    desc = pairAt(desc, pr, b.ws);
    keys = pairAt(keys, pr, b.ws);
    const size_t p = (size_t)y * W + x;
    const int ftzero = imax(cap, 15) | 1;
    const int xm = imax(x - 1, 0), xq = imin(x + 1, W - 1);
    const int gm = btGrad(img, step, W, H, xm, y, ftzero);
    const int g0 = btGrad(img, step, W, H, x, y, ftzero);
    const int gq = btGrad(img, step, W, H, xq, y, ftzero);
    const uint8_t* row = img + (size_t)y * step;
    desc[(size_t)z * W * H + p] = make_uint2(btInterval(g0, gm, gq), btInterval(row[x], row[xm], row[xq]));
    if (z == 0) keys[p] = 0xFFFFFFFFu;
}

__device__ __forceinline__ void store8(int16_t* dst, const int (&v)[8])
{
    uint4 o;
    o.x = (uint32_t)(uint16_t)v[0] | ((uint32_t)v[1] << 16);
    o.y = (uint32_t)(uint16_t)v[2] | ((uint32_t)v[3] << 16);
    o.z = (uint32_t)(uint16_t)v[4] | ((uint32_t)v[5] << 16);
    o.w = (uint32_t)(uint16_t)v[6] | ((uint32_t)v[7] << 16);
    *reinterpret_cast<uint4*>(dst) = o;
}

// Matching-cost volumes are stored complemented, E = SHRT_MAX - C (the form
// the path recurrence consumes): invalid and padding labels hold 0.
__device__ __forceinline__ void storeCost8(int16_t* dst, const int (&v)[8])
{
    int e[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) e[i] = kMax - v[i];
    store8(dst, e);
}

// Byte form (census levels whose costs and path costs fit a byte): 255 - C,
// 0 for SHRT_MAX; element index `at` addresses bytes.
template <bool NARROW>
__device__ __forceinline__ void storeCensus8(int16_t* C, size_t at, const int (&v)[8])
{
    if constexpr (NARROW) {
        uint32_t w[2] = {0u, 0u};
#pragma unroll
        for (int i = 0; i < 8; ++i)
            w[i >> 2] |= (uint32_t)(v[i] >= kMax ? 0 : 255 - v[i]) << (8 * (i & 3));
        *reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(C) + at) = make_uint2(w[0], w[1]);
    } else {
        storeCost8(C + at, v);
    }
}

// ---- Census cost over the full label range (sgm_kernels.cu:303-327) --------
// Block: 256 threads = NPX pixels x (Dp/8) label octets of one row segment; the
// right-view census words the segment can reach are staged in LDS.
// At a coarse pyramid level (geomMask) a label whose real disparity is outside
// [minD, min(maxD, x)] or whose right column is outside the image is stored as
// SHRT_MAX: that is exactly the reference's labelValid = geometry && C < MAX
// (sgm_kernels.cu:1135-1143), so the path kernels test "C < SHRT_MAX" alone.
template <bool NARROW>
__device__ __forceinline__ void kCensusCostBody(Tile3 tl, const uint32_t* cen, int16_t* C, int W,
                                                int H, int D, int Dp, int minD, int maxD,
                                                int geomMask, size_t wsStride)
{
    extern __shared__ uint32_t s_r[];
    cen = pairAt(cen, tl.z, wsStride);
    C = pairAt(C, tl.z, wsStride);
    const int tpp = Dp >> 3;
    const int npx = 256 / tpp;
    const int y = tl.y;
    const int x0 = tl.x * npx;
    const int base2 = x0 - minD - (Dp - 1);
    const int nR = npx + Dp - 1;
    const uint32_t* L = cen + (size_t)y * W;
    const uint32_t* R = cen + (size_t)W * H + (size_t)y * W;
    for (int i = threadIdx.x; i < nR; i += 256) {
        const int xx = base2 + i;
        s_r[i] = (xx >= 0 && xx < W) ? R[xx] : 0u;
    }
    __syncthreads();
    const int px = threadIdx.x / tpp;
    const int d0 = (threadIdx.x - px * tpp) * 8;
    const int x = x0 + px;
    if (x >= W) return;
    const uint32_t cl = L[x];
    const int hiReal = imin(maxD, x);
    if constexpr (NARROW) {
        // Byte form with label masks: the staged index (x - x0) + Dp - 1 - d is
        // always inside the staged row, so the eight reads share one address;
        // the byte is 255 - popc(l ^ r) = 223 + popc(~(l ^ r)).  Label i of the
        // thread (d = d0 + i) exists if d < D; its right column is in the image
        // if x - W + 1 <= minD + d <= x; at a coarse level it is also invalid
        // above hiReal.  Out-of-image columns cost 24 (byte 231) at single
        // scale and are invalid (byte 0) at coarse levels.
        const uint32_t* const Sc = s_r + (x - x0) + Dp - 1 - d0;
        uint32_t w[2] = {0u, 0u};
#pragma unroll
        for (int i = 0; i < 8; ++i)
            w[i >> 2] |= (223u + (uint32_t)__popc(~(cl ^ Sc[-i]))) << (8 * (i & 3));
        auto bits = [](int lo, int hi) {  // bits lo..hi of 0..7 (none if empty)
            lo = imax(lo, 0);
            hi = imin(hi, 7);
            return lo <= hi ? (0xFFu >> (7 - (hi - lo))) << lo : 0u;
        };
        auto bytes = [](uint32_t m) { return ((m * 0x00204081u) & 0x01010101u) * 0xFFu; };
        const uint32_t exist = bits(0, D - 1 - d0);
        const uint32_t in = bits(x - W + 1 - minD - d0, x - minD - d0);
        uint32_t keep = exist & in, out = 0u;
        if (geomMask) keep &= bits(0, hiReal - minD - d0);
        else out = exist & ~in;
        const uint32_t k0 = bytes(keep & 0xFu), k1 = bytes(keep >> 4);
        const uint32_t o0 = bytes(out & 0xFu) & 0xE7E7E7E7u, o1 = bytes(out >> 4) & 0xE7E7E7E7u;
        *reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(C) + ((size_t)y * W + x) * Dp + d0) =
            make_uint2((w[0] & k0) | o0, (w[1] & k1) | o1);
        return;
    }
    int v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int d = d0 + i;
        const int x2 = x - d - minD;
        const bool in = x2 >= 0 && x2 < W;
        v[i] = d >= D ? kMax : (in ? __popc(cl ^ s_r[x2 - base2]) : 24);
        if (geomMask && (!in || minD + d > hiReal)) v[i] = kMax;
    }
    storeCensus8<NARROW>(C, ((size_t)y * W + x) * Dp + d0, v);
}

// ---- refinement labels (sgm_kernels.cu:955-1000) ----------------------------
// The window of pixel x from its parent's raw disparity and confidence.
__device__ __forceinline__ void refineFrom(const RefineIn& r, int raw, float cf, int x, int D,
                                           int minD, int maxD, int& base, int& valid)
{
    const int pinv = (int)(int16_t)((minD - 1) * 16);
    if (raw == pinv || cf < r.minConf) {
        base = minD;
        valid = 0;
        return;
    }
    const int p16 = raw * 2;
    const int center = (p16 + 8) >> 4;
    const int res = p16 - center * 16;
    int b = center + (res < 0 ? -(D / 2) : -(D / 2) + 1);
    const int largest = imin(maxD, x) - D + 1;
    b = largest >= minD ? imin(imax(b, minD), largest) : minD;
    base = b;
    valid = 1;
}

__device__ __forceinline__ int parentRaw(const RefineIn& r, int x, int y)
{
    return r.pdisp[(size_t)imin(y >> 1, r.PH - 1) * r.pdpitch + imin(x >> 1, r.PW - 1)];
}

__device__ __forceinline__ float parentConf(const RefineIn& r, int x, int y)
{
    return r.pconf[(size_t)imin(y >> 1, r.PH - 1) * r.pcpitch + imin(x >> 1, r.PW - 1)];
}

__device__ __forceinline__ void refineLabel(const RefineIn& r, int x, int y, int D, int minD,
                                            int maxD, int& base, int& valid)
{
    refineFrom(r, parentRaw(r, x, y), parentConf(r, x, y), x, D, minD, maxD, base, valid);
}

// Label word of a refinement pixel: bits 0-15 the window base (int16), bit 16
// the parent-validity flag, bits 24-31 the left-image intensity, which the
// adaptive-P2 path step reads from here instead of a separate image load.
__device__ __forceinline__ int32_t labelWord(int base, int valid, uint8_t im)
{
    return (int32_t)(((uint32_t)base & 0xFFFFu) | ((uint32_t)valid << 16) | ((uint32_t)im << 24));
}

// ---- Census cost at refinement levels (sgm_kernels.cu:1112-1133) ------------
// Block: NPX pixels x (Dp/8) label octets of ROWS rows (2, 4 or 8; 4 unless a
// tuner picked another).  A label's right
// column x - real lies in [x0 - maxD, x0 + NPX - 1 - minD]; when that span is
// modest (STAGED) the right census words of the block's rows are staged in LDS
// first, otherwise read from global memory.  Every global load of a thread
// (staged words, parent disparity and confidence, left census word,
// intensity, for all its rows) is issued before the first wait.
template <bool STAGED, bool NARROW, int ROWS>
__device__ __forceinline__ void kCensusLocalBody(Tile3 tl, const uint32_t* cen, RefineIn r,
                                                 int32_t* labels, const uint8_t* img, int istep,
                                                 int16_t* C, int W, int H, int D, int Dp, int minD,
                                                 int maxD, Batch b)
{
    constexpr int NPRE = 2;  // staged words per thread and row loaded ahead
    extern __shared__ uint32_t s_r[];
    const size_t pr = tl.z;
    cen = pairAt(cen, pr, b.ws);
    labels = pairAt(labels, pr, b.ws);
    C = pairAt(C, pr, b.ws);
    img = pairAt(img, pr, b.left);
    r.pdisp = pairAt(r.pdisp, pr, b.pdisp);
    r.pconf = pairAt(r.pconf, pr, b.pconf);
    const int tpp = Dp >> 3;
    const int npx = 256 / tpp;
    const int y0 = tl.y * ROWS;
    const int x0 = tl.x * npx;
    const int xb = x0 - maxD;  // staged column 0
    const int span = npx + maxD - minD;
    const int px = threadIdx.x / tpp;
    const int d0 = (threadIdx.x - px * tpp) * 8;
    const int x = x0 + px;
    const int xc = imin(x, W - 1);
    uint32_t pre[ROWS][NPRE];
    int raw[ROWS];
    float cf[ROWS];
    uint32_t cl[ROWS];
    uint8_t im[ROWS];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int y = imin(y0 + rr, H - 1);
        if constexpr (STAGED) {
            const uint32_t* R = cen + (size_t)W * H + (size_t)y * W;
#pragma unroll
            for (int k = 0; k < NPRE; ++k) {
                const int i = threadIdx.x + k * 256;
                const int xx = xb + i;
                pre[rr][k] = (i < span && xx >= 0 && xx < W) ? R[xx] : 0u;
            }
        }
        raw[rr] = parentRaw(r, xc, y);
        cf[rr] = parentConf(r, xc, y);
        cl[rr] = cen[(size_t)y * W + xc];
        im[rr] = img[(size_t)y * istep + xc];
    }
    if constexpr (STAGED) {
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) {
#pragma unroll
            for (int k = 0; k < NPRE; ++k) {
                const int i = threadIdx.x + k * 256;
                if (i < span) s_r[rr * span + i] = pre[rr][k];
            }
            const uint32_t* R = cen + (size_t)W * H + (size_t)imin(y0 + rr, H - 1) * W;
            for (int i = threadIdx.x + NPRE * 256; i < span; i += 256) {
                const int xx = xb + i;
                s_r[rr * span + i] = (xx >= 0 && xx < W) ? R[xx] : 0u;
            }
        }
        __syncthreads();
    }
    if (x >= W) return;
    const int lo = imax(minD, x - W + 1), hi = imin(maxD, x);
    // Byte volumes from staged words over a full label range (uniform): the
    // staged index x - (base + d) - xb of every label lies in [0, span),
    // because refineFrom keeps base in [minD, maxD - D + 1] and d < D.  The
    // eight reads of a thread then share one address (immediate offsets), the
    // byte is 255 - popc(l ^ r) = 223 + popc(~(l ^ r)), and the label range
    // [lo, hi] and the parent flag are applied as one byte mask.
    const bool direct = STAGED && NARROW && Dp == D && maxD - minD + 1 >= D;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int y = y0 + rr;
        if (y >= H) break;
        int base, valid;
        refineFrom(r, raw[rr], cf[rr], x, D, minD, maxD, base, valid);
        const size_t p = (size_t)y * W + x;
        if (d0 == 0) labels[p] = labelWord(base, valid, im[rr]);
        if (STAGED && NARROW && direct) {
            const uint32_t* const Sc = s_r + rr * span + (x - base - d0 - xb);
            uint32_t w[2] = {0u, 0u};
#pragma unroll
            for (int i = 0; i < 8; ++i)
                w[i >> 2] |= (223u + (uint32_t)__popc(~(cl[rr] ^ Sc[-i]))) << (8 * (i & 3));
            // Labels d0 + i with lo <= base + d0 + i <= hi, as bits, then bytes.
            const int ilo = imax(lo - base - d0, 0), ihi = imin(hi - base - d0, 7);
            const uint32_t ok = (valid && ilo <= ihi) ? ((0xFFu >> (7 - (ihi - ilo))) << ilo) : 0u;
            auto bytes = [](uint32_t m) { return ((m * 0x00204081u) & 0x01010101u) * 0xFFu; };
            *reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(C) + p * Dp + d0) =
                make_uint2(w[0] & bytes(ok & 0xFu), w[1] & bytes(ok >> 4));
            continue;
        }
        const uint32_t* R = cen + (size_t)W * H + (size_t)y * W;
        const uint32_t* S = s_r + rr * span;
        int v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int d = d0 + i;
            const int real = base + d;
            const bool ok = valid && d < D && real >= lo && real <= hi;
            // ok => 0 <= x - real < W and x - real in the staged span.
            const int x2 = ok ? x - real : (STAGED ? xb : 0);
            const uint32_t cr = STAGED ? S[x2 - xb] : R[x2];
            v[i] = ok ? __popc(cl[rr] ^ cr) : kMax;
        }
        storeCensus8<NARROW>(C, p * Dp + d0, v);
    }
}

// Packed form of one BT descriptor: (v, v0, v1) of both channels as 16-bit
// pairs, gradient channel low, intensity channel high.
struct BTPacked {
    uint32_t v, v0, v1;
};

__device__ __forceinline__ BTPacked btPack(uint2 d)
{
    // Output bytes: 0 = x byte k, 1 = 0, 2 = y byte k, 3 = 0.
    constexpr uint32_t z = 0x0C000C00u;
    return {__builtin_amdgcn_perm(d.y, d.x, z | (4u << 16) | 0u),
            __builtin_amdgcn_perm(d.y, d.x, z | (5u << 16) | 1u),
            __builtin_amdgcn_perm(d.y, d.x, z | (6u << 16) | 2u)};
}

// A right-view column outside the image: every channel distance is at least
// 0x7FFF - 255, so the cost below is exactly 255 (sgm_kernels.cu:1037-1040).
__device__ __forceinline__ BTPacked btPoison() { return {0x7FFF7FFFu, 0x7FFF7FFFu, 0x7FFF7FFFu}; }

// One BT interval distance per 16-bit half: max(0, u - v1, v0 - u) equals
// max(u, v0) - min(u, v1) because v0 <= v1 (sgm_kernels.cu:1025-1027).
__device__ __forceinline__ uint32_t btDistPacked(const BTPacked& l, const BTPacked& r)
{
    typedef uint16_t u2 __attribute__((ext_vector_type(2)));
    auto V = [](uint32_t x) { return __builtin_bit_cast(u2, x); };
    const u2 c0 = __builtin_elementwise_max(V(l.v), V(r.v0)) - __builtin_elementwise_min(V(l.v), V(r.v1));
    const u2 c1 = __builtin_elementwise_max(V(r.v), V(l.v0)) - __builtin_elementwise_min(V(r.v), V(l.v1));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(c0, c1));
}

// btCost of two packed descriptors: min(255, grad + (int >> 2)).
__device__ __forceinline__ int btCostPacked(const BTPacked& l, const BTPacked& r)
{
    const uint32_t c = btDistPacked(l, r);
    return imin(255, (int)(c & 0xFFFFu) + (int)(c >> 18));
}

// ---- BT cost + horizontal window sum (sgm_kernels.cu:139-239, :354-379) ------
// Block: one 64-pixel row segment x 64 labels.  BT costs of the segment and its
// +-half halo (clamped columns) go to LDS as bytes (from the
// segment's descriptors, unpacked into LDS first); each thread then sums the
// window for 8 labels of one pixel and saturates at SHRT_MAX.
constexpr int kSegX = 64;
constexpr int kChunkD = 64;

__device__ __forceinline__ void kBTRowBody(Tile3 tl, const uint2* desc, int16_t* Hs, int W, int H,
                                           int Dp, int minD, int half, size_t wsStride)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t s_bt[];
    const int dc = imin(Dp, kChunkD);      // a power of two
    const int dsh = 31 - __clz(dc);
    const int nzc = Dp / dc;  // label chunks per pair
    const size_t pr = tl.z / nzc;
    desc = pairAt(desc, pr, wsStride);
    Hs = pairAt(Hs, pr, wsStride);
    const int y = tl.y;
    const int x0 = tl.x * kSegX;
    const int dc0 = (tl.z - (int)pr * nzc) * dc;
    const int ext = kSegX + 2 * half;
    const uint2* L = desc + (size_t)y * W;
    const uint2* R = desc + (size_t)W * H + (size_t)y * W;
    // Both views' descriptors of the segment, unpacked once into LDS: left
    // columns xx = clamp(x0 - half + o), right columns x2 = xx - d - minD for
    // the chunk's labels d (rbase .. rbase + nr - 1).
    const int btBytes = (ext * dc + 15) & ~15;
    BTPacked* s_l = reinterpret_cast<BTPacked*>(s_bt + btBytes);
    BTPacked* s_r = s_l + ext;
    const int xlo = clampi(x0 - half, 0, W - 1), xhi = clampi(x0 + kSegX - 1 + half, 0, W - 1);
    const int rbase = xlo - (dc0 + dc - 1) - minD;
    const int nr = xhi - xlo + dc;  // <= ext + dc - 1
    for (int o = threadIdx.x; o < ext; o += 256) s_l[o] = btPack(L[clampi(x0 - half + o, 0, W - 1)]);
    for (int j = threadIdx.x; j < nr; j += 256) {
        const int x2 = rbase + j;
        s_r[j] = (x2 >= 0 && x2 < W) ? btPack(R[x2]) : btPoison();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ext * dc; i += 256) {
        const int o = i >> dsh;
        const int dl = i - (o << dsh);
        const int xx = clampi(x0 - half + o, 0, W - 1);
        // An out-of-image right column holds the poison descriptor: cost 255.
        s_bt[i] = (uint8_t)btCostPacked(s_l[o], s_r[xx - (dc0 + dl) - minD - rbase]);
    }
    __syncthreads();
    const int tpp = dc >> 3;
    for (int t = threadIdx.x; t < kSegX * tpp; t += 256) {
        const int px = t / tpp;
        const int dl0 = (t - px * tpp) * 8;
        const int x = x0 + px;
        if (x >= W) continue;
        int v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = 0;
        for (int k = 0; k <= 2 * half; ++k) {
            const uint8_t* b = s_bt + (px + k) * dc + dl0;
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] += b[i];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = imin(kMax, v[i]);
        store8(Hs + ((size_t)y * W + x) * Dp + dc0 + dl0, v);
    }
}
__device__ __forceinline__ void kBoxVBody(Tile3 tl, const int16_t* Hs, int16_t* C, int W, int H,
                                          int D, int Dp, int half, int minD, int maxD, int geomMask,
                                          size_t wsStride)
{
    Hs = pairAt(Hs, tl.y, wsStride);
    C = pairAt(C, tl.y, wsStride);
    const size_t oct = (size_t)tl.x * 256 + threadIdx.x;
    const size_t nOct = (size_t)W * H * (Dp >> 3);
    if (oct >= nOct) return;
    const size_t p = oct / (Dp >> 3);
    const int d0 = (int)(oct - p * (Dp >> 3)) * 8;
    const int y = (int)(p / W);
    const int x = (int)(p - (size_t)y * W);
    int v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = 0;
    for (int k = -half; k <= half; ++k) {
        const int yy = clampi(y + k, 0, H - 1);
        const uint4 w = *reinterpret_cast<const uint4*>(Hs + ((size_t)yy * W + x) * Dp + d0);
        v[0] += (int16_t)(w.x & 0xFFFF); v[1] += (int)w.x >> 16;
        v[2] += (int16_t)(w.y & 0xFFFF); v[3] += (int)w.y >> 16;
        v[4] += (int16_t)(w.z & 0xFFFF); v[5] += (int)w.z >> 16;
        v[6] += (int16_t)(w.w & 0xFFFF); v[7] += (int)w.w >> 16;
    }
    const int hiReal = imin(maxD, x);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[i] = imin(kMax, v[i]);
        const int real = minD + d0 + i;
        // Padding labels, and at coarse pyramid levels geometry-invalid labels,
        // carry SHRT_MAX (see kCensusCost).
        if (d0 + i >= D || (geomMask && (real > hiReal || x - real < 0 || x - real >= W))) v[i] = kMax;
    }
    storeCost8(C + p * Dp + d0, v);
}

// ---- BT window cost at refinement levels (sgm_kernels.cu:1075-1110) ----------
// Window sums of 8 consecutive labels (real disparities r0 .. r0 + 7) of a
// pixel whose window columns x - HALF .. x + HALF all lie inside the image:
// per window row, the 2 HALF + 1 left and 2 HALF + 8 right descriptors the
// sums touch are loaded and unpacked once.
// CLAMP = false when a cost never reaches 255 (gradient values at most
// 2 * 95: 190 + 255 / 4 < 255): the two channel terms (grad, int >> 2) are summed
// in packed 16-bit halves (at most 25 * 255 each) and added at the end, and an
// out-of-image right column contributes exactly (255, 0).
template <int HALF, bool CLAMP>
__device__ __forceinline__ void btWindowInterior(const uint2* Ld, const uint2* Rd, int W, int H,
                                                 int x, int y, int r0, int sum[8])
{
    typedef uint16_t u2 __attribute__((ext_vector_type(2)));
    constexpr int NX = 2 * HALF + 1;
    constexpr int NR = NX + 7;
    const int xr0 = x - HALF - r0 - 7;  // right column of (window column 0, label 7)
    u2 acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = u2{0, 0};
#pragma unroll 1
    for (int dy = -HALF; dy <= HALF; ++dy) {
        const int yy = clampi(y + dy, 0, H - 1);
        const uint2* Lrow = Ld + (size_t)yy * W;
        const uint2* Rrow = Rd + (size_t)yy * W;
        uint2 lraw[NX], rraw[NR];
#pragma unroll
        for (int k = 0; k < NX; ++k) lraw[k] = Lrow[x - HALF + k];
#pragma unroll
        for (int m = 0; m < NR; ++m) rraw[m] = Rrow[clampi(xr0 + m, 0, W - 1)];
        BTPacked lp[NX], rp[NR];
        bool in[NR];
#pragma unroll
        for (int k = 0; k < NX; ++k) lp[k] = btPack(lraw[k]);
#pragma unroll
        for (int m = 0; m < NR; ++m) {
            const int x2 = xr0 + m;
            in[m] = x2 >= 0 && x2 < W;
            rp[m] = CLAMP && !in[m] ? btPoison() : btPack(rraw[m]);
        }
        // Window column k, label i: right column x - HALF + k - (r0 + i) = xr0 + (7 + k - i).
#pragma unroll
        for (int k = 0; k < NX; ++k)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int m = 7 + k - i;
                if constexpr (CLAMP) {
                    sum[i] += btCostPacked(lp[k], rp[m]);
                } else {
                    const u2 t = __builtin_bit_cast(u2, btDistPacked(lp[k], rp[m])) >> u2{0, 2};
                    acc[i] += in[m] ? t : u2{255, 0};
                }
            }
    }
    if constexpr (!CLAMP) {
#pragma unroll
        for (int i = 0; i < 8; ++i) sum[i] += (int)acc[i][0] + (int)acc[i][1];
    }
}

template <bool CLAMP>
__device__ __forceinline__ void kBTLocalBody(Tile3 tl, const uint2* desc, RefineIn r,
                                             int32_t* labels, const uint8_t* img, int istep,
                                             int16_t* C, int W, int H, int D, int Dp, int minD,
                                             int maxD, int half, Batch b)
{
    const size_t pr = tl.z;
    desc = pairAt(desc, pr, b.ws);
    labels = pairAt(labels, pr, b.ws);
    C = pairAt(C, pr, b.ws);
    img = pairAt(img, pr, b.left);
    r.pdisp = pairAt(r.pdisp, pr, b.pdisp);
    r.pconf = pairAt(r.pconf, pr, b.pconf);
    const int tpp = Dp >> 3;
    const int npx = 256 / tpp;
    const int y = tl.y;
    const int px = threadIdx.x / tpp;
    const int d0 = (threadIdx.x - px * tpp) * 8;
    const int x = tl.x * npx + px;
    if (x >= W) return;
    int base, valid;
    refineLabel(r, x, y, D, minD, maxD, base, valid);
    const size_t p = (size_t)y * W + x;
    if (d0 == 0) labels[p] = labelWord(base, valid, img[(size_t)y * istep + x]);
    const int lo = imax(minD, x - W + 1), hi = imin(maxD, x);
    const uint2* Ld = desc;
    const uint2* Rd = desc + (size_t)W * H;
    bool ok[8];
    int sum[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int real = base + d0 + i;
        ok[i] = valid && d0 + i < D && real >= lo && real <= hi;
        sum[i] = 0;
    }
    const bool interior = x - half >= 0 && x + half < W;
    if (interior && half == 2) {
        btWindowInterior<2, CLAMP>(Ld, Rd, W, H, x, y, base + d0, sum);
    } else if (interior && half == 1) {
        btWindowInterior<1, CLAMP>(Ld, Rd, W, H, x, y, base + d0, sum);
    } else {
        // Border columns (replicate-clamped window) and other block sizes.
        for (int dy = -half; dy <= half; ++dy) {
            const int yy = clampi(y + dy, 0, H - 1);
            const uint2* Lrow = Ld + (size_t)yy * W;
            const uint2* Rrow = Rd + (size_t)yy * W;
            for (int dx = -half; dx <= half; ++dx) {
                const int xx = clampi(x + dx, 0, W - 1);
                const uint2 li = Lrow[xx];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int x2 = xx - (base + d0 + i);
                    sum[i] += (x2 >= 0 && x2 < W) ? btCost(li, Rrow[x2]) : 255;
                }
            }
        }
    }
    int v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = ok[i] ? imin(sum[i], kMax - 1) : kMax;
    storeCost8(C + p * Dp + d0, v);
}

// Tiled form for 16-label windows (every StereoMatcher refinement level) when
// no cost can reach 255 (packed channel sums, see btWindowInterior) and the
// window is at most 7x7.  Block: kTileRows rows of kTilePx pixels.  The window
// sum of pixel (x, y) at real disparity r is the sum over its 2 HALF + 1
// window columns xx of the column sum V_y[r][xx] = sum over rows
// y - HALF .. y + HALF of cost(L(xx, yy), R(xx - r, yy)) (rows and columns
// replicate-clamped, 255 where xx - r leaves the image).  V depends on
// (y, r, xx) only: the block computes it for every real in the union
// [rlo, rhi] of its pixels' windows and every window column, in LDS, sliding
// it down one row at a time (one cost enters, one leaves), and each
// (pixel, label) adds 2 HALF + 1 entries.  A block whose union is wider than
// kTileR reals (a depth edge with far-apart windows) takes the plain loop.
constexpr int kTilePx = 64;
constexpr int kTileRows = 4;
constexpr int kTileR = 48;

template <int HALF>
__device__ __forceinline__ void kBTLocalTileBody(Tile3 tl, const uint2* desc, RefineIn r,
                                                 int32_t* labels, const uint8_t* img, int istep,
                                                 int16_t* C, int W, int H, int minD, int maxD,
                                                 Batch b)
{
    typedef uint16_t u2 __attribute__((ext_vector_type(2)));
    constexpr int NC = kTilePx + 2 * HALF;       // window columns of the block
    constexpr int NY = 2 * HALF + 1;             // window rows of one output row
    constexpr int NYT = kTileRows + 2 * HALF;    // window rows of the block
    constexpr int NR2 = NC + kTileR - 1;         // right columns the column sums touch
    constexpr int D = 16;
    static_assert(kTilePx == 64 && kTileRows * kTilePx == 256, "one wave per block row");
    __shared__ BTPacked s_l[NYT][NC];
    __shared__ BTPacked s_r[NYT][NR2];
    __shared__ uint32_t s_v[kTileR][NC];
    __shared__ int s_base[kTileRows][kTilePx];
    __shared__ int s_lo[kTileRows], s_hi[kTileRows];
    const size_t pr = tl.z;
    desc = pairAt(desc, pr, b.ws);
    labels = pairAt(labels, pr, b.ws);
    C = pairAt(C, pr, b.ws);
    img = pairAt(img, pr, b.left);
    r.pdisp = pairAt(r.pdisp, pr, b.pdisp);
    r.pconf = pairAt(r.pconf, pr, b.pconf);
    const int y0 = tl.y * kTileRows, x0 = tl.x * kTilePx, tid = threadIdx.x;
    const uint2* Ld = desc;
    const uint2* Rd = desc + (size_t)W * H;
    // Label windows: wave w holds row y0 + w.
    {
        const int row = tid >> 6, px = tid & 63;
        const int x = x0 + px, y = y0 + row;
        int base = 0, valid = 0;
        if (x < W && y < H) {
            refineLabel(r, x, y, D, minD, maxD, base, valid);
            labels[(size_t)y * W + x] = labelWord(base, valid, img[(size_t)y * istep + x]);
        }
        s_base[row][px] = valid ? base : INT_MIN;
        int lo = valid ? base : INT_MAX, hi = valid ? base + D - 1 : INT_MIN;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo = imin(lo, __shfl_xor(lo, off));
            hi = imax(hi, __shfl_xor(hi, off));
        }
        if (px == 0) {
            s_lo[row] = lo;
            s_hi[row] = hi;
        }
    }
    __syncthreads();
    int rlo = s_lo[0], rhi = s_hi[0];
#pragma unroll
    for (int k = 1; k < kTileRows; ++k) {
        rlo = imin(rlo, s_lo[k]);
        rhi = imax(rhi, s_hi[k]);
    }
    // Output of thread (px, q): labels q .. q + 3 of pixel px of one row.
    const int opx = tid >> 2, q = (tid & 3) * 4;
    const int ox = x0 + opx;
    auto store4 = [&](int y, const int (&v)[4]) {
        uint2 o;
        o.x = (uint32_t)(uint16_t)(kMax - v[0]) | ((uint32_t)(uint16_t)(kMax - v[1]) << 16);
        o.y = (uint32_t)(uint16_t)(kMax - v[2]) | ((uint32_t)(uint16_t)(kMax - v[3]) << 16);
        *reinterpret_cast<uint2*>(C + ((size_t)y * W + ox) * D + q) = o;
    };
    if (rlo > rhi) {
        // No valid pixel: every label is SHRT_MAX.
        const int v[4] = {kMax, kMax, kMax, kMax};
        if (ox < W)
            for (int row = 0; row < kTileRows && y0 + row < H; ++row) store4(y0 + row, v);
        return;
    }
    if (rhi - rlo + 1 <= kTileR) {
        const int xlo = clampi(x0 - HALF, 0, W - 1);
        const int xl2 = xlo - rhi;  // right column of s_r[.][0]
        const int nr2 = clampi(x0 + kTilePx - 1 + HALF, 0, W - 1) - rlo - xl2 + 1;  // <= NR2
        for (int i = tid; i < NYT * NC; i += 256) {
            const int dy = i / NC, c = i - dy * NC;
            const int yy = clampi(y0 - HALF + dy, 0, H - 1);
            s_l[dy][c] = btPack(Ld[(size_t)yy * W + clampi(x0 - HALF + c, 0, W - 1)]);
        }
        for (int i = tid; i < NYT * NR2; i += 256) {
            const int dy = i / NR2, j = i - dy * NR2;
            if (j >= nr2) continue;
            const int yy = clampi(y0 - HALF + dy, 0, H - 1);
            const int x2 = xl2 + j;
            s_r[dy][j] = (x2 >= 0 && x2 < W) ? btPack(Rd[(size_t)yy * W + x2]) : btPoison();
        }
        __syncthreads();
        const int RN = (rhi - rlo + 1) * NC;
        auto ev = [&](int dy, int c, int jr) {
            return __builtin_bit_cast(u2, btDistPacked(s_l[dy][c], s_r[dy][jr])) >> u2{0, 2};
        };
#pragma unroll 1
        for (int row = 0; row < kTileRows; ++row) {
            // Entry (rr, c) belongs to thread i % 256 for every row: it slides its
            // own column sums down in place.
#pragma unroll 1
            for (int i = tid; i < RN; i += 256) {
                const int rr = i / NC, c = i - rr * NC;
                const int x2 = clampi(x0 - HALF + c, 0, W - 1) - (rlo + rr);
                if (x2 >= 0 && x2 < W) {
                    const int jr = x2 - xl2;
                    u2 acc;
                    if (row == 0) {
                        acc = u2{0, 0};
#pragma unroll
                        for (int dy = 0; dy < NY; ++dy) acc += ev(dy, c, jr);
                    } else {
                        acc = __builtin_bit_cast(u2, s_v[rr][c]) + ev(row + NY - 1, c, jr) - ev(row - 1, c, jr);
                    }
                    s_v[rr][c] = __builtin_bit_cast(uint32_t, acc);
                } else if (row == 0) {
                    s_v[rr][c] = 255u * NY;  // (255 * NY, 0): every window row outside the image
                }
            }
            __syncthreads();
            const int y = y0 + row;
            if (y < H && ox < W) {
                const int base = s_base[row][opx];
                const int lo = imax(minD, ox - W + 1), hi = imin(maxD, ox);
                int v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int real = base + q + i;
                    v[i] = kMax;
                    if (base != INT_MIN && real >= lo && real <= hi) {
                        const int rr = real - rlo;
                        u2 sum = u2{0, 0};
#pragma unroll
                        for (int k = 0; k < NY; ++k) sum += __builtin_bit_cast(u2, s_v[rr][opx + k]);
                        v[i] = imin((int)sum[0] + (int)sum[1], kMax - 1);
                    }
                }
                store4(y, v);
            }
            __syncthreads();
        }
        return;
    }
    // Wide union (about 0.5% of blocks in the bench scenes): the plain
    // per-pixel loop, one thread per pixel, four labels at a time.
    const int prow = tid >> 6, ppx = tid & 63;
    const int x = x0 + ppx, y = y0 + prow;
    if (x >= W || y >= H) return;
    const int base = s_base[prow][ppx];
    const bool valid = base != INT_MIN;
    const int lo = imax(minD, x - W + 1), hi = imin(maxD, x);
#pragma unroll 1
    for (int d0 = 0; d0 < D; d0 += 8) {
        int sum[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) sum[i] = 0;
        if (valid) {
#pragma unroll 1
            for (int dy = -HALF; dy <= HALF; ++dy) {
                const int yy = clampi(y + dy, 0, H - 1);
                const uint2* Lrow = Ld + (size_t)yy * W;
                const uint2* Rrow = Rd + (size_t)yy * W;
#pragma unroll 1
                for (int dx = -HALF; dx <= HALF; ++dx) {
                    const int xx = clampi(x + dx, 0, W - 1);
                    const uint2 li = Lrow[xx];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const int x2 = xx - (base + d0 + i);
                        sum[i] += (x2 >= 0 && x2 < W) ? btCost(li, Rrow[x2]) : 255;
                    }
                }
            }
        }
        int v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int real = base + d0 + i;
            v[i] = valid && real >= lo && real <= hi ? imin(sum[i], kMax - 1) : kMax;
        }
        storeCost8(C + ((size_t)y * W + x) * D + d0, v);
    }
}

// ---- pyrDown (pyrdown.cu:16-52) ------------------------------------------------
__device__ __forceinline__ int reflect101(int i, int n)
{
    if (i < 0) i = -i;
    if (i >= n) i = 2 * n - 2 - i;
    return i < 0 ? 0 : i;
}

// Block: 64 x 8 destination pixels of one image, two adjacent rows per thread;
// the 131 x 19 source window they read (reflect-101 coordinates) is loaded into
// registers with every load in flight at once, then staged in LDS (rows padded
// to 132 bytes, so each row starts on a dword).
constexpr int kPyrRows = 8;

__device__ __forceinline__ void kPyrDownBody(Tile3 tl, const uint8_t* srcA, int sstepA,
                                             const uint8_t* srcB, int sstepB, size_t sStride, int sw,
                                             int sh, uint8_t* dstA, int dstepA, uint8_t* dstB,
                                             int dstepB, size_t dStride, int dw, int dh, int nv)
{
    constexpr int TW = 131, TH = 2 * kPyrRows + 3, TILE = TW * TH, NL = (TILE + 255) / 256;
    __shared__ __attribute__((aligned(16))) uint8_t s_t[TH][TW + 1];
    static_assert((TW + 1) % 4 == 0, "rows start on a dword");
    const size_t pr = tl.z / nv;              // pair
    const int view = (int)(tl.z - pr * nv);   // 0: A, 1: B
    const uint8_t* src = pairAt(view ? srcB : srcA, pr, sStride);
    const int sstep = view ? sstepB : sstepA;
    const int x0 = tl.x * 64, y0 = tl.y * kPyrRows;
    uint32_t v[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int i = threadIdx.x + k * 256;
        const int ty = i / TW, tx = i - ty * TW;
        if (i < TILE)
            v[k] = src[(size_t)reflect101(2 * y0 - 2 + ty, sh) * sstep + reflect101(2 * x0 - 2 + tx, sw)];
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int i = threadIdx.x + k * 256;
        const int ty = i / TW, tx = i - ty * TW;
        if (i < TILE) s_t[ty][tx] = (uint8_t)v[k];
    }
    __syncthreads();
    const int lx = threadIdx.x & 63;
    const int x = x0 + lx;
    if (x >= dw) return;
    uint8_t* dst = pairAt(view ? dstB : dstA, pr, dStride);
    const int dstep = view ? dstepB : dstepA;
    // Two vertically adjacent outputs per thread (rows 2rg, 2rg+1 of the
    // block): the 5-tap row sums of the 7 source rows they share are formed
    // once from two dword reads per row; the column filter then combines 5 of
    // them per output.  The separable integer sums equal the 5x5 ones.
    const int ly = 2 * (threadIdx.x >> 6);
    const int off = 2 * (lx & 1);  // byte offset of column 2lx in its dword
    int hs[7];
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        const uint32_t* row = reinterpret_cast<const uint32_t*>(s_t[2 * ly + r]) + (lx >> 1);
        const uint32_t d0 = row[0], d1 = row[1];
        const uint32_t u = __builtin_amdgcn_alignbyte(d1, d0, (uint32_t)off);  // bytes 2lx .. 2lx+3
        const int b4 = (int)((d1 >> (8 * off)) & 0xFFu);                         // byte 2lx+4
        hs[r] = (int)(u & 0xFFu) + b4 + 4 * ((int)((u >> 8) & 0xFFu) + (int)(u >> 24)) +
                6 * (int)((u >> 16) & 0xFFu);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int y = y0 + ly + k;
        if (y >= dh) break;
        const int* h = hs + 2 * k;
        const int sum = h[0] + h[4] + 4 * (h[1] + h[3]) + 6 * h[2];
        dst[(size_t)y * dstep + x] = (uint8_t)((sum + 128) >> 8);
    }
}

// ---- per-launch and queued kernels ----------------------------------------
// Variant codes: kStCensusCost 0/1 = int16 / byte volumes; kStCensusLocal
// staged | narrow << 1 | (rows 2/4/8 -> 0/1/2) << 2; kStBTLocal 0/1 = clamp
// off / on; kStBTLocalTile = the window's half size.
constexpr int censusLocalVariant(bool staged, bool narrow, int rows)
{
    return (staged ? 1 : 0) | (narrow ? 2 : 0) | ((rows == 2 ? 0 : (rows == 8 ? 2 : 1)) << 2);
}
constexpr int censusLocalRows(int v) { return (v >> 2) == 0 ? 2 : ((v >> 2) == 2 ? 8 : 4); }

template <int KIND, int V>
__device__ __forceinline__ void stageBody(const StageArgs& a, Tile3 tl)
{
    const uint32_t* cen = reinterpret_cast<const uint32_t*>(a.desc);
    if constexpr (KIND == kStPrepCensus)
        kPrepCensusBody(tl, a.left, a.lstep, a.right, a.rstep, reinterpret_cast<uint32_t*>(a.desc),
                        a.keys, a.W, a.H, a.batch);
    else if constexpr (KIND == kStPrepBT)
        kPrepBTBody(tl, a.left, a.lstep, a.right, a.rstep, a.desc, a.keys, a.W, a.H, a.cap, a.batch);
    else if constexpr (KIND == kStCensusCost)
        kCensusCostBody<V != 0>(tl, cen, a.C, a.W, a.H, a.D, a.Dp, a.minD, a.maxD, a.geom, a.batch.ws);
    else if constexpr (KIND == kStCensusLocal)
        kCensusLocalBody<(V & 1) != 0, (V & 2) != 0, censusLocalRows(V)>(
            tl, cen, a.r, a.labels, a.left, a.lstep, a.C, a.W, a.H, a.D, a.Dp, a.minD, a.maxD, a.batch);
    else if constexpr (KIND == kStBTRow)
        kBTRowBody(tl, a.desc, a.Hs, a.W, a.H, a.Dp, a.minD, a.half, a.batch.ws);
    else if constexpr (KIND == kStBoxV)
        kBoxVBody(tl, a.Hs, a.C, a.W, a.H, a.D, a.Dp, a.half, a.minD, a.maxD, a.geom, a.batch.ws);
    else if constexpr (KIND == kStBTLocal)
        kBTLocalBody<V != 0>(tl, a.desc, a.r, a.labels, a.left, a.lstep, a.C, a.W, a.H, a.D, a.Dp,
                             a.minD, a.maxD, a.half, a.batch);
    else if constexpr (KIND == kStBTLocalTile)
        kBTLocalTileBody<V>(tl, a.desc, a.r, a.labels, a.left, a.lstep, a.C, a.W, a.H, a.minD,
                            a.maxD, a.batch);
    else if constexpr (KIND == kStPyrDown)
        kPyrDownBody(tl, a.left, a.lstep, a.right, a.rstep, a.sStride, a.sw, a.sh, a.dstA, a.dstepA,
                     a.dstB, a.dstepB, a.dStride, a.W, a.H, a.nv);
}

// Per-launch kernels: tiles XCD-contiguous, except the BT row, box and plain
// local kernels, which keep the hardware order.
__global__ __launch_bounds__(256) void kPrepCensus(StageArgs a) { stageBody<kStPrepCensus, 0>(a, xcdTile()); }
__global__ __launch_bounds__(256) void kPrepBT(StageArgs a) { stageBody<kStPrepBT, 0>(a, xcdTile()); }
template <bool NARROW>
__global__ __launch_bounds__(256) void kCensusCost(StageArgs a)
{
    stageBody<kStCensusCost, NARROW ? 1 : 0>(a, xcdTile());
}
template <bool STAGED, bool NARROW, int ROWS>
__global__ __launch_bounds__(256) void kCensusLocal(StageArgs a)
{
    stageBody<kStCensusLocal, censusLocalVariant(STAGED, NARROW, ROWS)>(a, xcdTile());
}
__global__ __launch_bounds__(256) void kBTRow(StageArgs a) { stageBody<kStBTRow, 0>(a, hwTile()); }
__global__ __launch_bounds__(256) void kBoxV(StageArgs a) { stageBody<kStBoxV, 0>(a, hwTile()); }
template <bool CLAMP>
__global__ __launch_bounds__(256) void kBTLocal(StageArgs a)
{
    stageBody<kStBTLocal, CLAMP ? 1 : 0>(a, hwTile());
}
template <int HALF>
__global__ __launch_bounds__(256) void kBTLocalTile(StageArgs a)
{
    stageBody<kStBTLocalTile, HALF>(a, xcdTile());
}
__global__ __launch_bounds__(256) void kPyrDown(StageArgs a) { stageBody<kStPyrDown, 0>(a, xcdTile()); }

template <int KIND, int V>
__global__ __launch_bounds__(256) void kStageQ(StageQueue q)
{
    runQueuedStage(q, [](const StageArgs& a, Tile3 tl) { stageBody<KIND, V>(a, tl); });
}

template <bool STAGED, bool NARROW, int ROWS>
hipError_t launchCensusLocal(const StageLaunch& l, hipStream_t s)
{
    kCensusLocal<STAGED, NARROW, ROWS><<<l.grid, 256, l.lds, s>>>(l.a);
    return hipGetLastError();
}

template <int V>
hipError_t launchQ(int kind, const StageQueue& q, int blocks, size_t lds, hipStream_t s)
{
    switch (kind) {
    case kStCensusCost: kStageQ<kStCensusCost, V & 1><<<blocks, 256, lds, s>>>(q); break;
    case kStCensusLocal: kStageQ<kStCensusLocal, V><<<blocks, 256, lds, s>>>(q); break;
    case kStBTLocal: kStageQ<kStBTLocal, V & 1><<<blocks, 256, lds, s>>>(q); break;
    case kStBTLocalTile: kStageQ<kStBTLocalTile, V & 3><<<blocks, 256, lds, s>>>(q); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Output of a stage launcher: recorded (LevelCall::stageSink) or launched.
hipError_t emit(const LevelCall* c, const StageLaunch& l, hipStream_t s)
{
    if (c && c->stageSink) {
        c->stageSink->launches.push_back(l);
        return hipSuccess;
    }
    return launchStage(l, s);
}

StageLaunch stageOf(int kind, int variant, dim3 grid, size_t lds)
{
    StageLaunch l;
    memset(&l, 0, sizeof(l));
    l.kind = kind;
    l.variant = variant;
    l.grid = grid;
    l.lds = lds;
    return l;
}

} // namespace

hipError_t launchPostStage(const StageLaunch& l, hipStream_t s);
hipError_t launchPostQueued(const StageQueue& q, int blocks, hipStream_t s);

hipError_t launchStage(const StageLaunch& l, hipStream_t s)
{
    const int v = l.variant;
    switch (l.kind) {
    case kStPrepCensus: kPrepCensus<<<l.grid, 256, l.lds, s>>>(l.a); break;
    case kStPrepBT: kPrepBT<<<l.grid, 256, l.lds, s>>>(l.a); break;
    case kStCensusCost:
        if (v) kCensusCost<true><<<l.grid, 256, l.lds, s>>>(l.a);
        else kCensusCost<false><<<l.grid, 256, l.lds, s>>>(l.a);
        break;
    case kStCensusLocal:
        switch (v) {
        case censusLocalVariant(true, true, 2): return launchCensusLocal<true, true, 2>(l, s);
        case censusLocalVariant(true, true, 4): return launchCensusLocal<true, true, 4>(l, s);
        case censusLocalVariant(true, true, 8): return launchCensusLocal<true, true, 8>(l, s);
        case censusLocalVariant(true, false, 2): return launchCensusLocal<true, false, 2>(l, s);
        case censusLocalVariant(true, false, 4): return launchCensusLocal<true, false, 4>(l, s);
        case censusLocalVariant(true, false, 8): return launchCensusLocal<true, false, 8>(l, s);
        case censusLocalVariant(false, true, 2): return launchCensusLocal<false, true, 2>(l, s);
        case censusLocalVariant(false, true, 4): return launchCensusLocal<false, true, 4>(l, s);
        case censusLocalVariant(false, true, 8): return launchCensusLocal<false, true, 8>(l, s);
        case censusLocalVariant(false, false, 2): return launchCensusLocal<false, false, 2>(l, s);
        case censusLocalVariant(false, false, 4): return launchCensusLocal<false, false, 4>(l, s);
        case censusLocalVariant(false, false, 8): return launchCensusLocal<false, false, 8>(l, s);
        default: return hipErrorInvalidValue;
        }
    case kStBTRow: kBTRow<<<l.grid, 256, l.lds, s>>>(l.a); break;
    case kStBoxV: kBoxV<<<l.grid, 256, l.lds, s>>>(l.a); break;
    case kStBTLocal:
        if (v) kBTLocal<true><<<l.grid, 256, l.lds, s>>>(l.a);
        else kBTLocal<false><<<l.grid, 256, l.lds, s>>>(l.a);
        break;
    case kStBTLocalTile:
        switch (v) {
        case 0: kBTLocalTile<0><<<l.grid, 256, l.lds, s>>>(l.a); break;
        case 1: kBTLocalTile<1><<<l.grid, 256, l.lds, s>>>(l.a); break;
        case 2: kBTLocalTile<2><<<l.grid, 256, l.lds, s>>>(l.a); break;
        case 3: kBTLocalTile<3><<<l.grid, 256, l.lds, s>>>(l.a); break;
        default: return hipErrorInvalidValue;
        }
        break;
    case kStPyrDown: kPyrDown<<<l.grid, 256, l.lds, s>>>(l.a); break;
    case kStPost: return launchPostStage(l, s);
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchStageQueued(int kind, int variant, const StageQueue& q, int blocks, size_t lds,
                             hipStream_t s)
{
    if (blocks <= 0) return hipSuccess;
    switch (kind) {
    case kStPrepCensus: kStageQ<kStPrepCensus, 0><<<blocks, 256, lds, s>>>(q); break;
    case kStPrepBT: kStageQ<kStPrepBT, 0><<<blocks, 256, lds, s>>>(q); break;
    case kStBTRow: kStageQ<kStBTRow, 0><<<blocks, 256, lds, s>>>(q); break;
    case kStBoxV: kStageQ<kStBoxV, 0><<<blocks, 256, lds, s>>>(q); break;
    case kStPyrDown: kStageQ<kStPyrDown, 0><<<blocks, 256, lds, s>>>(q); break;
    case kStPost: return launchPostQueued(q, blocks, s);
    default:
        switch (variant) {
        case 0: return launchQ<0>(kind, q, blocks, lds, s);
        case 1: return launchQ<1>(kind, q, blocks, lds, s);
        case 2: return launchQ<2>(kind, q, blocks, lds, s);
        case 3: return launchQ<3>(kind, q, blocks, lds, s);
        case 4: return launchQ<4>(kind, q, blocks, lds, s);
        case 5: return launchQ<5>(kind, q, blocks, lds, s);
        case 6: return launchQ<6>(kind, q, blocks, lds, s);
        case 7: return launchQ<7>(kind, q, blocks, lds, s);
        case 8: return launchQ<8>(kind, q, blocks, lds, s);
        case 9: return launchQ<9>(kind, q, blocks, lds, s);
        case 10: return launchQ<10>(kind, q, blocks, lds, s);
        case 11: return launchQ<11>(kind, q, blocks, lds, s);
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launchPrep(const LevelCall& c, Workspace& ws, hipStream_t s)
{
    const bool census = c.cost == 1;
    StageLaunch l = census
        ? stageOf(kStPrepCensus, 0, dim3((c.W + 63) / 64, (c.H + kCenRows - 1) / kCenRows, 2 * c.batch.n), 0)
        : stageOf(kStPrepBT, 0, dim3((c.W + 63) / 64, (c.H + 3) / 4, 2 * c.batch.n), 0);
    StageArgs& a = l.a;
    a.left = c.left;
    a.lstep = c.lstep;
    a.right = c.right;
    a.rstep = c.rstep;
    a.desc = ws.desc;
    a.keys = ws.keys;
    a.W = c.W;
    a.H = c.H;
    a.cap = c.cap;
    a.batch = c.batch;
    return emit(&c, l, s);
}

// Right census words per row a refinement block may stage in LDS (4 rows: 32 KB).
constexpr long kMaxStaged = 2048;

hipError_t launchCost(const LevelCall& c, Workspace& ws, hipStream_t s)
{
    const int tpp = c.Dp / 8;
    const int npx = 256 / tpp;
    const int half = c.block / 2;
    StageArgs a;
    memset(&a, 0, sizeof(a));
    a.desc = ws.desc;
    a.labels = ws.labels;
    a.C = ws.C;
    a.Hs = ws.Alr;
    a.left = c.left;
    a.lstep = c.lstep;
    a.W = c.W;
    a.H = c.H;
    a.D = c.D;
    a.Dp = c.Dp;
    a.minD = c.minD;
    a.maxD = c.maxD;
    a.half = half;
    a.geom = c.mode == kCoarse;
    a.batch = c.batch;
    auto go = [&](int kind, int variant, dim3 grid, size_t lds) {
        StageLaunch l = stageOf(kind, variant, grid, lds);
        l.a = a;
        return emit(&c, l, s);
    };
    if (c.mode == kRefine) {
        a.r = RefineIn{c.pdisp, c.pconf, c.pdpitch, c.pcpitch, c.PW, c.PH, c.minConf};
        const dim3 grid((c.W + npx - 1) / npx, c.H, c.batch.n);
        const long span = (long)npx + c.maxD - c.minD;  // staged right columns per row
        if (c.cost == 1) {
            const int rows = c.tiles.localRows ? c.tiles.localRows : 4;
            const bool staged = span <= kMaxStaged;
            const dim3 g(grid.x, (c.H + rows - 1) / rows, grid.z);
            return go(kStCensusLocal, censusLocalVariant(staged, c.narrow != 0, rows), g,
                      staged ? (size_t)span * 4 * rows : 0);
        }
        if (((c.cap > 15 ? c.cap : 15) | 1) <= 95 && c.Dp == 16 && c.D == 16 && half <= 3)
            return go(kStBTLocalTile, half,
                      dim3((c.W + kTilePx - 1) / kTilePx, (c.H + kTileRows - 1) / kTileRows, c.batch.n), 0);
        // Clamped sums when a cost may reach 255.
        return go(kStBTLocal, ((c.cap > 15 ? c.cap : 15) | 1) > 95 ? 1 : 0, grid, 0);
    }
    if (c.cost == 1)
        return go(kStCensusCost, c.narrow ? 1 : 0, dim3((c.W + npx - 1) / npx, c.H, c.batch.n),
                  (size_t)(npx + c.Dp - 1) * sizeof(uint32_t));
    const int dc = c.Dp < kChunkD ? c.Dp : kChunkD;
    const int ext = kSegX + 2 * half;
    const size_t lds = (size_t)((ext * dc + 15) & ~15) + (size_t)(2 * ext + dc) * sizeof(BTPacked);
    hipError_t e = go(kStBTRow, 0, dim3((c.W + kSegX - 1) / kSegX, c.H, (c.Dp / dc) * c.batch.n), lds);
    if (e != hipSuccess) return e;
    const size_t nOct = (size_t)c.W * c.H * (c.Dp / 8);
    return go(kStBoxV, 0, dim3((unsigned)((nOct + 255) / 256), c.batch.n), 0);
}

hipError_t launchPyrDownBatch(const uint8_t* srcA, int sstepA, const uint8_t* srcB, int sstepB,
                              size_t sStride, int sw, int sh, uint8_t* dstA, int dstepA,
                              uint8_t* dstB, int dstepB, size_t dStride, int dw, int dh, int n,
                              hipStream_t stream, StageSink* sink)
{
    const int nv = srcB ? 2 : 1;
    StageLaunch l = stageOf(kStPyrDown, 0, dim3((dw + 63) / 64, (dh + kPyrRows - 1) / kPyrRows, nv * n), 0);
    StageArgs& a = l.a;
    a.left = srcA;
    a.lstep = sstepA;
    a.right = srcB;
    a.rstep = sstepB;
    a.sStride = sStride;
    a.sw = sw;
    a.sh = sh;
    a.dstA = dstA;
    a.dstepA = dstepA;
    a.dstB = dstB;
    a.dstepB = dstepB;
    a.dStride = dStride;
    a.W = dw;
    a.H = dh;
    a.nv = nv;
    if (sink) {
        sink->launches.push_back(l);
        return hipSuccess;
    }
    return launchStage(l, stream);
}

hipError_t launchPyrDown2(const uint8_t* srcA, int sstepA, const uint8_t* srcB, int sstepB,
                          int sw, int sh, uint8_t* dstA, int dstepA, uint8_t* dstB, int dstepB,
                          int dw, int dh, hipStream_t stream)
{
    return launchPyrDownBatch(srcA, sstepA, srcB, sstepB, 0, sw, sh, dstA, dstepA, dstB, dstepB,
                              0, dw, dh, 1, stream, nullptr);
}

hipError_t launchPyrDown(const uint8_t* src, int sstep, int sw, int sh, uint8_t* dst, int dstep,
                         int dw, int dh, hipStream_t stream)
{
    return launchPyrDown2(src, sstep, nullptr, 0, sw, sh, dst, dstep, nullptr, 0, dw, dh, stream);
}

} // namespace pyr
