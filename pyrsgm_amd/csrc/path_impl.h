// path_impl.h -- SGM path recurrences for gfx950 (included by kernels_path_*.hip).
//
// One wave64 runs 64/G independent lines (rows for L->R and R->L, columns for
// T->B); a line is a G-lane group holding K labels per lane (label d = g*K+k).
// One pixel step: take the previous pixel's path costs for d, d-1, d+1 (own
// registers plus one DPP shift each way, or an LDS exchange when a refinement
// line's label window moved), add the penalties, and reduce the minimum over
// the labels with four DPP min instructions (plus permlane swaps for 32/64-lane
// groups).  All three directions run in one launch and each writes its own
// path-cost volume; winner-take-all is a separate, fully parallel kernel.
//
// Memory: every stream (costs, label words, left image, path-cost output) is
// addressed through a buffer resource whose per-step offset is an instruction
// immediate or a scalar register, so a step spends no vector instructions on
// addressing.  Costs are loaded two chunks of U steps ahead (three register
// buffers in rotation); path costs go to an LDS tile and are written back
// with 16-byte stores once per chunk.  Lanes of lines past the image edge get
// an offset beyond the buffer's range, so their loads return 0 and their
// stores are dropped.
//
// Semantics restated from the reference: single scale sgm_kernels.cu:520-578
// (L->R), :587-638 (T->B), :665-735 (R->L recurrence); pyramid :1191-1261,
// :1263-1334, :1336-1404.  Two equivalences keep the step short:
//  * The host rejects P1 < 0 and P2 < 0.  With P1 >= 0 the reference's
//    "L_prev[d+-1] < SHRT_MAX ? L_prev[d+-1] + P1 : SHRT_MAX" gating and its
//    d == 0 / d == D-1 edges reduce to "fill SHRT_MAX, then add P1": such an
//    operand is >= SHRT_MAX >= min(SHRT_MAX, minPrev + P2'), so it can only
//    tie a result that is SHRT_MAX anyway.
//  * A line without an active path (first pixel, or the previous pixel had no
//    label with a path) keeps all-zero previous costs and minPrev = 0; the
//    recurrence then yields best = 0 and L = C, which is the reference's
//    "!prevActive -> L = C" (sgm_kernels.cu:1231-1232), and for the single
//    scale paths the initial L = C (sgm_kernels.cu:547, :607, :702).
#pragma once

#include "device_common.h"
#include "internal.h"

namespace pyr {
namespace path {

enum Dir : int { kLR = 0, kTB = 1, kRL = 2 };

constexpr uint32_t kRsrcWord3 = 0x00020000u;
constexpr uint32_t kOffLimit = 0x80000000u;  // offset given to lanes of absent lines

// Pixel steps fetched ahead (compile-time tuning knob): about 32 labels per
// lane in flight by default.
#ifndef PYRSGM_PREFETCH_STEPS
#define PYRSGM_PREFETCH_STEPS 16
#endif
template <int K>
constexpr int prefetchSteps() { return K <= 2 ? PYRSGM_PREFETCH_STEPS : 32 / K; }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             (int)kRsrcWord3);
}

template <int K>
__device__ __forceinline__ void bload(uint32_t (&w)[(K + 1) / 2], __amdgpu_buffer_rsrc_t r,
                                      uint32_t voff, uint32_t soff)
{
    if constexpr (K == 1) {
        w[0] = __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
    } else if constexpr (K == 2) {
        w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    } else if constexpr (K == 4) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        w[0] = v[0];
        w[1] = v[1];
    } else {
#pragma unroll
        for (int i = 0; i < K / 8; ++i) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16 * i, soff, 0);
            w[4 * i + 0] = v[0];
            w[4 * i + 1] = v[1];
            w[4 * i + 2] = v[2];
            w[4 * i + 3] = v[3];
        }
    }
}

template <int K>
__device__ __forceinline__ void bstore(const int (&v)[K], __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                       uint32_t soff)
{
    auto pk = [&](int i) { return (uint32_t)(uint16_t)v[i] | ((uint32_t)v[i + 1] << 16); };
    if constexpr (K == 1) {
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)v[0], r, voff, soff, 0);
    } else if constexpr (K == 2) {
        __builtin_amdgcn_raw_buffer_store_b32(pk(0), r, voff, soff, 0);
    } else if constexpr (K == 4) {
        typedef uint32_t u2 __attribute__((ext_vector_type(2)));
        u2 o;
        o[0] = pk(0);
        o[1] = pk(2);
        __builtin_amdgcn_raw_buffer_store_b64(o, r, voff, soff, 0);
    } else {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int i = 0; i < K / 8; ++i) {
            u4 o;
            o[0] = pk(8 * i);
            o[1] = pk(8 * i + 2);
            o[2] = pk(8 * i + 4);
            o[3] = pk(8 * i + 6);
            __builtin_amdgcn_raw_buffer_store_b128(o, r, voff + 16 * i, soff, 0);
        }
    }
}

// K labels of one lane into the LDS tile (int16 each).
template <int K>
__device__ __forceinline__ void tstore(const int (&v)[K], char* p)
{
    auto pk = [&](int i) { return (uint32_t)(uint16_t)v[i] | ((uint32_t)v[i + 1] << 16); };
    if constexpr (K == 1) {
        *reinterpret_cast<uint16_t*>(p) = (uint16_t)v[0];
    } else if constexpr (K == 2) {
        *reinterpret_cast<uint32_t*>(p) = pk(0);
    } else if constexpr (K == 4) {
        typedef uint32_t u2 __attribute__((ext_vector_type(2)));
        u2 o;
        o[0] = pk(0);
        o[1] = pk(2);
        *reinterpret_cast<u2*>(p) = o;
    } else {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int i = 0; i < K / 8; ++i) {
            u4 o;
            o[0] = pk(8 * i);
            o[1] = pk(8 * i + 2);
            o[2] = pk(8 * i + 4);
            o[3] = pk(8 * i + 6);
            reinterpret_cast<u4*>(p)[i] = o;
        }
    }
}

// Costs and path costs are in [0, SHRT_MAX]: zero extension is exact.
template <int K>
__device__ __forceinline__ int unpk(const uint32_t (&w)[(K + 1) / 2], int k)
{
    const uint32_t x = w[k >> 1];
    return (int)((k & 1) ? (x >> 16) : (x & 0xFFFFu));
}

// Row-internal DPP moves without a fill value fold into the consuming
// instruction (v_min_i32_dpp).
template <int CTRL>
__device__ __forceinline__ int dmov(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}

// Minimum over each group of G consecutive lanes (G in 1..64, power of two),
// returned to every lane of the group.
template <int G>
__device__ __forceinline__ int gmin(int v)
{
    if constexpr (G >= 2) v = imin(v, dmov<kDppQuadXor1>(v));
    if constexpr (G >= 4) v = imin(v, dmov<kDppQuadXor2>(v));
    if constexpr (G >= 8) v = imin(v, dmov<kDppRowHalfMirror>(v));
    if constexpr (G >= 16) v = imin(v, dmov<kDppRowMirror>(v));
    if constexpr (G >= 32) {
        const auto s = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
        v = imin((int)s[0], (int)s[1]);
    }
    if constexpr (G >= 64) {
        const auto s = __builtin_amdgcn_permlane32_swap((uint32_t)v, (uint32_t)v, false, false);
        v = imin((int)s[0], (int)s[1]);
    }
    return v;
}

template <int G>
__device__ __forceinline__ uint32_t gumin(uint32_t v)
{
    if constexpr (G >= 2) v = umin(v, (uint32_t)dmov<kDppQuadXor1>((int)v));
    if constexpr (G >= 4) v = umin(v, (uint32_t)dmov<kDppQuadXor2>((int)v));
    if constexpr (G >= 8) v = umin(v, (uint32_t)dmov<kDppRowHalfMirror>((int)v));
    if constexpr (G >= 16) v = umin(v, (uint32_t)dmov<kDppRowMirror>((int)v));
    if constexpr (G >= 32) {
        const auto s = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = umin(s[0], s[1]);
    }
    if constexpr (G >= 64) {
        const auto s = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        v = umin(s[0], s[1]);
    }
    return v;
}

// Maximum over each group of G consecutive lanes (G in 16..64), returned to
// every lane of the group.  Inputs are >= 0, so the zero that bound_ctrl feeds
// to out-of-row lanes never wins.
template <int G>
__device__ __forceinline__ int gmax(int v)
{
    v = imax(v, dmov<kDppQuadXor1>(v));
    v = imax(v, dmov<kDppQuadXor2>(v));
    v = imax(v, dmov<kDppRowHalfMirror>(v));
    v = imax(v, dmov<kDppRowMirror>(v));
    if constexpr (G >= 32) {
        const auto s = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
        v = imax((int)s[0], (int)s[1]);
    }
    if constexpr (G >= 64) {
        const auto s = __builtin_amdgcn_permlane32_swap((uint32_t)v, (uint32_t)v, false, false);
        v = imax((int)s[0], (int)s[1]);
    }
    return v;
}

// Lane n of each 16-lane row, broadcast to the row (DPP row_newbcast).
template <int N>
__device__ __forceinline__ int rowBcast(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + N, 0xF, 0xF, true);
}

template <int N>
struct IntC { static constexpr int value = N; };

// f(IntC<0>{}), ..., f(IntC<N-1>{}): a compile-time step index for each step.
template <int N, int I = 0, class F>
__device__ __forceinline__ void unrollSteps(F& f)
{
    if constexpr (I < N) {
        f(IntC<I>{});
        unrollSteps<N, I + 1>(f);
    }
}

// Value of the next-lower / next-higher lane of the G-lane group; the group's
// edge lane gets 0 (the complemented "no path" value).
template <int G>
__device__ __forceinline__ int lowerOr0(int v, int g)
{
    if constexpr (G == 16) {
        (void)g;
        return dmov<kDppRowShr1>(v);
    } else if constexpr (G == 64) {
        (void)g;
        return dmov<kDppWaveShr1>(v);
    } else {
        const int t = dmov<kDppWaveShr1>(v);
        return g == 0 ? 0 : t;
    }
}

template <int G>
__device__ __forceinline__ int upperOr0(int v, int g)
{
    if constexpr (G == 16) {
        (void)g;
        return dmov<kDppRowShl1>(v);
    } else if constexpr (G == 64) {
        (void)g;
        return dmov<kDppWaveShl1>(v);
    } else {
        const int t = dmov<kDppWaveShl1>(v);
        return g == G - 1 ? 0 : t;
    }
}

// One chunk of U pixel steps in flight: per-step complemented costs of this
// lane's labels, and (refinement / adaptive P2) one per-step word per lane,
// lane slot r (= lane % 16) holding step sc + r of its line.
template <int K, int U>
struct StepBuf {
    uint32_t c[U][(K + 1) / 2];
    int w;
};

// Uniform (per wave, per chunk) addressing of one pixel-ordered stream.
template <int U>
struct ChunkAddr {
    uint64_t origin;  // byte offset of the chunk's lowest-address pixel on line0
    uint32_t soff[U];
    uint32_t nr;      // bytes reachable from origin by the wave's present lines
    int sLast;        // last step of the chunk inside the line
};

// Volumes are stored complemented: the cost kernels write E = SHRT_MAX - C and
// the path kernels write H = SHRT_MAX - L.  With H the recurrence reads
//   H = max(0, H'[d] + f, max(H'[d-1], H'[d+1]) + f - P1, E - P2'),
//   f = E - max_d H'[d]
// (H' the previous pixel's values), which is the reference's
//   L = min(SHRT_MAX, C + min(L'[d], L'[d+-1] + P1, min(SHRT_MAX, minPrev + P2')) - minPrev)
// term by term: min <-> max, SHRT_MAX <-> 0, minPrev = SHRT_MAX - max H'.
// A lane outside the line reads 0 (= SHRT_MAX) from DPP's bound control, so
// the d = 0 / d = D-1 edges need no fill instructions.  The pyramid levels use
// it whenever Cmax + 2 * max(P2') < SHRT_MAX (host-checked): then an active
// line's best is below SHRT_MAX and the reference's "best < MAX ? ... : MAX"
// and "!prevActive -> L = C" cases reduce to the same expression (a line with
// no active path holds H = 0 everywhere, so H = E).  Otherwise (EXACT) the
// step keeps the reference's explicit tests in the L domain.
//
// Path costs of a chunk are staged in an LDS tile and written back with
// 16-byte stores once per chunk (a 2-byte store per lane and step would cost
// one memory instruction per step).  Tile layout: horizontal paths
// [line][step slot][label] with each line's row padded by one pixel's labels
// (the lines of a wave then write disjoint LDS banks at each step), vertical
// path [step][line][label].  A horizontal line's row and a vertical step's row are
// each one contiguous run of the path-cost volume.
template <int G, int K>
struct Tile {
    static constexpr int U = prefetchSteps<K>();
    static constexpr int LPW = 64 / G;
    static constexpr int VB = G * K * 2;               // bytes of one pixel's labels
    static constexpr int ROWH = U * VB + VB;           // horizontal: padded line row, bytes
    static constexpr int BYTES_H = LPW * ROWH;
    static constexpr int BYTES_V = U * LPW * VB;
    static constexpr int BYTES = BYTES_H > BYTES_V ? BYTES_H : BYTES_V;
    static constexpr int PIECES = LPW * U * VB / 16;   // 16-byte pieces per chunk
    static_assert(PIECES % 64 == 0, "whole pieces per lane");
};

// MODE: kSingle, kCoarse or kRefine.
template <int G, int K, int MODE, bool ADAPT, bool EXACT, int DIR>
__device__ __forceinline__ void runLine(const PathArgs& a, int lineGroup, const int* s_lut,
                                        int* s_xchg, char* s_tile)
{
    constexpr int LPW = 64 / G;
    constexpr int GK = G * K;
    constexpr int U = prefetchSteps<K>();
    static_assert(U <= 16, "per-step words use one 16-lane row slot per step");
    constexpr bool VERT = DIR == kTB;
    constexpr bool RL = DIR == kRL;
    constexpr bool PYR = MODE != kSingle;
    constexpr bool REFINE = MODE == kRefine;
    constexpr bool WORDS = REFINE || ADAPT;  // per-step words are loaded
    constexpr int VB = GK * 2;  // bytes of one pixel's label vector
    using TL = Tile<G, K>;

    const int lane = threadIdx.x;
    const int j = lane / G;
    const int g = lane - j * G;
    const int r16 = lane & 15;
    const int W = a.W, H = a.H, D = a.D;
    const int nLines = VERT ? W : H;
    const int line0 = lineGroup * LPW;
    const int line = line0 + j;
    const bool lineOk = line < nLines;
    const int nValid = imin(LPW, nLines - line0);
    int s0 = 0, s1 = VERT ? H : W;
    if constexpr (!PYR && DIR == kLR) {
        s0 = imax(a.startX, 0);
        s1 = imin(a.endX, W);
    }
    if (s0 >= s1) return;

    // Stream layouts: offset(row, col) = row * rowPitch + col * elem.
    const uint32_t volRow = (uint32_t)W * VB;
    const uint32_t wRow = REFINE ? (uint32_t)W * 4 : (uint32_t)a.istep;
    constexpr uint32_t wElem = REFINE ? 4 : 1;
    auto lineStride = [&](uint32_t rowPitch, uint32_t elem) { return VERT ? elem : rowPitch; };
    auto stepStride = [&](uint32_t rowPitch, uint32_t elem) { return VERT ? rowPitch : elem; };
    const uint32_t vVol = lineOk ? (uint32_t)j * lineStride(volRow, VB) + (uint32_t)g * K * 2 : kOffLimit;
    const uint32_t wLine = (uint32_t)j * lineStride(wRow, wElem);
    const uint32_t wStep = stepStride(wRow, wElem);
    int16_t* const out = DIR == kLR ? a.Alr : (DIR == kRL ? a.Arl : a.Atb);

    // full: the chunk [sc, sc+U) lies inside the line, so every offset is a
    // compile-time multiple of the step stride (an instruction immediate for
    // the horizontal paths).  Otherwise steps past the line re-read its last
    // step; the look-ahead of the last chunk starts at or past s1.
    auto chunk = [&](int scIn, uint32_t rowPitch, uint32_t elem, uint32_t vecBytes, bool full) {
        ChunkAddr<U> c;
        const int sc = full ? scIn : imin(scIn, s1 - 1);
        const int sLast = full ? sc + U - 1 : imin(sc + U - 1, s1 - 1);
        const int so = RL ? sLast : sc;
        const int row = VERT ? so : line0;
        const int col = VERT ? line0 : (RL ? W - 1 - so : so);
        c.origin = (uint64_t)row * rowPitch + (uint64_t)col * elem;
        const uint32_t st = stepStride(rowPitch, elem);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int su = full ? sc + u : imin(sc + u, s1 - 1);
            c.soff[u] = (uint32_t)(RL ? so - su : su - so) * st;
        }
        c.nr = (uint32_t)(nValid - 1) * lineStride(rowPitch, elem) + (uint32_t)(sLast - sc) * st + vecBytes;
        c.sLast = sLast;
        return c;
    };

    auto issue = [&](StepBuf<K, U>& buf, int scIn, bool full) {
        if constexpr (WORDS) {
            // Lane slot r16 loads step sc + r16 of its line; slots past the chunk
            // or the line, and lanes of absent lines, read 0 (offset out of range).
            const ChunkAddr<U> cw = chunk(scIn, wRow, wElem, wElem, full);
            const int sc = full ? scIn : imin(scIn, s1 - 1);
            const int rel = RL ? cw.sLast - sc - r16 : r16;
            const bool ok = lineOk && r16 < U && (full || sc + r16 <= cw.sLast);
            const uint32_t voff = ok ? wLine + (uint32_t)rel * wStep : kOffLimit;
            const char* base = REFINE ? reinterpret_cast<const char*>(a.labels) : reinterpret_cast<const char*>(a.img);
            const auto r = rsrc(base + cw.origin, cw.nr);
            buf.w = REFINE ? (int)__builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0)
                           : (int)__builtin_amdgcn_raw_buffer_load_b8(r, voff, 0, 0);
        }
        const ChunkAddr<U> cv = chunk(scIn, volRow, VB, VB, full);
        const auto rC = rsrc(reinterpret_cast<const char*>(a.C) + cv.origin, cv.nr);
#pragma unroll
        for (int u = 0; u < U; ++u) bload<K>(buf.c[u], rC, vVol, cv.soff[u]);
    };

    // Path state of this lane's labels: complemented H (fast form) or L with
    // the reference's activity flag (EXACT).
    int Hs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) Hs[k] = 0;
    int maxH = 0;                  // fast form: max over the line of Hs
    int minPrev = 0;               // EXACT
    bool active = false;           // EXACT
    int carryBase = a.minD;        // label-window base of the previous step (slot row)
    int carryIm4 = 0;              // 4 * intensity of the previous step
    const int P1 = a.P1;

    // Previous costs of the line re-indexed after its label window moved by
    // sh labels: l0 = prev[d + sh], y = max(prev[d + sh - 1], prev[d + sh + 1]),
    // out-of-window labels reading `none`.
    auto reindex = [&](const int (&prev)[K], int sh, int none, int (&l0)[K], int (&lm)[K],
                       int (&lp)[K]) {
        if constexpr (K == 1) {
            const int sm = g + sh;
            const int lane0 = j * G;
            const int i0 = lane0 + imin(imax(sm, 0), D - 1);
            const int im = lane0 + imin(imax(sm - 1, 0), D - 1);
            const int ip = lane0 + imin(imax(sm + 1, 0), D - 1);
            const int v0 = __builtin_amdgcn_ds_bpermute(i0 << 2, prev[0]);
            const int vm = __builtin_amdgcn_ds_bpermute(im << 2, prev[0]);
            const int vp = __builtin_amdgcn_ds_bpermute(ip << 2, prev[0]);
            l0[0] = (sm >= 0 && sm < D) ? v0 : none;
            lm[0] = (sm >= 1 && sm <= D) ? vm : none;
            lp[0] = (sm >= -1 && sm < D - 1) ? vp : none;
        } else {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < K; ++k) s_xchg[lane * K + k] = prev[k];
            __syncthreads();
            const int* lineX = s_xchg + j * GK;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int sm = g * K + k + sh;
                const int v0 = lineX[imin(imax(sm, 0), D - 1)];
                const int vm = lineX[imin(imax(sm - 1, 0), D - 1)];
                const int vp = lineX[imin(imax(sm + 1, 0), D - 1)];
                l0[k] = (sm >= 0 && sm < D) ? v0 : none;
                lm[k] = (sm >= 1 && sm <= D) ? vm : none;
                lp[k] = (sm >= -1 && sm < D - 1) ? vp : none;
            }
        }
    };

    // Pixel step u of a chunk.  cw: complemented costs; p2: this step's P2'
    // (line-uniform); shv: per-step window shifts of this lane's line (slot u is
    // read on moved steps only); bigMask bit u: some line shifts by more than 1.
    auto step = [&](auto uc, const uint32_t (&cw)[(K + 1) / 2], int p2, bool moved,
                    uint32_t bigMask, int shv, int tileByte) {
        constexpr int u = decltype(uc)::value;
        int e[K];
#pragma unroll
        for (int k = 0; k < K; ++k) e[k] = unpk<K>(cw, k);
        int Hc[K];
        if constexpr (!EXACT) {
            // Neighbours of an unmoved window; a moved one overrides them (one
            // uniform branch per step, taken on the rare moved steps).
            int h0[K], y[K];
            const int lowN = lowerOr0<G>(Hs[K - 1], g);
            const int upN = upperOr0<G>(Hs[0], g);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                h0[k] = Hs[k];
                y[k] = imax(k > 0 ? Hs[k - 1] : lowN, k < K - 1 ? Hs[k + 1] : upN);
            }
            if (REFINE && __builtin_expect(moved, 0)) {
                if (G == 16 && K == 1 && !((bigMask >> u) & 1u)) {
                    // Every line of the wave moved by -1, 0 or +1 labels: the
                    // re-indexed values are row shifts of the previous ones (zero
                    // outside the 16 labels), selected per line.
                    const int sh = rowBcast<u>(shv);
                    const int tm2 = dmov<kDppRowShr2>(Hs[0]), tp2 = dmov<kDppRowShl2>(Hs[0]);
                    const bool neg = sh < 0, pos = sh > 0;
                    h0[0] = neg ? lowN : (pos ? upN : Hs[0]);
                    const int lm = neg ? tm2 : (pos ? Hs[0] : lowN);
                    const int lp = neg ? Hs[0] : (pos ? tp2 : upN);
                    y[0] = imax(lm, lp);
                } else {
                    int lm[K], lp[K];
                    reindex(Hs, rowBcast<u>(shv), 0, h0, lm, lp);
#pragma unroll
                    for (int k = 0; k < K; ++k) y[k] = imax(lm[k], lp[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int f = e[k] - maxH;
                const int t3 = (int)__builtin_elementwise_sub_sat((uint32_t)e[k], (uint32_t)p2);
                Hc[k] = imax(imax(h0[k] + f, y[k] + f - P1), t3);
            }
            int m = Hc[0];
#pragma unroll
            for (int k = 1; k < K; ++k) m = imax(m, Hc[k]);
            maxH = gmax<G>(m);
#pragma unroll
            for (int k = 0; k < K; ++k) Hs[k] = Hc[k];
        } else {
            // L domain; Hs holds L here (zeros while no path is active).
            int l0[K], lm[K], lp[K];
            const int lmS = from_lower<G>(Hs[K - 1], kMax, g);
            const int lpS = from_upper<G>(Hs[0], kMax, g);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                l0[k] = Hs[k];
                lm[k] = k > 0 ? Hs[k - 1] : lmS;
                lp[k] = k < K - 1 ? Hs[k + 1] : lpS;
            }
            if (REFINE && __builtin_expect(moved, 0))
                reindex(Hs, active ? rowBcast<u>(shv) : 0, kMax, l0, lm, lp);
            int Lc[K];
            const int cap = imin(kMax, minPrev + p2);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int c = kMax - e[k];
                const int m3 = imin(imin(l0[k], lm[k] + P1), lp[k] + P1);
                const int best = imin(m3, cap);
                int rec = best < kMax ? imin(kMax, c + best - minPrev) : kMax;
                rec = active ? rec : c;
                // Label valid <=> its cost is below SHRT_MAX (the cost kernels
                // store SHRT_MAX for every geometry-invalid label).
                Lc[k] = c < kMax ? rec : kMax;
                Hc[k] = kMax - Lc[k];
            }
            int m = Lc[0];
#pragma unroll
            for (int k = 1; k < K; ++k) m = imin(m, Lc[k]);
            const int minCur = gmin<G>(m);
            active = minCur < kMax;
#pragma unroll
            for (int k = 0; k < K; ++k) Hs[k] = active ? Lc[k] : 0;
            minPrev = active ? minCur : 0;
        }
        tstore<K>(Hc, s_tile + tileByte);
    };

    // Write the chunk's tile to the path-cost volume: lane l stores 16-byte
    // pieces l, l+64, ...  A piece belongs to one tile row (a line of a
    // horizontal path, a step of the vertical one), which is one contiguous run
    // of the volume; pieces of absent lines or steps past the line get an
    // out-of-range offset and are dropped.
    auto flush = [&](int sc, int nSteps, bool full) {
        const ChunkAddr<U> co = chunk(sc, volRow, VB, VB, full);
        const auto rOut = rsrc(reinterpret_cast<char*>(out) + co.origin, co.nr);
        constexpr int PPR = VERT ? LPW * VB / 16 : U * VB / 16;  // pieces per tile row
        const int nRows = VERT ? nSteps : nValid;
        const int nCols = VERT ? nValid : nSteps;                // pixels present in a row
#pragma unroll
        for (int i = 0; i < TL::PIECES / 64; ++i) {
            const int pc = lane + 64 * i;
            const int row = pc / PPR;
            const int ob = (pc - row * PPR) * 16;                 // byte offset inside the row
            const int tbyte = row * (VERT ? LPW * VB : TL::ROWH) + ob;
            typedef uint32_t u4 __attribute__((ext_vector_type(4)));
            const u4 v = *reinterpret_cast<const u4*>(s_tile + tbyte);
            const bool ok = row < nRows && ob / VB < nCols;
            const uint32_t go = ok ? (uint32_t)row * volRow + (uint32_t)ob : kOffLimit;
            __builtin_amdgcn_raw_buffer_store_b128(v, rOut, go, 0, 0);
        }
    };

    // Process one prefetched chunk [sc, sc+U); full: all U steps exist.
    auto process = [&](const StepBuf<K, U>& buf, int sc, bool full) {
        int p2v = 0, shv = 0;
        uint32_t moveMask = 0, bigMask = 0;
        if constexpr (REFINE) {
            const int base = (int)(int16_t)(buf.w & 0xFFFF);
            const int prevBase = dpp<kDppRowShr1>(carryBase, base);
            shv = base - prevBase;
            // Bit u: some line of the wave moves at step u (any shift / |shift| > 1).
            auto fold = [](uint64_t b) {
                const uint64_t f2 = b | (b >> 32);
                return (uint32_t)(f2 | (f2 >> 16)) & 0xFFFFu;
            };
            moveMask = fold(__ballot(shv != 0));
            bigMask = fold(__ballot(shv > 1 || shv < -1));
            carryBase = rowBcast<U - 1>(base);
        }
        if constexpr (ADAPT) {
            const int im4 = REFINE ? (int)((uint32_t)buf.w >> 22) : buf.w << 2;
            const int prevIm4 = dpp<kDppRowShr1>(carryIm4, im4);
            const uint32_t d4 = __builtin_amdgcn_sad_u16((uint32_t)im4, (uint32_t)prevIm4, 0u);
            p2v = *reinterpret_cast<const int*>(reinterpret_cast<const char*>(s_lut) + d4);
            carryIm4 = rowBcast<U - 1>(im4);
        }
        const int nSteps = full ? U : imin(U, s1 - sc);
        auto run = [&](auto uc) {
            constexpr int u = decltype(uc)::value;
            if (full || u < nSteps) {
                const int p2 = ADAPT ? rowBcast<u>(p2v) : a.P2;
                const bool moved = REFINE && ((moveMask >> u) & 1u);
                // Tile slot of step u: horizontal rows run in address order (R->L
                // reversed), the vertical tile is step-major.
                int tb;
                if constexpr (VERT)
                    tb = (u * LPW + j) * VB + g * K * 2;
                else
                    tb = j * TL::ROWH + (RL ? (full ? U - 1 - u : nSteps - 1 - u) : u) * VB + g * K * 2;
                step(uc, buf.c[u], p2, moved, bigMask, shv, tb);
            }
        };
        unrollSteps<U>(run);
    };

    // Three register buffers in rotation: chunk k's loads are issued before
    // chunk k-2 is processed, so two chunks of steps hide their latency, and
    // each chunk's tile is written back right after its steps.  Triples of full
    // chunks run with no conditional loads, stores or steps (a guarded path
    // makes the compiler's wait counts drain the look-ahead: vmcnt retires
    // loads and stores together, in issue order); the remaining steps follow
    // one chunk at a time.
    StepBuf<K, U> bufA, bufB, bufC;
    int sc = s0;
    issue(bufA, sc, sc + U <= s1);
    issue(bufB, sc + U, sc + 2 * U <= s1);
    for (; sc + 5 * U <= s1; sc += 3 * U) {
        issue(bufC, sc + 2 * U, true);
        process(bufA, sc, true);
        flush(sc, U, true);
        issue(bufA, sc + 3 * U, true);
        process(bufB, sc + U, true);
        flush(sc + U, U, true);
        issue(bufB, sc + 4 * U, true);
        process(bufC, sc + 2 * U, true);
        flush(sc + 2 * U, U, true);
    }
    for (; sc < s1; sc += U) {
        issue(bufC, sc + 2 * U, false);
        process(bufA, sc, false);
        flush(sc, imin(U, s1 - sc), false);
        bufA = bufB;
        bufB = bufC;
    }
}

template <int G, int K, int MODE, bool ADAPT, bool EXACT>
__global__ __launch_bounds__(64) void kPaths(PathArgs a)
{
    __shared__ int s_lut[256];
    __shared__ int s_xchg[MODE == kRefine ? 64 * K : 1];
    __shared__ __attribute__((aligned(16))) char s_tile[Tile<G, K>::BYTES];
    if constexpr (ADAPT) {
        for (int i = threadIdx.x; i < 256; i += 64) s_lut[i] = a.lut[i];
        __syncthreads();
    }
    const int b = blockIdx.x;
    if (b < a.nLRBlocks)
        runLine<G, K, MODE, ADAPT, EXACT, kLR>(a, b, s_lut, s_xchg, s_tile);
    else if (b < a.nLRBlocks + a.nRLBlocks)
        runLine<G, K, MODE, ADAPT, EXACT, kRL>(a, b - a.nLRBlocks, s_lut, s_xchg, s_tile);
    else
        runLine<G, K, MODE, ADAPT, EXACT, kTB>(a, b - a.nLRBlocks - a.nRLBlocks, s_lut, s_xchg, s_tile);
}

inline int linesPerWave(int Dp) { return Dp <= 16 ? 4 : (Dp <= 32 ? 2 : 1); }

template <int MODE, bool ADAPT, bool EXACT>
hipError_t dispatchPaths(const PathArgs& a, hipStream_t s)
{
    const int lpw = linesPerWave(a.Dp);
    PathArgs b = a;
    b.nLRBlocks = a.numPaths >= 2 ? (a.H + lpw - 1) / lpw : 0;
    b.nRLBlocks = (a.H + lpw - 1) / lpw;
    const int nTB = a.numPaths >= 3 ? (a.W + lpw - 1) / lpw : 0;
    const int blocks = b.nLRBlocks + b.nRLBlocks + nTB;
    switch (a.Dp) {
    case 16: kPaths<16, 1, MODE, ADAPT, EXACT><<<blocks, 64, 0, s>>>(b); break;
    case 32: kPaths<32, 1, MODE, ADAPT, EXACT><<<blocks, 64, 0, s>>>(b); break;
    case 64: kPaths<64, 1, MODE, ADAPT, EXACT><<<blocks, 64, 0, s>>>(b); break;
    case 128: kPaths<64, 2, MODE, ADAPT, EXACT><<<blocks, 64, 0, s>>>(b); break;
    case 256: kPaths<64, 4, MODE, ADAPT, EXACT><<<blocks, 64, 0, s>>>(b); break;
    case 512: kPaths<64, 8, MODE, ADAPT, EXACT><<<blocks, 64, 0, s>>>(b); break;
    case 1024: kPaths<64, 16, MODE, ADAPT, EXACT><<<blocks, 64, 0, s>>>(b); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t launchPathsMode(const PathArgs& a, int adaptive, hipStream_t s)
{
    if (MODE != kSingle && a.exact)
        return adaptive ? dispatchPaths<MODE, true, true>(a, s) : dispatchPaths<MODE, false, true>(a, s);
    return adaptive ? dispatchPaths<MODE, true, false>(a, s) : dispatchPaths<MODE, false, false>(a, s);
}

} // namespace path
} // namespace pyr
