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
// addressed through a buffer resource whose per-step offset is a scalar
// register, so a step spends no vector instructions on addressing.  The next
// U steps are loaded while the current U are processed (two register buffers
// used in turn).  Lanes of lines past the image edge get an offset beyond the
// buffer's range, so their loads return 0 and their stores are dropped.
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

template <int K>
struct StepIn {
    uint32_t c[(K + 1) / 2];
    int lw;  // refinement label word
    int im;  // left-image intensity (adaptive P2)
};

// Uniform (per wave, per chunk) addressing of one pixel-ordered stream.
template <int U>
struct ChunkAddr {
    uint64_t origin;  // byte offset of the chunk's lowest-address pixel on line0
    uint32_t soff[U];
    uint32_t nr;      // bytes reachable from origin by the wave's present lines
};

// MODE: kSingle, kCoarse or kRefine.  EXACT (pyramid only): keep the
// reference's explicit "no active path" flag and SHRT_MAX tests, needed when
// Cmax + 2 * max(P2') could reach SHRT_MAX.
template <int G, int K, int MODE, bool ADAPT, bool EXACT, int DIR>
__device__ __forceinline__ void runLine(const PathArgs& a, int lineGroup, const int* s_lut,
                                        int* s_xchg)
{
    constexpr int LPW = 64 / G;
    constexpr int GK = G * K;
    constexpr int U = prefetchSteps<K>();
    constexpr bool VERT = DIR == kTB;
    constexpr bool RL = DIR == kRL;
    constexpr bool PYR = MODE != kSingle;
    constexpr bool REFINE = MODE == kRefine;
    constexpr int VB = GK * 2;  // bytes of one pixel's label vector

    const int lane = threadIdx.x;
    const int j = lane / G;
    const int g = lane - j * G;
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
    const uint32_t volRow = (uint32_t)W * VB, labRow = (uint32_t)W * 4, imRow = (uint32_t)a.istep;
    auto lineStride = [&](uint32_t rowPitch, uint32_t elem) { return VERT ? elem : rowPitch; };
    auto stepStride = [&](uint32_t rowPitch, uint32_t elem) { return VERT ? rowPitch : elem; };
    const uint32_t vVol = lineOk ? (uint32_t)j * lineStride(volRow, VB) + (uint32_t)g * K * 2 : kOffLimit;
    const uint32_t vLab = lineOk ? (uint32_t)j * lineStride(labRow, 4) : kOffLimit;
    const uint32_t vImg = lineOk ? (uint32_t)j * lineStride(imRow, 1) : kOffLimit;
    int16_t* const out = DIR == kLR ? a.Alr : (DIR == kRL ? a.Arl : a.Atb);

    auto chunk = [&](int sc, uint32_t rowPitch, uint32_t elem, uint32_t vecBytes) {
        ChunkAddr<U> c;
        const int sLast = imin(sc + U - 1, s1 - 1);
        const int so = RL ? sLast : sc;
        const int row = VERT ? so : line0;
        const int col = VERT ? line0 : (RL ? W - 1 - so : so);
        c.origin = (uint64_t)row * rowPitch + (uint64_t)col * elem;
        const uint32_t st = stepStride(rowPitch, elem);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int su = imin(sc + u, s1 - 1);
            c.soff[u] = (uint32_t)(RL ? so - su : su - so) * st;
        }
        c.nr = (uint32_t)(nValid - 1) * lineStride(rowPitch, elem) + (uint32_t)(sLast - sc) * st + vecBytes;
        return c;
    };

    auto issue = [&](StepIn<K>(&buf)[U], int sc) {
        const ChunkAddr<U> cv = chunk(sc, volRow, VB, VB);
        const auto rC = rsrc(reinterpret_cast<const char*>(a.C) + cv.origin, cv.nr);
#pragma unroll
        for (int u = 0; u < U; ++u) bload<K>(buf[u].c, rC, vVol, cv.soff[u]);
        if constexpr (REFINE) {
            const ChunkAddr<U> cl = chunk(sc, labRow, 4, 4);
            const auto r = rsrc(reinterpret_cast<const char*>(a.labels) + cl.origin, cl.nr);
#pragma unroll
            for (int u = 0; u < U; ++u)
                buf[u].lw = (int)__builtin_amdgcn_raw_buffer_load_b32(r, vLab, cl.soff[u], 0);
        }
        if constexpr (ADAPT) {
            const ChunkAddr<U> ci = chunk(sc, imRow, 1, 1);
            const auto r = rsrc(a.img + ci.origin, ci.nr);
#pragma unroll
            for (int u = 0; u < U; ++u)
                buf[u].im = (int)__builtin_amdgcn_raw_buffer_load_b8(r, vImg, ci.soff[u], 0);
        }
    };

    // Path state of this lane's line.  Fast form (EXACT == false): a line
    // without an active path holds SHRT_MAX everywhere, so the next step's
    // best is SHRT_MAX and min(MAX, C + best - minPrev) = C.  Exact form: a
    // line without an active path holds zeros and an explicit flag.
    int Lp[K];
#pragma unroll
    for (int k = 0; k < K; ++k) Lp[k] = EXACT ? 0 : kMax;
    int minPrev = EXACT ? 0 : kMax;
    int prevBase = a.minD;
    bool active = false;
    int prevIm = 0;
    const int P1 = a.P1;

    auto step = [&](const StepIn<K>& st, int p2, bool moved, __amdgpu_buffer_rsrc_t rOut,
                    uint32_t soffOut) {
        int c[K];
#pragma unroll
        for (int k = 0; k < K; ++k) c[k] = unpk<K>(st.c, k);
        int Lc[K];
        if constexpr (!PYR) {
            // L = min(MAX, C + min(m3, min(MAX, minPrev + P2')) - minPrev) with
            // m3 = min(L[d], L[d-1]+P1, L[d+1]+P1) >= minPrev, split into
            // min3(C + m3 - minPrev, C + min(MAX - minPrev, P2'), MAX): no select,
            // and padding labels (C = SHRT_MAX) stay SHRT_MAX.
            const int lmS = from_lower<G>(Lp[K - 1], kMax, g);
            const int lpS = from_upper<G>(Lp[0], kMax, g);
            const int cm = 0 - minPrev;
            const int q = imin(kMax - minPrev, p2);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int Lm = k > 0 ? Lp[k - 1] : lmS;
                const int Lq = k < K - 1 ? Lp[k + 1] : lpS;
                const int m3 = imin(imin(Lm + P1, Lq + P1), Lp[k]);
                Lc[k] = imin(imin(c[k] + cm + m3, c[k] + q), kMax);
            }
        } else {
            int l0[K], lm[K], lp[K];
            if constexpr (REFINE) {
                const int base = (int)(int16_t)(st.lw & 0xFFFF);
                const int sh = base - prevBase;
                prevBase = base;
                if (__builtin_expect(moved, 0)) {
                    // Some line's label window moved: re-index its previous costs.
                    const int she = (!EXACT || active) ? sh : 0;
                    if constexpr (K == 1) {
                        const int sm = g + she;
                        const int lane0 = j * G;
                        const int i0 = lane0 + imin(imax(sm, 0), D - 1);
                        const int im = lane0 + imin(imax(sm - 1, 0), D - 1);
                        const int ip = lane0 + imin(imax(sm + 1, 0), D - 1);
                        const int v0 = __builtin_amdgcn_ds_bpermute(i0 << 2, Lp[0]);
                        const int vm = __builtin_amdgcn_ds_bpermute(im << 2, Lp[0]);
                        const int vp = __builtin_amdgcn_ds_bpermute(ip << 2, Lp[0]);
                        l0[0] = (sm >= 0 && sm < D) ? v0 : kMax;
                        lm[0] = ((sm >= 1 && sm <= D) ? vm : kMax) + P1;
                        lp[0] = ((sm >= -1 && sm < D - 1) ? vp : kMax) + P1;
                    } else {
                        __syncthreads();
#pragma unroll
                        for (int k = 0; k < K; ++k) s_xchg[lane * K + k] = Lp[k];
                        __syncthreads();
                        const int* lineX = s_xchg + j * GK;
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            const int sm = g * K + k + she;
                            const int v0 = lineX[imin(imax(sm, 0), D - 1)];
                            const int vm = lineX[imin(imax(sm - 1, 0), D - 1)];
                            const int vp = lineX[imin(imax(sm + 1, 0), D - 1)];
                            l0[k] = (sm >= 0 && sm < D) ? v0 : kMax;
                            lm[k] = ((sm >= 1 && sm <= D) ? vm : kMax) + P1;
                            lp[k] = ((sm >= -1 && sm < D - 1) ? vp : kMax) + P1;
                        }
                    }
                }
            }
            if (!REFINE || !moved) {
                const int lmS = from_lower<G>(Lp[K - 1], kMax, g);
                const int lpS = from_upper<G>(Lp[0], kMax, g);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    l0[k] = Lp[k];
                    lm[k] = (k > 0 ? Lp[k - 1] : lmS) + P1;
                    lp[k] = (k < K - 1 ? Lp[k + 1] : lpS) + P1;
                }
            }
            const int cm = 0 - minPrev;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int m3 = imin(imin(l0[k], lm[k]), lp[k]);
                if constexpr (EXACT) {
                    const int best = imin(m3, imin(kMax, minPrev + p2));
                    int rec = best < kMax ? imin(kMax, c[k] + cm + best) : kMax;
                    rec = active ? rec : c[k];
                    // Label valid <=> its cost is below SHRT_MAX (the cost kernels
                    // store SHRT_MAX for every geometry-invalid label).
                    Lc[k] = c[k] < kMax ? rec : kMax;
                } else {
                    // Fast form: on active lines minPrev + P2' < SHRT_MAX (host-checked)
                    // so best < SHRT_MAX and L = min(MAX, C + min(m3, minPrev+P2') - minPrev)
                    // = min3(C + m3 - minPrev, C + P2', MAX); m3 >= minPrev, so an invalid
                    // label (C = SHRT_MAX) yields SHRT_MAX, and an inactive line
                    // (all SHRT_MAX, minPrev = SHRT_MAX) yields C.
                    Lc[k] = imin(imin(c[k] + cm + m3, c[k] + p2), kMax);
                }
            }
        }

        int m = Lc[0];
#pragma unroll
        for (int k = 1; k < K; ++k) m = imin(m, Lc[k]);
        const int minCur = gmin<G>(m);
        bstore<K>(Lc, rOut, vVol, soffOut);

        if constexpr (PYR && EXACT) {
            active = minCur < kMax;
#pragma unroll
            for (int k = 0; k < K; ++k) Lp[k] = active ? Lc[k] : 0;
            minPrev = active ? minCur : 0;
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) Lp[k] = Lc[k];
            minPrev = minCur;
        }
    };

    // Process one prefetched chunk [sc, sc+U).
    int lastBase = a.minD;
    auto process = [&](const StepIn<K>(&buf)[U], int sc) {
        int p2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (ADAPT) {
                p2[u] = s_lut[abs(buf[u].im - (u ? buf[u - 1].im : prevIm))];
            } else {
                p2[u] = a.P2;
            }
        }
        if constexpr (ADAPT) prevIm = buf[U - 1].im;
        // Steps at which any line's label window moves, known before the chain runs.
        uint32_t moveMask = 0;
        if constexpr (REFINE) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = (int)(int16_t)(buf[u].lw & 0xFFFF);
                if (__ballot(b != (u ? (int)(int16_t)(buf[u - 1].lw & 0xFFFF) : lastBase)) != 0)
                    moveMask |= 1u << u;
            }
            lastBase = (int)(int16_t)(buf[U - 1].lw & 0xFFFF);
        }
        const ChunkAddr<U> co = chunk(sc, volRow, VB, VB);
        const auto rOut = rsrc(reinterpret_cast<char*>(out) + co.origin, co.nr);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (sc + u < s1) step(buf[u], p2[u], (moveMask >> u) & 1u, rOut, co.soff[u]);
        }
    };

    StepIn<K> bufA[U], bufB[U];
    issue(bufA, s0);
    for (int sc = s0; sc < s1; sc += 2 * U) {
        if (sc + U < s1) issue(bufB, sc + U);
        process(bufA, sc);
        if (sc + U >= s1) break;
        if (sc + 2 * U < s1) issue(bufA, sc + 2 * U);
        process(bufB, sc + U);
    }
}

template <int G, int K, int MODE, bool ADAPT, bool EXACT>
__global__ __launch_bounds__(64) void kPaths(PathArgs a)
{
    __shared__ int s_lut[256];
    __shared__ int s_xchg[MODE == kRefine ? 64 * K : 1];
    if constexpr (ADAPT) {
        for (int i = threadIdx.x; i < 256; i += 64) s_lut[i] = a.lut[i];
        __syncthreads();
    }
    const int b = blockIdx.x;
    if (b < a.nLRBlocks)
        runLine<G, K, MODE, ADAPT, EXACT, kLR>(a, b, s_lut, s_xchg);
    else if (b < a.nLRBlocks + a.nRLBlocks)
        runLine<G, K, MODE, ADAPT, EXACT, kRL>(a, b - a.nLRBlocks, s_lut, s_xchg);
    else
        runLine<G, K, MODE, ADAPT, EXACT, kTB>(a, b - a.nLRBlocks - a.nRLBlocks, s_lut, s_xchg);
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
