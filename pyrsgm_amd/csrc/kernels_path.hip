// kernels_path.hip -- SGM path aggregation for gfx950.
//
// One wave64 runs 64/G independent lines (rows for L->R / R->L, columns for
// T->B); each line is a G-lane group holding K labels per lane.  A pixel step
// is: take the previous pixel's path costs for d, d-1, d+1 (own registers plus
// one DPP shift each way, or an LDS exchange when a refinement line's label
// window moved), add the penalties, and reduce the minimum over the labels
// with four DPP min instructions (plus permlane swaps for 32/64-lane groups).
//
// Memory: every stream (costs, path costs, label words, left image) is read
// with raw buffer loads whose per-step offset is a scalar register, so a step
// spends no vector instructions on addressing.  The next U steps are in
// flight while the current U are processed (two register buffers, used in
// turn).  Lanes of lines past the image edge get an offset beyond the
// buffer's range, so their loads return 0 and their stores are dropped.
//
// Launches:
//   kPathsLRTB: L->R rows and T->B columns in ONE launch (independent paths);
//               writes Alr and Atb.
//   kPathRL:    R->L rows; adds Alr/Atb with the reference's saturation order
//               and parks each pixel's summed costs in an LDS ring; every 16
//               steps the wave does winner-take-all for the 16*64/G parked
//               pixels, 1..4 lanes per pixel (uniqueness, sub-pixel,
//               confidence, and the right-view projection as an order-free
//               atomicMin key).
//
// Semantics restated from the reference:
//   single scale: sgm_kernels.cu:520-578 (L->R), :587-638 (T->B), :665-819 (R->L+WTA)
//   pyramid:      sgm_kernels.cu:1191-1261, :1263-1334, :1336-1478
// The host rejects P1 < 0 and P2 < 0; with P1 >= 0 the reference's
// "L_prev[d+-1] < SHRT_MAX ? L_prev[d+-1] + P1 : SHRT_MAX" gating and its
// d == 0 / d == D-1 edge cases reduce to "fill SHRT_MAX, then add P1": an
// operand >= SHRT_MAX can never be the strict minimum next to the jump term
// min(SHRT_MAX, minPrev + P2'), and when it ties, the result is SHRT_MAX
// either way.
#include "device_common.h"
#include "internal.h"

namespace pyr {
namespace {

enum Dir : int { kLR = 0, kTB = 1, kRL = 2 };

constexpr int kBatch = 16;          // R->L pixels parked per line before WTA
constexpr uint32_t kRsrcWord3 = 0x00020000u;
constexpr uint32_t kOffLimit = 0x80000000u;  // offset given to lanes of absent lines

// Pixel steps fetched ahead: about 32 labels per lane in flight.
template <int K>
constexpr int prefetchSteps() { return K <= 2 ? 16 : 32 / K; }

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

template <int K>
__device__ __forceinline__ int unpk(const uint32_t (&w)[(K + 1) / 2], int k)
{
    const uint32_t x = w[k >> 1];
    return (k & 1) ? ((int)x >> 16) : (int)(int16_t)(x & 0xFFFFu);
}

// Row-internal DPP moves without a fill value: the compiler folds them into
// the consuming VALU instruction (v_min_i32_dpp).
template <int CTRL>
__device__ __forceinline__ int dmov(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}

template <int G>
__device__ __forceinline__ int gmin(int v)
{
    v = imin(v, dmov<kDppQuadXor1>(v));
    v = imin(v, dmov<kDppQuadXor2>(v));
    v = imin(v, dmov<kDppRowHalfMirror>(v));
    v = imin(v, dmov<kDppRowMirror>(v));
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

template <int K>
struct StepIn {
    uint32_t c[(K + 1) / 2], lr[(K + 1) / 2], tb[(K + 1) / 2];
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

template <int G, int K, bool PYR, bool ADAPT, int DIR>
__device__ __forceinline__ void runLine(const PathArgs& a, int lineGroup, const int* s_lut,
                                        int* s_xchg, int16_t* s_ring, uint32_t* s_meta)
{
    constexpr int LPW = 64 / G;
    constexpr int GK = G * K;
    constexpr int U = prefetchSteps<K>();
    constexpr bool VERT = DIR == kTB;
    constexpr bool RL = DIR == kRL;
    constexpr int RP = GK + 2;   // ring pitch (int16 per parked pixel)
    constexpr int VB = GK * 2;   // bytes of one pixel's label vector

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

    const bool hasLR = RL && a.numPaths >= 2;
    const bool hasTB = RL && a.numPaths >= 3;
    const bool refine = PYR && a.labels != nullptr;

    // Stream layouts: offset(row, col) = row * rowPitch + col * elem.
    const uint32_t volRow = (uint32_t)W * VB, labRow = (uint32_t)W * 4, imRow = (uint32_t)a.istep;
    auto lineStride = [&](uint32_t rowPitch, uint32_t elem) { return VERT ? elem : rowPitch; };
    auto stepStride = [&](uint32_t rowPitch, uint32_t elem) { return VERT ? rowPitch : elem; };
    const uint32_t vVol = lineOk ? (uint32_t)j * lineStride(volRow, VB) + (uint32_t)g * K * 2 : kOffLimit;
    const uint32_t vLab = lineOk ? (uint32_t)j * lineStride(labRow, 4) : kOffLimit;
    const uint32_t vImg = lineOk ? (uint32_t)j * lineStride(imRow, 1) : kOffLimit;

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
        const auto rL = rsrc(reinterpret_cast<const char*>(a.Alr) + cv.origin, cv.nr);
        const auto rT = rsrc(reinterpret_cast<const char*>(a.Atb) + cv.origin, cv.nr);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            bload<K>(buf[u].c, rC, vVol, cv.soff[u]);
            if constexpr (RL) {
                if (hasLR) bload<K>(buf[u].lr, rL, vVol, cv.soff[u]);
                if (hasTB) bload<K>(buf[u].tb, rT, vVol, cv.soff[u]);
            }
        }
        if constexpr (PYR) {
            if (refine) {
                const ChunkAddr<U> cl = chunk(sc, labRow, 4, 4);
                const auto r = rsrc(reinterpret_cast<const char*>(a.labels) + cl.origin, cl.nr);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    buf[u].lw = (int)__builtin_amdgcn_raw_buffer_load_b32(r, vLab, cl.soff[u], 0);
            }
        }
        if constexpr (ADAPT) {
            const ChunkAddr<U> ci = chunk(sc, imRow, 1, 1);
            const auto r = rsrc(a.img + ci.origin, ci.nr);
#pragma unroll
            for (int u = 0; u < U; ++u)
                buf[u].im = (int)__builtin_amdgcn_raw_buffer_load_b8(r, vImg, ci.soff[u], 0);
        }
    };

    // Path state of this lane's line.
    int Lp[K];
#pragma unroll
    for (int k = 0; k < K; ++k) Lp[k] = kMax;
    int minPrev = kMax;
    int prevBase = a.minD;
    bool prevActive = false;
    int prevIm = 0;
    const int P1 = a.P1;

    // Coarse pyramid levels: fixed label window [minD, minD+D).
    auto geomOk = [&](int d, int x) {
        const int real = a.minD + d;
        return d < D && real <= imin(a.maxD, x) && real >= x - W + 1;
    };

    auto step = [&](const StepIn<K>& st, int s, int t, int p2, __amdgpu_buffer_rsrc_t rOut,
                    uint32_t soffOut) {
        int c[K];
#pragma unroll
        for (int k = 0; k < K; ++k) c[k] = unpk<K>(st.c, k);
        int Lc[K];
        bool lv[K];
        int base = a.minD;
        if constexpr (!PYR) {
            if (s == s0) {
#pragma unroll
                for (int k = 0; k < K; ++k) Lc[k] = c[k];
            } else {
                const int lmS = from_lower<G>(Lp[K - 1], kMax, g);
                const int lpS = from_upper<G>(Lp[0], kMax, g);
                const int lj = imin(kMax, minPrev + p2);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int Lm = k > 0 ? Lp[k - 1] : lmS;
                    const int Lq = k < K - 1 ? Lp[k + 1] : lpS;
                    const int best = imin(imin(Lm + P1, Lq + P1), imin(Lp[k], lj));
                    Lc[k] = sat16(c[k] + best - minPrev);
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                lv[k] = g * K + k < D;
                if (GK != 16 && GK != 32 && !lv[k]) Lc[k] = kMax;
            }
        } else {
            if (refine) base = (int)(int16_t)(st.lw & 0xFFFF);
            const int x = VERT ? line : (RL ? W - 1 - s : s);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                // Refinement costs are SHRT_MAX exactly where the label's geometry is invalid.
                lv[k] = c[k] < kMax && (refine || geomOk(g * K + k, x));
            }
            const int sh = base - prevBase;
            int l0[K], lm[K], lp[K];
            if (__ballot(prevActive && sh != 0) != 0) {
                // A line's label window moved: re-index its previous costs through LDS.
                __syncthreads();
#pragma unroll
                for (int k = 0; k < K; ++k) s_xchg[lane * K + k] = Lp[k];
                __syncthreads();
                const int* lineX = s_xchg + j * GK;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int sm = g * K + k + sh;
                    const int v0 = lineX[imin(imax(sm, 0), D - 1)];
                    const int vm = lineX[imin(imax(sm - 1, 0), D - 1)];
                    const int vp = lineX[imin(imax(sm + 1, 0), D - 1)];
                    l0[k] = (sm >= 0 && sm < D) ? v0 : kMax;
                    lm[k] = ((sm >= 1 && sm <= D) ? vm : kMax) + P1;
                    lp[k] = ((sm >= -1 && sm < D - 1) ? vp : kMax) + P1;
                }
            } else {
                const int lmS = from_lower<G>(Lp[K - 1], kMax, g);
                const int lpS = from_upper<G>(Lp[0], kMax, g);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    l0[k] = Lp[k];
                    lm[k] = (k > 0 ? Lp[k - 1] : lmS) + P1;
                    lp[k] = (k < K - 1 ? Lp[k + 1] : lpS) + P1;
                }
            }
            const int lj = imin(kMax, minPrev + p2);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int best = imin(imin(l0[k], lm[k]), imin(lp[k], lj));
                const int rec = best < kMax ? sat16(c[k] + best - minPrev) : kMax;
                Lc[k] = lv[k] ? (prevActive ? rec : c[k]) : kMax;
            }
        }

        int m = Lc[0];
#pragma unroll
        for (int k = 1; k < K; ++k) m = imin(m, Lc[k]);
        const int minCur = gmin<G>(m);

        if constexpr (!RL) {
            bstore<K>(Lc, rOut, vVol, soffOut);
        } else {
            int my[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int lr = hasLR ? unpk<K>(st.lr, k) : 0;
                const int tb = hasTB ? unpk<K>(st.tb, k) : 0;
                if constexpr (!PYR) {
                    const int acc = hasTB ? (int)(int16_t)imin(kMax, lr + tb) : lr;
                    my[k] = (int)(int16_t)imin(kMax, acc + Lc[k]);
                } else {
                    // Path costs are in [0, SHRT_MAX] with SHRT_MAX meaning "no path".
                    const int old = hasTB ? (imax(lr, tb) == kMax ? kMax : imin(kMax - 1, lr + tb)) : lr;
                    my[k] = imax(old, Lc[k]) == kMax ? kMax : imin(kMax - 1, old + Lc[k]);
                }
            }
            int16_t* r = s_ring + (j * kBatch + t) * RP + g * K;
            if constexpr (K == 1) {
                r[0] = (int16_t)my[0];
            } else {
#pragma unroll
                for (int k = 0; k < K; k += 2)
                    *reinterpret_cast<uint32_t*>(r + k) =
                        (uint32_t)(uint16_t)my[k] | ((uint32_t)my[k + 1] << 16);
            }
            if constexpr (PYR) {
                if (g == 0) {
                    const int x = W - 1 - s;
                    const int lo = imax(a.minD, x - W + 1), hi = imin(a.maxD, x);
                    const int valid = refine ? (st.lw >> 16) & 1 : 1;
                    const int fl = imax(0, lo - base), ll = imin(D - 1, hi - base);
                    const int vcount = (valid && ll >= fl) ? ll - fl + 1 : 0;
                    s_meta[j * kBatch + t] = (uint32_t)(base & 0xFFFF) | ((uint32_t)vcount << 16);
                }
            }
        }

#pragma unroll
        for (int k = 0; k < K; ++k) Lp[k] = Lc[k];
        minPrev = minCur;
        if constexpr (PYR) {
            prevBase = base;
            prevActive = minCur < kMax;
        }
    };

    // Winner-take-all over the parked pixels of steps [sb, sb+16).
    auto batch = [&](int sb) {
        __syncthreads();
        constexpr int Q = G / 16;   // lanes per parked pixel
        constexpr int PL = GK / Q;  // labels scanned per lane
        const int q = lane / Q;
        const int sub = lane - q * Q;
        const int jj = q / kBatch;
        const int tt = q - jj * kBatch;
        const int s = sb + tt;
        const int lineB = line0 + jj;
        const bool ok = s < s1 && lineB < nLines;
        const int16_t* row = s_ring + q * RP;
        const int d0 = sub * PL;

        uint32_t kmin = 0xFFFFFFFFu;
#pragma unroll
        for (int i = 0; i < PL; i += 2) {
            const uint32_t w = *reinterpret_cast<const uint32_t*>(row + d0 + i);
            const int d = d0 + i;
            if (d < D) kmin = umin(kmin, (w << 16) | (uint32_t)d);
            if (d + 1 < D) kmin = umin(kmin, (w & 0xFFFF0000u) | (uint32_t)(d + 1));
        }
        kmin = quad_umin<Q>(kmin);
        const int minCost = (int)(int16_t)(kmin >> 16);
        const int bestD = (int)(kmin & 0xFFFFu);
        int alt = kMax;
#pragma unroll
        for (int i = 0; i < PL; i += 2) {
            const uint32_t w = *reinterpret_cast<const uint32_t*>(row + d0 + i);
            const int d = d0 + i;
            if (d < D && abs(d - bestD) > 1) alt = imin(alt, (int)(int16_t)(w & 0xFFFF));
            if (d + 1 < D && abs(d + 1 - bestD) > 1) alt = imin(alt, (int)w >> 16);
        }
        alt = quad_min<Q>(alt);

        if (sub == 0 && ok) {
            const int x = W - 1 - s;
            const int y = lineB;
            const size_t p = (size_t)y * W + x;
            const int16_t invalid = (int16_t)((a.minD - 1) * 16);
            int16_t out = invalid;
            float cf = 0.0f;
            if constexpr (!PYR) {
                const int minX1 = imax(a.minD + D, 0);
                const int maxX1 = W + imin(a.minD, 0);
                bool valid = s > 0 && x >= minX1 && x < maxX1;
                if (a.uniq > 0 && minCost < kMax / 2 && alt * (100 - a.uniq) < minCost * 100)
                    valid = false;
                if (valid && minCost < kMax / 2) {
                    int dint = bestD * 16;
                    if (bestD > 0 && bestD < D - 1) {
                        const int c0 = row[bestD - 1], c1 = row[bestD], c2 = row[bestD + 1];
                        const int den = imax(c0 + c2 - 2 * c1, 1);
                        dint += div_trunc((c0 - c2) * 16 + den, den * 2);
                    }
                    out = (int16_t)(dint + a.minD * 16);
                    const float denom = fmaxf((float)minCost, 1.0f);
                    cf = fminf(fmaxf(((float)alt - (float)minCost) / denom, 0.0f), 1.0f);
                    const int x2 = x - bestD - a.minD;
                    if (x2 >= 0 && x2 < W)
                        atomicMin(&a.keys[(size_t)y * W + x2],
                                  ((uint32_t)(minCost + 32768) << 16) | (uint32_t)(0xFFFF - x));
                }
            } else {
                const uint32_t meta = s_meta[jj * kBatch + tt];
                const int base = (int)(int16_t)(meta & 0xFFFF);
                const int vc = (int)(meta >> 16);
                bool valid = vc >= 3 && minCost < kMax && alt < kMax;
                if (valid && a.uniq > 0 && alt * (100 - a.uniq) < minCost * 100) valid = false;
                const int realBest = base + bestD;
                const int physMax = imin(a.maxD, x);
                if ((bestD == 0 && base > a.minD) || (bestD == D - 1 && base + D - 1 < physMax))
                    valid = false;
                float c = 0.0f;
                if (valid) {
                    const float denom = fmaxf((float)minCost, 1.0f);
                    c = fminf(fmaxf(((float)alt - (float)minCost) / denom, 0.0f), 1.0f);
                    if (c < a.minConf) valid = false;
                }
                if (valid) {
                    int d16 = realBest * 16;
                    if (bestD > 0 && bestD + 1 < D) {
                        const int c0 = row[bestD - 1], c1 = row[bestD], c2 = row[bestD + 1];
                        if (c0 < kMax && c1 < kMax && c2 < kMax) {
                            const int den = imax(c0 + c2 - 2 * c1, 1);
                            d16 += div_trunc((c0 - c2) * 16 + den, den * 2);
                        }
                    }
                    out = (int16_t)d16;
                    cf = c;
                    const int x2 = x - realBest;
                    if (x2 >= 0 && x2 < W)
                        atomicMin(&a.keys[(size_t)y * W + x2],
                                  ((uint32_t)(minCost + 32768) << 16) | (uint32_t)(0xFFFF - x));
                }
            }
            a.disp[p] = out;
            a.conf[p] = cf;
        }
        __syncthreads();
    };

    // Process one prefetched chunk [sc, sc+U).
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
        ChunkAddr<U> co{};
        // The R->L pass stores nothing per step; its (unused) resource has no range.
        __amdgpu_buffer_rsrc_t rOut = rsrc(a.C, 0);
        if constexpr (!RL) {
            co = chunk(sc, volRow, VB, VB);
            rOut = rsrc(reinterpret_cast<char*>(DIR == kLR ? a.Alr : a.Atb) + co.origin, co.nr);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int s = sc + u;
            if (s < s1) step(buf[u], s, (s - s0) & (kBatch - 1), p2[u], rOut, co.soff[u]);
        }
        if constexpr (RL) {
            const int done = sc + U - s0;
            if ((done & (kBatch - 1)) == 0 || sc + U >= s1) batch(s0 + ((sc - s0) & ~(kBatch - 1)));
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

template <int G, int K, bool PYR, bool ADAPT>
__global__ __launch_bounds__(64) void kPathsLRTB(PathArgs a)
{
    __shared__ int s_lut[256];
    __shared__ int s_xchg[64 * K];
    if constexpr (ADAPT) {
        for (int i = threadIdx.x; i < 256; i += 64) s_lut[i] = a.lut[i];
        __syncthreads();
    }
    if ((int)blockIdx.x < a.nLRBlocks)
        runLine<G, K, PYR, ADAPT, kLR>(a, blockIdx.x, s_lut, s_xchg, nullptr, nullptr);
    else
        runLine<G, K, PYR, ADAPT, kTB>(a, blockIdx.x - a.nLRBlocks, s_lut, s_xchg, nullptr, nullptr);
}

template <int G, int K, bool PYR, bool ADAPT>
__global__ __launch_bounds__(64) void kPathRL(PathArgs a)
{
    constexpr int LPW = 64 / G;
    __shared__ int s_lut[256];
    __shared__ int s_xchg[64 * K];
    __shared__ __attribute__((aligned(16))) int16_t s_ring[LPW * kBatch * (G * K + 2)];
    __shared__ uint32_t s_meta[LPW * kBatch];
    if constexpr (ADAPT) {
        for (int i = threadIdx.x; i < 256; i += 64) s_lut[i] = a.lut[i];
        __syncthreads();
    }
    runLine<G, K, PYR, ADAPT, kRL>(a, blockIdx.x, s_lut, s_xchg, s_ring, s_meta);
}

template <bool PYR, bool ADAPT>
hipError_t dispatchLRTB(const PathArgs& a, int nLR, int nTB, hipStream_t s)
{
    PathArgs b = a;
    b.nLRBlocks = nLR;
    const int blocks = nLR + nTB;
    switch (a.Dp) {
    case 16: kPathsLRTB<16, 1, PYR, ADAPT><<<blocks, 64, 0, s>>>(b); break;
    case 32: kPathsLRTB<32, 1, PYR, ADAPT><<<blocks, 64, 0, s>>>(b); break;
    case 64: kPathsLRTB<64, 1, PYR, ADAPT><<<blocks, 64, 0, s>>>(b); break;
    case 128: kPathsLRTB<64, 2, PYR, ADAPT><<<blocks, 64, 0, s>>>(b); break;
    case 256: kPathsLRTB<64, 4, PYR, ADAPT><<<blocks, 64, 0, s>>>(b); break;
    case 512: kPathsLRTB<64, 8, PYR, ADAPT><<<blocks, 64, 0, s>>>(b); break;
    case 1024: kPathsLRTB<64, 16, PYR, ADAPT><<<blocks, 64, 0, s>>>(b); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool PYR, bool ADAPT>
hipError_t dispatchRL(const PathArgs& a, hipStream_t s)
{
    const int lpw = a.Dp <= 16 ? 4 : (a.Dp <= 32 ? 2 : 1);
    const int blocks = (a.H + lpw - 1) / lpw;
    switch (a.Dp) {
    case 16: kPathRL<16, 1, PYR, ADAPT><<<blocks, 64, 0, s>>>(a); break;
    case 32: kPathRL<32, 1, PYR, ADAPT><<<blocks, 64, 0, s>>>(a); break;
    case 64: kPathRL<64, 1, PYR, ADAPT><<<blocks, 64, 0, s>>>(a); break;
    case 128: kPathRL<64, 2, PYR, ADAPT><<<blocks, 64, 0, s>>>(a); break;
    case 256: kPathRL<64, 4, PYR, ADAPT><<<blocks, 64, 0, s>>>(a); break;
    case 512: kPathRL<64, 8, PYR, ADAPT><<<blocks, 64, 0, s>>>(a); break;
    case 1024: kPathRL<64, 16, PYR, ADAPT><<<blocks, 64, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace

hipError_t launchPathsLRTB(const PathArgs& a, int mode, int adaptive, int runLR, int runTB,
                           hipStream_t s)
{
    const int lpw = a.Dp <= 16 ? 4 : (a.Dp <= 32 ? 2 : 1);
    const int nLR = runLR ? (a.H + lpw - 1) / lpw : 0;
    const int nTB = runTB ? (a.W + lpw - 1) / lpw : 0;
    if (nLR + nTB == 0) return hipSuccess;
    const bool pyr = mode != kSingle;
    if (pyr) return adaptive ? dispatchLRTB<true, true>(a, nLR, nTB, s) : dispatchLRTB<true, false>(a, nLR, nTB, s);
    return adaptive ? dispatchLRTB<false, true>(a, nLR, nTB, s) : dispatchLRTB<false, false>(a, nLR, nTB, s);
}

hipError_t launchPathRL(const PathArgs& a, int mode, int adaptive, hipStream_t s)
{
    const bool pyr = mode != kSingle;
    if (pyr) return adaptive ? dispatchRL<true, true>(a, s) : dispatchRL<true, false>(a, s);
    return adaptive ? dispatchRL<false, true>(a, s) : dispatchRL<false, false>(a, s);
}

} // namespace pyr
