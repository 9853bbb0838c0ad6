// kernels_path.hip -- SGM path aggregation for gfx950.
//
// One wave64 runs 64/G independent lines (rows for L->R / R->L, columns for
// T->B); each line is a G-lane group holding K labels per lane.  A pixel step
// is: fetch the previous pixel's path costs for d, d-1, d+1 (own registers +
// one DPP shift each way, or an LDS exchange when a refinement line's label
// window moved), add the penalties, and reduce the minimum over the labels with
// four DPP moves (plus permlane swaps for 32/64-lane groups).  The loop is
// software-pipelined: the costs of the next U pixel steps are loaded while the
// current U are processed.
//
// Launches:
//   kPathsLRTB: L->R rows and T->B columns in ONE launch (they are independent);
//               writes Alr and Atb.
//   kPathRL:    R->L rows; adds Alr/Atb with the reference's saturation order and
//               parks each pixel's summed costs in an LDS ring; every 16 steps
//               the wave turns around and does winner-take-all for the 16*64/G
//               parked pixels, one pixel per 1..4 lanes (uniqueness, sub-pixel,
//               confidence, and the right-view projection as an order-free
//               atomicMin key).
//
// Semantics restated from the reference:
//   single scale: sgm_kernels.cu:520-578 (L->R), :587-638 (T->B), :665-819 (R->L+WTA)
//   pyramid:      sgm_kernels.cu:1191-1261, :1263-1334, :1336-1478
#include "device_common.h"
#include "internal.h"

namespace pyr {
namespace {

enum Dir : int { kLR = 0, kTB = 1, kRL = 2 };

constexpr int kBatch = 16;  // R->L pixels parked per line before WTA

// Pixel steps fetched ahead: keeps ~32 labels per lane in flight.
template <int K>
constexpr int prefetchSteps() { return K <= 2 ? 16 : 32 / K; }

template <int K>
struct StepIn {
    Packed<K> c, lr, tb;
    int lw;   // refinement label word
    int im;   // left-image intensity (adaptive P2)
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
    constexpr int RP = GK + 2;  // ring pitch (int16 per parked pixel)

    const int lane = threadIdx.x;
    const int j = lane / G;
    const int g = lane - j * G;
    const int W = a.W, H = a.H, D = a.D, Dp = a.Dp;
    const int nLines = VERT ? W : H;
    const int line = lineGroup * LPW + j;
    const bool lineOk = line < nLines;
    const int ln = lineOk ? line : nLines - 1;
    int s0 = 0, s1 = VERT ? H : W;
    if constexpr (!PYR && DIR == kLR) {
        s0 = imax(a.startX, 0);
        s1 = imin(a.endX, W);
    }
    if (s0 >= s1) return;

    const bool hasLR = RL && a.numPaths >= 2;
    const bool hasTB = RL && a.numPaths >= 3;
    const bool refine = PYR && a.labels != nullptr;

    auto colOf = [&](int s) { return VERT ? ln : (RL ? W - 1 - s : s); };
    auto rowOf = [&](int s) { return VERT ? s : ln; };

    StepIn<K> cur[U], nxt[U];
    auto load = [&](StepIn<K>(&buf)[U], int sc) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int s = imin(sc + u, s1 - 1);
            const size_t p = (size_t)rowOf(s) * W + colOf(s);
            const size_t off = p * Dp + (size_t)g * K;
            load_packed<K>(buf[u].c, a.C + off);
            if constexpr (RL) {
                if (hasLR) load_packed<K>(buf[u].lr, a.Alr + off);
                if (hasTB) load_packed<K>(buf[u].tb, a.Atb + off);
            }
            if constexpr (PYR) buf[u].lw = refine ? a.labels[p] : 0;
            if constexpr (ADAPT) buf[u].im = a.img[(size_t)rowOf(s) * a.istep + colOf(s)];
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

    auto step = [&](const StepIn<K>& st, int s, int t) {
        int c[K];
#pragma unroll
        for (int k = 0; k < K; ++k) c[k] = unpack<K>(st.c, k);
        int p2 = a.P2;
        if constexpr (ADAPT) {
            p2 = s_lut[abs(st.im - prevIm)];
            prevIm = st.im;
        }
        int Lc[K];
        bool lv[K];
        int base = a.minD, vcount = 0;
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
                    const int d = g * K + k;
                    const int Lm = k > 0 ? Lp[k - 1] : lmS;
                    const int Lq = k < K - 1 ? Lp[k + 1] : lpS;
                    const int lm = d == 0 ? kMax : Lm + P1;
                    const int lp = d >= D - 1 ? kMax : Lq + P1;
                    const int best = imin(imin(lm, lp), imin(Lp[k], lj));
                    Lc[k] = sat16(c[k] + best - minPrev);
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                lv[k] = g * K + k < D;
                if (!lv[k]) Lc[k] = kMax;
            }
        } else {
            int valid = 1;
            if (refine) {
                base = (int)(int16_t)(st.lw & 0xFFFF);
                valid = (st.lw >> 16) & 1;
            }
            const int x = colOf(s);
            const int lo = imax(a.minD, x - W + 1);
            const int hi = imin(a.maxD, x);
            const int fl = imax(0, lo - base);
            const int ll = imin(D - 1, hi - base);
            vcount = (valid && ll >= fl) ? ll - fl + 1 : 0;
            int cc[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int d = g * K + k;
                const int real = base + d;
                lv[k] = valid && d < D && real >= lo && real <= hi && c[k] < kMax;
                cc[k] = lv[k] ? c[k] : kMax;
            }
            const int sh = base - prevBase;
            int l0[K], lm[K], lp[K];
            if (__ballot(prevActive && sh != 0) != 0) {
                // A line's label window moved: exchange its previous costs through LDS.
                __syncthreads();
#pragma unroll
                for (int k = 0; k < K; ++k) s_xchg[lane * K + k] = Lp[k];
                __syncthreads();
                const int* lineX = s_xchg + j * GK;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int sm = g * K + k + sh;
                    const int v0 = (sm >= 0 && sm < D) ? lineX[sm] : kMax;
                    const int vm = (sm - 1 >= 0 && sm - 1 < D) ? lineX[sm - 1] : kMax;
                    const int vp = (sm + 1 >= 0 && sm + 1 < D) ? lineX[sm + 1] : kMax;
                    l0[k] = v0;
                    lm[k] = vm < kMax ? vm + P1 : kMax;
                    lp[k] = vp < kMax ? vp + P1 : kMax;
                }
            } else {
                const int lmS = from_lower<G>(Lp[K - 1], kMax, g);
                const int lpS = from_upper<G>(Lp[0], kMax, g);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int Lm = k > 0 ? Lp[k - 1] : lmS;
                    const int Lq = k < K - 1 ? Lp[k + 1] : lpS;
                    l0[k] = Lp[k];
                    lm[k] = Lm < kMax ? Lm + P1 : kMax;
                    lp[k] = Lq < kMax ? Lq + P1 : kMax;
                }
            }
            const int lj = imin(kMax, minPrev + p2);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int best = imin(imin(l0[k], lm[k]), imin(lp[k], lj));
                const int rec = best < kMax ? sat16(cc[k] + best - minPrev) : kMax;
                int v = prevActive ? (lv[k] ? rec : kMax) : cc[k];
                if (vcount == 0) v = kMax;
                Lc[k] = v;
            }
        }

        int m = Lc[0];
#pragma unroll
        for (int k = 1; k < K; ++k) m = imin(m, Lc[k]);
        const int minCur = group_min<G>(m);

        if constexpr (!RL) {
            if (lineOk) {
                const size_t p = (size_t)rowOf(s) * W + colOf(s);
                store_i16<K>((DIR == kLR ? a.Alr : a.Atb) + p * Dp + (size_t)g * K, Lc);
            }
        } else {
            int my[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int lr = hasLR ? unpack<K>(st.lr, k) : 0;
                const int tb = hasTB ? unpack<K>(st.tb, k) : 0;
                if constexpr (!PYR) {
                    const int acc = hasTB ? (int)(int16_t)imin(kMax, lr + tb) : lr;
                    my[k] = (int)(int16_t)imin(kMax, acc + Lc[k]);
                } else {
                    const int old = hasTB ? ((lr < kMax && tb < kMax) ? imin(kMax - 1, lr + tb) : kMax)
                                          : lr;
                    my[k] = (lv[k] && Lc[k] < kMax && (!hasLR || old < kMax))
                                ? imin(kMax - 1, old + Lc[k]) : kMax;
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
                if (g == 0) s_meta[j * kBatch + t] = (uint32_t)(base & 0xFFFF) | ((uint32_t)vcount << 16);
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
        const int lineB = lineGroup * LPW + jj;
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

    // Main loop: batches of 16 steps, each made of 16/U prefetch chunks.
    load(cur, s0);
    for (int sb = s0; sb < s1; sb += kBatch) {
#pragma unroll
        for (int ch = 0; ch < kBatch / U; ++ch) {
            const int sc = sb + ch * U;
            if (sc + U < s1) load(nxt, sc + U);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (sc + u < s1) step(cur[u], sc + u, ch * U + u);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) cur[u] = nxt[u];
        }
        if constexpr (RL) batch(sb);
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
