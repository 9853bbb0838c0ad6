// wta_impl.h -- winner-take-all of one pixel from its three path-cost vectors,
// shared by the winner-take-all kernel (kernels_wta.hip) and the R->L path pass
// that finishes each pixel as soon as its R->L costs exist (path_impl.h).
//
// A pixel's labels are spread over Q lanes, 16 labels per lane (Q = 1 in the
// fused R->L pass).  Per pixel:
//   * sum the three path costs in the reference's order and saturation:
//     single scale acc = min(MAX, Lr + Tb) then min(MAX, acc + Rl)
//     (sgm_kernels.cu:634, :740-741); pyramid
//     old = (Lr<MAX && Tb<MAX) ? min(MAX-1, Lr+Tb) : MAX, then the same with Rl
//     (sgm_kernels.cu:1325-1328, :1407-1411); numPaths 2 and 1 drop T->B / L->R;
//   * argmin over the packed key (cost << 16 | d): smallest cost, then smallest d
//     (sgm_kernels.cu:750-753), reduced across the Q lanes with DPP;
//   * second-best cost over |d - best| > 1, uniqueness, sub-pixel parabola,
//     confidence, validity rules (single scale :759-808; pyramid :1422-1470);
//   * the right-view projection disp2 as atomicMin of
//     ((cost + 32768) << 16) | (0xFFFF - x): the smallest cost wins and, among
//     equal costs, the largest x -- the pixel the reference's sequential
//     right-to-left update keeps (sgm_kernels.cu:794-797, :1462-1465).
#pragma once

#include "device_common.h"
#include "internal.h"

namespace pyr {
namespace wta {

// Global-memory accesses through pointers held in an argument block: the
// compiler cannot infer their address space, and a generic (flat) access also
// counts against the LDS counter, so a wave waits on both.
template <class T>
__device__ __forceinline__ T ldGlobal(const T* p)
{
    return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ void stGlobal(T* p, T v)
{
    *(__attribute__((address_space(1))) T*)p = v;
}
__device__ __forceinline__ void minGlobal(uint32_t* p, uint32_t v)
{
    __hip_atomic_fetch_min((__attribute__((address_space(1))) uint32_t*)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// The label search of one pixel: best (smallest) summed cost and its label,
// the best cost outside labels best-1 .. best+1, and the costs at best-1,
// best, best+1 (SHRT_MAX where a label does not exist).
struct Pick {
    int minCost, bestD, alt, c0, c1, c2;
};

// Byte volumes: 16 labels d0 .. d0+15 of the pixel in this lane, as the three
// stored 16-byte vectors (255 - L; a path that the level does not run reads
// 0xFF bytes).  Every path cost is below 255 (narrowVolumes), so each sum
// above equals "SHRT_MAX if a used path is SHRT_MAX at the label, else the
// plain sum", and the label's sum is 255 * 3 - S with S the sum of the stored
// bytes.  S is formed for label pairs in packed 16-bit fields (Sv[k]: labels
// 2k, 2k+1); a field is 0 where a used path is SHRT_MAX.  The argmin over
// (cost, d) is then the maximum of the packed key S * 16 + (15 - label).
template <int Q>
__device__ __forceinline__ Pick searchNarrow(const uint32_t (&wr)[4], const uint32_t (&wl)[4],
                                             const uint32_t (&wt)[4], int d0, int D)
{
    typedef uint16_t h2 __attribute__((ext_vector_type(2)));
    auto H2 = [](uint32_t x) { return __builtin_bit_cast(h2, x); };
    auto U32 = [](h2 x) { return __builtin_bit_cast(uint32_t, x); };
    uint32_t Sv[8], Km[8];  // masked pair sums; 0xFFFF per field where all used paths are valid
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t sel = (k & 1) ? 0x0C030C02u : 0x0C010C00u;  // bytes 2k, 2k+1 -> 16-bit fields
        const uint32_t a = __builtin_amdgcn_perm(0u, wr[k >> 1], sel);
        const uint32_t b = __builtin_amdgcn_perm(0u, wl[k >> 1], sel);
        const uint32_t c = __builtin_amdgcn_perm(0u, wt[k >> 1], sel);
        const h2 mn = __builtin_elementwise_min(__builtin_elementwise_min(H2(a), H2(b)), H2(c));
        // 0 - min(mn, 1) per field as two packed instructions (the compiler
        // would otherwise form it with compares and selects).  The 1s come
        // from a register: a packed instruction reads an inline constant's
        // high half from its bits 16-31, which are 0.
        asm("v_pk_min_u16 %0, %1, %2\n\tv_pk_sub_u16 %0, 0, %0" : "=&v"(Km[k]) : "v"(U32(mn)), "s"(0x00010001u));
        Sv[k] = (a + b + c) & Km[k];
    }
    const bool pad = d0 >= D;  // D is a multiple of 16: a lane is all labels or all padding
    const int tot = 255 * 3;   // S counts 255 for every unused path
    auto costOf = [&](uint32_t sf) { return sf ? tot - (int)sf : kMax; };
    // Packed keys S * 16 + (15 - label), 0 where invalid; their maximum.
    h2 best2 = (h2){0, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t lab = (uint32_t)(15 - 2 * k) | ((uint32_t)(14 - 2 * k) << 16);
        best2 = __builtin_elementwise_max(best2, H2((Sv[k] << 4) | (lab & Km[k])));
    }
    const uint32_t best = umax(best2[0], best2[1]);
    uint32_t kmin;
    if (pad)
        kmin = 0xFFFFFFFFu;
    else if (best == 0)
        kmin = ((uint32_t)kMax << 16) | (uint32_t)d0;
    else
        kmin = ((uint32_t)costOf(best >> 4) << 16) | (uint32_t)(d0 + 15 - (int)(best & 15u));
    kmin = gumin<Q>(kmin);
    Pick r;
    r.minCost = (int)(int16_t)(kmin >> 16);
    r.bestD = (int)(kmin & 0xFFFFu);
    const int bl = r.bestD - d0;  // best label's index in this lane (may be outside [0, 16))
    // Field f of the packed sums (kMax outside this lane), by a 3-level select
    // tree on the pair index (a run-time index into Sv would put Sv in memory).
    auto field = [&](int f) {
        if (pad || f < 0 || f >= 16) return kMax;
        const int k = f >> 1;
        const uint32_t a01 = (k & 1) ? Sv[1] : Sv[0], a23 = (k & 1) ? Sv[3] : Sv[2];
        const uint32_t a45 = (k & 1) ? Sv[5] : Sv[4], a67 = (k & 1) ? Sv[7] : Sv[6];
        const uint32_t a03 = (k & 2) ? a23 : a01, a47 = (k & 2) ? a67 : a45;
        const uint32_t x = (k & 4) ? a47 : a03;
        return costOf((f & 1) ? (x >> 16) : (x & 0xFFFFu));
    };
    r.c0 = field(bl - 1);
    r.c2 = field(bl + 1);
    // alt: best sum outside labels bl-1 .. bl+1.
    const int sh = imin(imax(bl - 1, -3), 17);
    const uint32_t X = (sh >= 0 ? (7u << sh) : (7u >> (-sh))) & 0xFFFFu;
    h2 alt2 = (h2){0, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        // 0xFFFF in the fields of excluded labels 2k, 2k+1.
        const uint32_t ex = (__builtin_amdgcn_ubfe(X, 2 * k, 1) |
                             (__builtin_amdgcn_ubfe(X, 2 * k + 1, 1) << 16)) * 0xFFFFu;
        alt2 = __builtin_elementwise_max(alt2, H2(Sv[k] & ~ex));
    }
    const uint32_t bestAlt = pad ? 0u : umax(alt2[0], alt2[1]);
    r.alt = gmin<Q>(costOf(bestAlt));
    if constexpr (Q > 1) {
        r.c0 = gmin<Q>(r.c0);
        r.c2 = gmin<Q>(r.c2);
    }
    r.c1 = r.minCost;
    return r;
}

// Path volumes hold H = SHRT_MAX - L; returns the 16 path costs L.
__device__ __forceinline__ void load16(int (&v)[16], const int16_t* p)
{
    typedef uint32_t v4 __attribute__((ext_vector_type(4)));
    const v4 a = ldGlobal(reinterpret_cast<const v4*>(p));
    const v4 b = ldGlobal(reinterpret_cast<const v4*>(p) + 1);
    const uint32_t w[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[2 * i] = kMax - (int)(w[i] & 0xFFFFu);
        v[2 * i + 1] = kMax - (int)(w[i] >> 16);
    }
}

// int16 volumes: labels d0 .. d0+15 read from the three volumes at `off`.
template <int Q, bool PYR>
__device__ __forceinline__ Pick searchWide(const int16_t* Arl, const int16_t* Alr, const int16_t* Atb,
                                           size_t off, int np, int d0, int D)
{
    int my[16];
    {
        int rl[16], lr[16], tb[16];
        load16(rl, Arl + off);
        if (np >= 2) load16(lr, Alr + off);
        if (np >= 3) load16(tb, Atb + off);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if constexpr (!PYR) {
                int acc = np >= 2 ? lr[i] : 0;
                if (np >= 3) acc = (int)(int16_t)imin(kMax, acc + tb[i]);
                my[i] = (int)(int16_t)imin(kMax, acc + rl[i]);
            } else {
                int old = np >= 2 ? lr[i] : 0;
                if (np >= 3) old = imax(lr[i], tb[i]) == kMax ? kMax : imin(kMax - 1, lr[i] + tb[i]);
                my[i] = np >= 2 ? (imax(old, rl[i]) == kMax ? kMax : imin(kMax - 1, old + rl[i]))
                                : rl[i];
            }
        }
    }
    uint32_t kmin = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        if (d0 + i < D) kmin = umin(kmin, ((uint32_t)(uint16_t)my[i] << 16) | (uint32_t)(d0 + i));
    kmin = gumin<Q>(kmin);
    Pick r;
    r.minCost = (int)(int16_t)(kmin >> 16);
    r.bestD = (int)(kmin & 0xFFFFu);
    r.alt = kMax;
    r.c0 = kMax;
    r.c1 = kMax;
    r.c2 = kMax;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int d = d0 + i;
        if (d < D && abs(d - r.bestD) > 1) r.alt = imin(r.alt, my[i]);
        if (d == r.bestD - 1) r.c0 = my[i];
        if (d == r.bestD) r.c1 = my[i];
        if (d == r.bestD + 1) r.c2 = my[i];
    }
    r.alt = gmin<Q>(r.alt);
    if constexpr (Q > 1) {
        r.c0 = gmin<Q>(r.c0);
        r.c1 = gmin<Q>(r.c1);
        r.c2 = gmin<Q>(r.c2);
    }
    return r;
}

// Row of pixel index p of a W-wide level, for p < 2^32 and W < 2^16: invW is
// 1/W rounded to double and raised by one ulp, so invW >= 1/W and the product
// lies in [p/W, p/W + 2^-18); floor(p/W) = q with p mod W <= W - 1 leaves a
// gap of 1/W >= 2^-16 below q + 1, so the truncation is exact.
__device__ __forceinline__ uint32_t rowOf(uint32_t p, double invW)
{
    return (uint32_t)((double)p * invW);
}

// Validity rules, sub-pixel disparity, confidence and the right-view key of
// pixel (x, y) = p of a W-wide level; writes dispV[p] and confV[p].  lw: the
// pixel's refinement label word (window base in bits 0-15, parent valid in
// bit 16), read at refinement levels only.
template <int MODE>
__device__ __forceinline__ void resolve(const PathArgs& a, const Pick& k, int x, int y, size_t p,
                                        int lw, int16_t* dispV, float* confV, uint32_t* keys)
{
    constexpr bool PYR = MODE != kSingle;
    const int W = a.W, D = a.D;
    const int minCost = k.minCost, bestD = k.bestD, alt = k.alt;
    const int16_t invalid = (int16_t)((a.minD - 1) * 16);
    int16_t out = invalid;
    float cf = 0.0f;
    if constexpr (!PYR) {
        const int minX1 = imax(a.minD + D, 0);
        const int maxX1 = W + imin(a.minD, 0);
        // x = W-1 starts the R->L pass and is never matched (sgm_kernels.cu:710-716).
        bool valid = x < W - 1 && x >= minX1 && x < maxX1;
        if (a.uniq > 0 && minCost < kMax / 2 && alt * (100 - a.uniq) < minCost * 100) valid = false;
        if (valid && minCost < kMax / 2) {
            int dint = bestD * 16;
            if (bestD > 0 && bestD < D - 1) {
                const int den = imax(k.c0 + k.c2 - 2 * k.c1, 1);
                dint += div_trunc((k.c0 - k.c2) * 16 + den, den * 2);
            }
            out = (int16_t)(dint + a.minD * 16);
            const float denom = fmaxf((float)minCost, 1.0f);
            cf = fminf(fmaxf(((float)alt - (float)minCost) / denom, 0.0f), 1.0f);
            const int x2 = x - bestD - a.minD;
            if (x2 >= 0 && x2 < W)
                minGlobal(&keys[(size_t)y * W + x2], ((uint32_t)(minCost + 32768) << 16) | (uint32_t)(0xFFFF - x));
        }
    } else {
        int base = a.minD, valid = 1;
        if constexpr (MODE == kRefine) {
            base = (int)(int16_t)(lw & 0xFFFF);
            valid = (lw >> 16) & 1;
        }
        const int lo = imax(a.minD, x - W + 1), hi = imin(a.maxD, x);
        const int fl = imax(0, lo - base), ll = imin(D - 1, hi - base);
        const int vc = (valid && ll >= fl) ? ll - fl + 1 : 0;
        bool good = vc >= 3 && minCost < kMax && alt < kMax;
        if (good && a.uniq > 0 && alt * (100 - a.uniq) < minCost * 100) good = false;
        const int realBest = base + bestD;
        const int physMax = imin(a.maxD, x);
        if ((bestD == 0 && base > a.minD) || (bestD == D - 1 && base + D - 1 < physMax)) good = false;
        float c = 0.0f;
        if (good) {
            const float denom = fmaxf((float)minCost, 1.0f);
            c = fminf(fmaxf(((float)alt - (float)minCost) / denom, 0.0f), 1.0f);
            if (c < a.minConf) good = false;
        }
        if (good) {
            int d16 = realBest * 16;
            if (bestD > 0 && bestD + 1 < D && k.c0 < kMax && k.c1 < kMax && k.c2 < kMax) {
                const int den = imax(k.c0 + k.c2 - 2 * k.c1, 1);
                d16 += div_trunc((k.c0 - k.c2) * 16 + den, den * 2);
            }
            out = (int16_t)d16;
            cf = c;
            const int x2 = x - realBest;
            if (x2 >= 0 && x2 < W)
                minGlobal(&keys[(size_t)y * W + x2], ((uint32_t)(minCost + 32768) << 16) | (uint32_t)(0xFFFF - x));
        }
    }
    stGlobal(dispV + p, out);
    stGlobal(confV + p, cf);
}

} // namespace wta
} // namespace pyr
