// kernels_wta.hip -- path sums + winner-take-all for one level.
//
// Volumes are int16 (SHRT_MAX - L) or, in the byte form, uint8 (255 - L, 0
// for SHRT_MAX); both are converted to L on load, so the rest is shared.
//
// Q = Dp/16 lanes per pixel, 16 labels per lane, pixels consecutive across
// lanes, so every volume read is a run of full 16-byte loads.  Per pixel:
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
#include "path_impl.h"

namespace pyr {
namespace {

using path::gmin;
using path::gumin;

// Path volumes hold H = SHRT_MAX - L; returns the 16 path costs L.
__device__ __forceinline__ void load16(int (&v)[16], const int16_t* p)
{
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[2 * i] = kMax - (int)(w[i] & 0xFFFFu);
        v[2 * i + 1] = kMax - (int)(w[i] >> 16);
    }
}

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
    int16_t* const dispV = pairAt(a.disp, pr, wsS);
    float* const confV = pairAt(a.conf, pr, wsS);
    uint32_t* const keys = pairAt(a.keys, pr, wsS);
    const size_t t = bx * 256 + threadIdx.x;
    const size_t p = t / Q;
    const int sub = (int)(t - p * Q);
    const bool ok = p < P;
    const size_t pc = ok ? p : P - 1;
    const size_t off = pc * a.Dp + (size_t)sub * 16;
    const int np = a.numPaths;
    // The pixel's label word is loaded with the volumes (pc is in bounds), so
    // that the wave waits on memory once, not again after the label search.
    const int lwPre = MODE == kRefine ? labels[pc] : 0;

    const int d0 = sub * 16;
    uint32_t kmin;
    int minCost, bestD, alt, c0, c1, c2;
    if constexpr (NARROW) {
        // Byte volumes keep every path cost below 255 (narrowVolumes), so each
        // sum above equals "SHRT_MAX if a used path is SHRT_MAX at the label,
        // else the plain sum", and the label's sum is 255 * np - S with
        // S = sum of the stored bytes.  S is formed for label pairs in packed
        // 16-bit fields (Sv[k]: labels 2k, 2k+1); a field is 0 where a used
        // path is SHRT_MAX.  The argmin over (cost, d) is then the maximum of
        // the packed key S * 16 + (15 - label).
        const uint4 vr = *reinterpret_cast<const uint4*>(reinterpret_cast<const uint8_t*>(Arl) + off);
        const uint4 vl = np >= 2 ? *reinterpret_cast<const uint4*>(reinterpret_cast<const uint8_t*>(Alr) + off)
                                 : make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        const uint4 vt = np >= 3 ? *reinterpret_cast<const uint4*>(reinterpret_cast<const uint8_t*>(Atb) + off)
                                 : make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        const uint32_t wr[4] = {vr.x, vr.y, vr.z, vr.w};
        const uint32_t wl[4] = {vl.x, vl.y, vl.z, vl.w};
        const uint32_t wt[4] = {vt.x, vt.y, vt.z, vt.w};
        // Unused paths read 0xFF bytes: they pass the validity test, and their
        // 255 per label is taken off again by the 255 * 3 below.
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
            Km[k] = U32((h2){0, 0} - __builtin_elementwise_min(mn, (h2){1, 1}));
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
        if (pad)
            kmin = 0xFFFFFFFFu;
        else if (best == 0)
            kmin = ((uint32_t)kMax << 16) | (uint32_t)d0;
        else
            kmin = ((uint32_t)costOf(best >> 4) << 16) | (uint32_t)(d0 + 15 - (int)(best & 15u));
        kmin = gumin<Q>(kmin);
        minCost = (int)(int16_t)(kmin >> 16);
        bestD = (int)(kmin & 0xFFFFu);
        const int bl = bestD - d0;  // best label's index in this lane (may be outside [0, 16))
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
        c0 = field(bl - 1);
        c2 = field(bl + 1);
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
        alt = gmin<Q>(costOf(bestAlt));
        if constexpr (Q > 1) {
            c0 = gmin<Q>(c0);
            c2 = gmin<Q>(c2);
        }
        c1 = minCost;
    } else {
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
        kmin = 0xFFFFFFFFu;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (d0 + i < D) kmin = umin(kmin, ((uint32_t)(uint16_t)my[i] << 16) | (uint32_t)(d0 + i));
        kmin = gumin<Q>(kmin);
        minCost = (int)(int16_t)(kmin >> 16);
        bestD = (int)(kmin & 0xFFFFu);
        alt = kMax;
        c0 = kMax;
        c1 = kMax;
        c2 = kMax;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int d = d0 + i;
            if (d < D && abs(d - bestD) > 1) alt = imin(alt, my[i]);
            if (d == bestD - 1) c0 = my[i];
            if (d == bestD) c1 = my[i];
            if (d == bestD + 1) c2 = my[i];
        }
        alt = gmin<Q>(alt);
        if constexpr (Q > 1) {
            c0 = gmin<Q>(c0);
            c1 = gmin<Q>(c1);
            c2 = gmin<Q>(c2);
        }
    }
    if (sub != 0 || !ok) return;

    const int y = (int)(p / W);
    const int x = (int)(p - (size_t)y * W);
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
                const int den = imax(c0 + c2 - 2 * c1, 1);
                dint += div_trunc((c0 - c2) * 16 + den, den * 2);
            }
            out = (int16_t)(dint + a.minD * 16);
            const float denom = fmaxf((float)minCost, 1.0f);
            cf = fminf(fmaxf(((float)alt - (float)minCost) / denom, 0.0f), 1.0f);
            const int x2 = x - bestD - a.minD;
            if (x2 >= 0 && x2 < W)
                atomicMin(&keys[(size_t)y * W + x2],
                          ((uint32_t)(minCost + 32768) << 16) | (uint32_t)(0xFFFF - x));
        }
    } else {
        int base = a.minD, valid = 1;
        if constexpr (MODE == kRefine) {
            const int lw = lwPre;
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
            if (bestD > 0 && bestD + 1 < D && c0 < kMax && c1 < kMax && c2 < kMax) {
                const int den = imax(c0 + c2 - 2 * c1, 1);
                d16 += div_trunc((c0 - c2) * 16 + den, den * 2);
            }
            out = (int16_t)d16;
            cf = c;
            const int x2 = x - realBest;
            if (x2 >= 0 && x2 < W)
                atomicMin(&keys[(size_t)y * W + x2],
                          ((uint32_t)(minCost + 32768) << 16) | (uint32_t)(0xFFFF - x));
        }
    }
    dispV[p] = out;
    confV[p] = cf;
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
