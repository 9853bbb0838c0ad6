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
// addressing.  A second wave per workgroup moves all data: it loads costs two
// chunks of U steps ahead into LDS and writes the finished path costs back
// with 16-byte stores, so the recurrence's wave issues no global memory
// instructions.  Lanes of lines past the image edge get an offset beyond the
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

#include <type_traits>

#include "device_common.h"
#include "internal.h"
#include "wta_impl.h"

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

// Register sets of the memory wave (compile-time tuning knob): a chunk's
// global loads are issued this many chunks before the chunk is scattered into
// LDS, which is two chunks before the recurrence reads it.
#ifndef PYRSGM_MEM_SETS
#define PYRSGM_MEM_SETS 2
#endif
constexpr int kMemSets = PYRSGM_MEM_SETS;
static_assert(kMemSets >= 2, "the memory wave keeps at least two chunks of loads in flight");

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

// K labels of one lane into the LDS tile (int16 each, or one byte each in
// the byte-volume form, which has K = 1).
template <int K, bool NARROW = false>
__device__ __forceinline__ void tstore(const int (&v)[K], char* p)
{
    auto pk = [&](int i) { return (uint32_t)(uint16_t)v[i] | ((uint32_t)v[i + 1] << 16); };
    if constexpr (NARROW) {
        static_assert(K == 1, "byte volumes have one label per lane");
        *reinterpret_cast<uint8_t*>(p) = (uint8_t)v[0];
    } else if constexpr (K == 1) {
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

// gmax split before its last cross-row level: for 32-lane lines gmaxPartial
// leaves, in every lane, the maximum over half of its line and lastPair gives
// that value and the other half's (a permlane swap), so that gmax(v) =
// max(a, b); the path step folds that level into its subtraction:
// x - gmax(v) = min(x - a, x - b).  A 16-lane line keeps the whole reduction
// in gmaxPartial (a = b): x - row_mirror(a) would be a v_subrev_u32_dpp,
// whose swizzle gfx950 applies to x instead of a (see dppOpaque).  A 64-lane
// line is the whole wave: its row maxima meet in lane 63 (row_bcast:15 into
// rows 1 and 3, then row_bcast:31 into rows 2 and 3) and the maximum is read
// as a scalar (a = b), with no permlane swaps or their copies and wait states.
template <int G>
__device__ __forceinline__ int gmaxPartial(int v)
{
    v = imax(v, dmov<kDppQuadXor1>(v));
    v = imax(v, dmov<kDppQuadXor2>(v));
    v = imax(v, dmov<kDppRowHalfMirror>(v));
    v = imax(v, dmov<kDppRowMirror>(v));
    if constexpr (G >= 64) {
        // Unsigned maxima with 0 in the masked-off rows (the identity of an
        // unsigned maximum, and v >= 0), so that each move folds into its maximum.
        auto up = [](int x, int o) { return (int)__builtin_elementwise_max((uint32_t)x, (uint32_t)o); };
        v = up(v, __builtin_amdgcn_update_dpp(0, v, kDppRowBcast15, 0xA, 0xF, false));
        v = up(v, __builtin_amdgcn_update_dpp(0, v, kDppRowBcast31, 0xC, 0xF, false));
        v = __builtin_amdgcn_readlane(v, 63);
    }
    return v;
}

// v, with the compiler's knowledge of how it was formed cut off (it is
// neither moved into a later block nor re-associated with its users).
__device__ __forceinline__ int opaqueV(int v)
{
    asm("" : "+v"(v));
    return v;
}
// The same for a wave-uniform value kept in a scalar register.
__device__ __forceinline__ int opaqueS(int v)
{
    asm("" : "+s"(v));
    return v;
}

template <int G>
__device__ __forceinline__ void lastPair(int p, int& a, int& b)
{
    if constexpr (G == 16 || G == 64) {
        a = p;
        b = p;
    } else {
        const auto s = __builtin_amdgcn_permlane16_swap((uint32_t)p, (uint32_t)p, false, false);
        a = (int)s[0];
        b = (int)s[1];
    }
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
// A workgroup is two waves.  The compute wave runs the recurrence of 64/G
// lines and touches only registers and LDS; the memory wave streams the
// chunk's costs and per-step words into LDS two chunks ahead and writes the
// finished path costs back.  Each chunk of U steps ends with one workgroup
// barrier.  LDS plan:
//  * costs: 3 slots (chunks k, k+1 in use, k+2 being filled); in a slot each
//    compute lane owns its K labels for the U steps: 16 dwords (a dword per
//    label pair) when K > 1, 16 halfwords when K = 1, padded by 16 bytes so
//    that the lanes' 16-byte reads hit disjoint banks.  The 16-bit form keeps
//    the K = 1 workgroup at 15 KB, so that ten fit in a CU's LDS;
//  * per-step words: 3 slots of 64 dwords, [line][step slot];
//  * path costs: 2 tiles (chunk k being written, chunk k-1 being stored).
//    Horizontal paths [line][step slot][label], each line's row padded by one
//    pixel's labels (the lines of a wave then write disjoint banks at each
//    step); vertical path [step][line][label].  A horizontal line's row and a
//    vertical step's row are each one contiguous run of the path-cost volume,
//    written with 16-byte stores.
template <int G, int K, bool NARROW = false, bool PK = false, bool MODR = false>
struct Lds {
    static_assert(!NARROW || K == 1, "byte volumes have one label per lane");
    static_assert(!PK || (NARROW && G == 16), "paired lines need byte volumes of 16 labels");
    static constexpr int U = prefetchSteps<K>();
    static constexpr int LPW = 64 / G;
    static constexpr int LPWC = PK ? 2 * LPW : LPW;    // lines per wave (two per lane when paired)
    static constexpr int ESZ = NARROW ? 1 : 2;         // bytes per label in volumes and tiles
    static constexpr int VB = G * K * ESZ;             // bytes of one pixel's labels
    static constexpr int W32 = (K + 1) / 2;            // cost dwords per lane and step
    // A lane's U steps of costs: bytes (byte volumes, 16 bytes), 16-bit costs
    // (K = 1, 32 bytes) or W32 dwords per step (64 bytes); read as CQ 16-byte
    // pieces.
    static constexpr bool PACK16 = K == 1 && !NARROW;
    static constexpr int CQ = NARROW ? (PK ? 2 : 1) : (PACK16 ? 2 : 4);
    static_assert(U * (NARROW ? (PK ? 2 : 1) : (PACK16 ? 2 : 4 * W32)) == CQ * 16,
                  "a compute lane's chunk of costs is CQ pieces");
    // Bytes per compute lane in a cost slot: 16 bytes of padding keep the
    // lanes' 16-byte reads on distinct banks.  Byte slots are unpadded (16-byte
    // reads at a 16-byte stride are conflict-free) and every group of 16 lanes
    // is followed by LPAD bytes: the memory wave writes byte t of a pixel's
    // 16-label piece to lane d0 + t of one group, and without the skew the
    // groups of one line (d0 = 0, 16, 32, 48) or of four lines would hit the
    // same banks at every t.
    static constexpr int LANE = NARROW ? 16 * CQ : CQ * 16 + 16;
    static constexpr int LPAD = NARROW ? 16 : 0;
    static constexpr int CSLOT = 64 * LANE + 4 * LPAD;
    __device__ static constexpr int laneOff(int l) { return l * LANE + (l / 16) * LPAD; }
    // Per-step words, [line][step slot]: the memory wave writes 64 per lane set
    // (two sets for paired lines).
    static constexpr int WSLOT = 64 * 4 * (PK ? 2 : 1);
    static constexpr int ROWH = U * VB + VB;           // horizontal tile row, bytes
    static constexpr int TSLOT = (LPWC * ROWH > U * LPWC * VB) ? LPWC * ROWH : U * LPWC * VB;
    static constexpr int PIECES = LPWC * U * VB / 16;  // 16-byte pieces per chunk
    static_assert(PIECES % 64 == 0, "whole pieces per memory lane");
    static constexpr int NQ = PIECES / 64;
    static constexpr int OFF_W = 3 * CSLOT;
    static constexpr int OFF_T = OFF_W + 3 * WSLOT;
    // Modular refinement step (MODR): per ring slot, two mask words per (line,
    // step) laid out like the per-step words (M0 | MM << 16, then MP: 16-bit
    // lane masks), and per chunk (4 in flight) the rotation of each of the
    // memory wave's pieces.
    static constexpr int OFF_M = OFF_T + 2 * TSLOT;
    static constexpr int MSLOT = MODR ? 2 * WSLOT : 0;
    static constexpr int OFF_R = OFF_M + 3 * MSLOT;
    static constexpr int BYTES = OFF_R + (MODR ? 4 * 64 : 0);
};

// Workgroup barrier ordering LDS only: the memory wave's global loads stay in
// flight across it.
__device__ __forceinline__ void wgBarrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

typedef uint32_t u4v __attribute__((ext_vector_type(4)));

// Orders this wave's own LDS accesses (compiler-level; the hardware executes a
// wave's LDS operations in order).
__device__ __forceinline__ void waveLdsFence()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront", "local");
    asm volatile("" ::: "memory");
}

// The modular refinement step (see runLine): 16-label byte-volume refinement
// lines in the fast form, when the launch asks for it (MOD).
template <int G, int K, int MODE, bool EXACT, bool NARROW, bool PK, bool MOD>
constexpr bool modularRefine()
{
    return MOD && MODE == kRefine && !EXACT && NARROW && !PK && G == 16 && K == 1;
}

// Byte i of a 16-byte vector goes to byte (i - R) & 15: rotation toward lower
// bytes by R (0..15).
__device__ __forceinline__ u4v rotDown16(const u4v& x, int R)
{
    const bool q1 = (R & 4) != 0, q2 = (R & 8) != 0;
    // Two stages of four selects each, on named scalars kept opaque: a select
    // between elements of one vector or array becomes an element read at a
    // variable index, a 4-way select per element (about twice the instructions).
    auto opq = [](uint32_t v) { return (uint32_t)opaqueV((int)v); };
    const uint32_t e0 = opq(x[0]), e1 = opq(x[1]), e2 = opq(x[2]), e3 = opq(x[3]);
    const uint32_t y0 = opq(q1 ? e1 : e0), y1 = opq(q1 ? e2 : e1), y2 = opq(q1 ? e3 : e2),
                   y3 = opq(q1 ? e0 : e3);
    const uint32_t z0 = q2 ? y2 : y0, z1 = q2 ? y3 : y1, z2 = q2 ? y0 : y2, z3 = q2 ? y1 : y3;
    const uint32_t b = (uint32_t)(R & 3);
    return u4v{__builtin_amdgcn_alignbyte(z1, z0, b), __builtin_amdgcn_alignbyte(z2, z1, b),
               __builtin_amdgcn_alignbyte(z3, z2, b), __builtin_amdgcn_alignbyte(z0, z3, b)};
}

// MODE: kSingle, kCoarse or kRefine.
template <int G, int K, int MODE, bool ADAPT, bool EXACT, bool NARROW, bool PK, int DIR,
          bool MOD = false, bool FUSE = false>
__device__ __forceinline__ void runLine(const PathArgs& a, size_t pair, int lineGroup,
                                        const int* s_lut, int* s_xchg, char* s_lds)
{
    constexpr int LPW = 64 / G;
    constexpr int LPWC = PK ? 2 * LPW : LPW;  // lines per wave
    constexpr int GK = G * K;
    constexpr int U = prefetchSteps<K>();
    static_assert(U <= 16, "per-step words use one 16-lane row slot per step");
    constexpr bool VERT = DIR == kTB;
    constexpr bool RL = DIR == kRL;
    constexpr bool PYR = MODE != kSingle;
    constexpr bool REFINE = MODE == kRefine;
    constexpr bool WORDS = REFINE || ADAPT;  // per-step words are loaded
    // Modular refinement step: compute lane g holds the label whose real
    // disparity is congruent to g modulo 16 (see the step below).
    constexpr bool MODR = modularRefine<G, K, MODE, EXACT, NARROW, PK, MOD>();
    // FUSE: R->L lines of a pyramid level whose winner-take-all is done here
    // (the L->R and T->B volumes were written by an earlier launch).
    static_assert(!FUSE || (RL && PYR && NARROW && G == 16 && K == 1 && !EXACT),
                  "the fused winner-take-all runs on 16-label byte-volume R->L lines");
    using TL = Lds<G, K, NARROW, PK, MODR>;
    constexpr int ESZ = TL::ESZ;
    constexpr int VB = GK * ESZ;  // bytes of one pixel's label vector
    constexpr int W32 = TL::W32;

    // Waves 1, 3: memory; 0, 2: recurrence.  Read as a scalar: the compiler
    // does not know that a wave's threads agree.
    const bool memWave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x & 64)) != 0;
    const int lane = threadIdx.x & 63;
    const int j = lane / G;
    const int g = lane - j * G;
    const int r16 = lane & 15;
    const int W = a.W, H = a.H, D = a.D;
    const int nLines = VERT ? W : H;
    const int line0 = lineGroup * LPWC;
    const int nValid = imin(LPWC, nLines - line0);
    int s0 = 0, s1 = VERT ? H : W;
    if constexpr (!PYR && DIR == kLR) {
        s0 = imax(a.startX, 0);
        s1 = imin(a.endX, W);
    }
    if (s0 >= s1) return;
    if (nValid <= 0) {
        // A workgroup's spare line group (past the last line of its direction):
        // both waves keep the workgroup's barriers in step, one per chunk plus
        // the first, as a working pair does.
        const int nc = (s1 - s0 + U - 1) / U;
        for (int c = 0; c <= nc; ++c) wgBarrier();
        return;
    }

    // Stream layouts: offset(row, col) = row * rowPitch + col * elem.
    const uint32_t volRow = (uint32_t)W * VB;
    const uint32_t wRow = REFINE ? (uint32_t)W * 4 : (uint32_t)a.istep;
    constexpr uint32_t wElem = REFINE ? 4 : 1;
    auto lineStride = [&](uint32_t rowPitch, uint32_t elem) { return VERT ? elem : rowPitch; };
    auto stepStride = [&](uint32_t rowPitch, uint32_t elem) { return VERT ? rowPitch : elem; };
    const uint32_t wStep = stepStride(wRow, wElem);
    // This pair's streams.
    const int16_t* const costV = pairAt(a.C, pair, a.batch.ws);
    const int32_t* const labelsV = pairAt(a.labels, pair, a.batch.ws);
    const uint8_t* const imgV = pairAt(a.img, pair, a.batch.left);
    int16_t* const out = pairAt(DIR == kLR ? a.Alr : (DIR == kRL ? a.Arl : a.Atb), pair, a.batch.ws);

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

    const int nChunks = (s1 - s0 + U - 1) / U;
    const int nFull = (s1 - s0) / U;
    auto chunkStart = [&](int k) { return s0 + k * U; };
    auto chunkFull = [&](int k) { return chunkStart(k) + U <= s1; };

    // ---- memory wave ------------------------------------------------------
    // Piece pc (16 bytes) of a chunk: tile row pc / PPR (a horizontal line or a
    // vertical step), byte ob inside that row.  The same mapping addresses the
    // cost loads and the path-cost stores.
    constexpr int PPR = VERT ? LPWC * VB / 16 : U * VB / 16;
    constexpr int NW = PK ? 2 : 1;  // per-step word loads per memory lane
    struct MBuf {
        u4v q[TL::NQ];
        int w[NW];
    };
    auto mload = [&](MBuf& m, int k) {
        const int sc = chunkStart(k);
        const bool full = chunkFull(k);
        const int nSteps = imin(U, s1 - sc);
        const int nRows = VERT ? nSteps : nValid;
        const int nCols = VERT ? nValid : nSteps;
        const ChunkAddr<U> cv = chunk(sc, volRow, VB, VB, full);
        const auto rC = rsrc(reinterpret_cast<const char*>(costV) + cv.origin, cv.nr);
#pragma unroll
        for (int i = 0; i < TL::NQ; ++i) {
            const int pc = lane + 64 * i;
            const int row = pc / PPR;
            const int ob = (pc - row * PPR) * 16;
            const bool ok = row < nRows && ob / VB < nCols;
            const uint32_t go = ok ? (uint32_t)row * volRow + (uint32_t)ob : kOffLimit;
            m.q[i] = __builtin_amdgcn_raw_buffer_load_b128(rC, go, 0, 0);
        }
        if constexpr (WORDS) {
            // Lane (line jm, slot r) loads the word of step sc + r; slots past the
            // chunk or the line, and absent lines, read 0 (offset out of range).
            const ChunkAddr<U> cw = chunk(sc, wRow, wElem, wElem, full);
            const char* base = REFINE ? reinterpret_cast<const char*>(labelsV) : reinterpret_cast<const char*>(imgV);
            const auto rw = rsrc(base + cw.origin, cw.nr);
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                const int jm = (lane >> 4) + 4 * i, r = lane & 15;
                const int rel = RL ? cw.sLast - sc - r : r;
                const bool ok = jm < nValid && r < nSteps;
                const uint32_t voff = ok ? (uint32_t)jm * lineStride(wRow, wElem) + (uint32_t)rel * wStep : kOffLimit;
                m.w[i] = REFINE ? (int)__builtin_amdgcn_raw_buffer_load_b32(rw, voff, 0, 0)
                                : (int)__builtin_amdgcn_raw_buffer_load_b8(rw, voff, 0, 0);
            }
        }
    };
    // Scatter a loaded chunk into cost slot k % 3: label d of step u goes to
    // compute lane (line * G + d / K), dword u * W32 + (d % K) / 2.  MODR: the
    // label-order costs of a pixel are rotated into lane order first, and the
    // step's three lane masks go to the mask slots with the same layout.
    int mcarry = 0;  // MODR: window base of each line's last step so far (all lanes of its row)
    auto mstore = [&](const MBuf& m, int k) {
        const int sc = chunkStart(k);
        const int nSteps = imin(U, s1 - sc);
        const int nRows = VERT ? nSteps : nValid;
        const int nCols = VERT ? nValid : nSteps;
        char* const slot = s_lds + (k % 3) * TL::CSLOT;
        int info = 0;  // MODR: lane (line, r): base & 15 of step r's window
        if constexpr (MODR) {
            const int base = (int)(int16_t)(m.w[0] & 0xFFFF);
            const int prevBase = dppOpaque<kDppRowShr1>(mcarry, base);
            mcarry = rowBcast<U - 1>(base);
            const int sh = base - prevBase;
            const int rot = base & 15;
            info = rot;
            // Lane masks of step r of line j (this lane's word): the previous
            // pixel's path cost for real base + o + e (label o, e = 0, -1, +1)
            // exists iff o + sh + e lies in [0, 15], i.e. label o is in
            //   f(s) = 0xFFFF >> s (s >= 0), (0xFFFF << -s) & 0xFFFF (s < 0),
            //          0 (|s| >= 16),  s = sh + e,
            // the bits of 0xFFFF0000 >> (16 + s) for s clamped to [-16, 16].
            // Label o is held by compute lane (o + rot) & 15: both 16-bit halves
            // of lanes(s) are f(s) rotated left by rot (f duplicated into both
            // halves, then a 32-bit rotation right by 16 - rot).
            auto lanes = [&](int e) {
                const uint32_t k = (uint32_t)imin(imax(sh + e + 16, 0), 32);
                const uint32_t f = (uint32_t)(0xFFFF0000ull >> k);
                const uint32_t d = __builtin_amdgcn_perm(f, f, 0x01000100u);
                return __builtin_amdgcn_alignbit(d, d, (uint32_t)(16 - rot));
            };
            char* const pm = s_lds + TL::OFF_M + (k % 3) * TL::MSLOT + lane * 4;
            // M0 | MM << 16, and MP (its high half repeats the low one; the
            // recurrence reads bits 0-15 only).
            *reinterpret_cast<uint32_t*>(pm) = (lanes(0) & 0xFFFFu) | (lanes(-1) & 0xFFFF0000u);
            *reinterpret_cast<uint32_t*>(pm + TL::WSLOT) = lanes(1);
        }
#pragma unroll
        for (int i = 0; i < TL::NQ; ++i) {
            const int pc = lane + 64 * i;
            const int row = pc / PPR;
            const int ob = (pc - row * PPR) * 16;
            const int col = ob / VB;
            int pinfo = 0;
            if constexpr (MODR) {
                // Rotation of this piece's (line, step), from the lane that
                // loaded that step's word (every lane takes part).
                const int jj = VERT ? col : row;
                const int u = VERT ? row : (RL ? nSteps - 1 - col : col);
                const int src = imin(imax(jj, 0), LPW - 1) * 16 + imin(imax(u, 0), 15);
                pinfo = __builtin_amdgcn_ds_bpermute(src << 2, info);
                *reinterpret_cast<uint8_t*>(s_lds + TL::OFF_R + (k & 3) * 64 + pc) = (uint8_t)pinfo;
            }
            if (row < nRows && col < nCols) {
                const int jj = VERT ? col : row;
                const int u = VERT ? row : (RL ? nSteps - 1 - col : col);
                const int d0 = (ob - col * VB) / ESZ;
                // Paired lines: line jj goes to half jj / LPW of lane (jj % LPW, label).
                char* const p = slot + TL::laneOff((jj % LPW) * G + d0 / K) + (jj / LPW) * 16;
                u4v q = m.q[i];
                if constexpr (MODR) {
                    // Lane g takes label (g - rot) & 15, the label whose real
                    // disparity (base + label) is congruent to g modulo 16.
                    q = rotDown16(q, (16 - pinfo) & 15);
                }
                if constexpr (NARROW) {
#pragma unroll
                    for (int t = 0; t < 16; ++t)
                        *reinterpret_cast<uint8_t*>(p + t * TL::LANE + u) =
                            (uint8_t)(q[t >> 2] >> (8 * (t & 3)));
                } else if constexpr (K == 1) {
#pragma unroll
                    for (int t = 0; t < 8; ++t)
                        *reinterpret_cast<uint16_t*>(p + t * TL::LANE + u * 2) =
                            (uint16_t)((t & 1) ? (q[t >> 1] >> 16) : q[t >> 1]);
                } else if constexpr (K == 2) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) *reinterpret_cast<uint32_t*>(p + t * TL::LANE + u * 4) = q[t];
                } else if constexpr (K == 4) {
                    typedef uint32_t u2v __attribute__((ext_vector_type(2)));
                    *reinterpret_cast<u2v*>(p + u * 8) = u2v{q[0], q[1]};
                    *reinterpret_cast<u2v*>(p + TL::LANE + u * 8) = u2v{q[2], q[3]};
                } else {
                    *reinterpret_cast<u4v*>(p + u * W32 * 4 + ((d0 % K) / 2) * 4) = q;
                }
            }
        }
        if constexpr (WORDS) {
#pragma unroll
            for (int i = 0; i < NW; ++i)
                *reinterpret_cast<int*>(s_lds + TL::OFF_W + (k % 3) * TL::WSLOT + (64 * i + lane) * 4) = m.w[i];
        }
    };
    // Write path-cost tile k % 2 to the volume; pieces of absent lines or of
    // steps past the line get an out-of-range offset and are dropped.
    auto flushTile = [&](int k) {
        const int sc = chunkStart(k);
        const bool full = chunkFull(k);
        const int nSteps = imin(U, s1 - sc);
        const int nRows = VERT ? nSteps : nValid;
        const int nCols = VERT ? nValid : nSteps;
        const char* const tile = s_lds + TL::OFF_T + (k & 1) * TL::TSLOT;
        const ChunkAddr<U> co = chunk(sc, volRow, VB, VB, full);
        const auto rOut = rsrc(reinterpret_cast<char*>(out) + co.origin, co.nr);
#pragma unroll
        for (int i = 0; i < TL::NQ; ++i) {
            const int pc = lane + 64 * i;
            const int row = pc / PPR;
            const int ob = (pc - row * PPR) * 16;
            const int tbyte = row * (VERT ? LPWC * VB : TL::ROWH) + ob;
            u4v v = *reinterpret_cast<const u4v*>(tile + tbyte);
            if constexpr (MODR) {
                // Lane order back to label order.
                const int rot = *reinterpret_cast<const uint8_t*>(s_lds + TL::OFF_R + (k & 3) * 64 + pc);
                v = rotDown16(v, rot);
            }
            const bool ok = row < nRows && ob / VB < nCols;
            const uint32_t go = ok ? (uint32_t)row * volRow + (uint32_t)ob : kOffLimit;
            __builtin_amdgcn_raw_buffer_store_b128(v, rOut, go, 0, 0);
        }
    };

    // ---- fused winner-take-all (FUSE) -----------------------------------------
    // Instead of storing tile k, the memory wave finishes its pixels: piece pc
    // of the tile is one pixel's 16 R->L costs (row pc / PPR = line, column
    // (pc % PPR) in address order), and the pixel's L->R and T->B vectors and
    // label word are loaded into register set k % 2 two iterations earlier.
    struct ABuf {
        u4v l[FUSE ? TL::NQ : 1], t[FUSE ? TL::NQ : 1];
        int w[FUSE ? TL::NQ : 1];
    };
    auto aload = [&](ABuf& b, int k) {
        if constexpr (FUSE) {
            const int sc = chunkStart(k);
            const bool full = chunkFull(k);
            const int nSteps = imin(U, s1 - sc);
            const ChunkAddr<U> co = chunk(sc, volRow, VB, VB, full);
            const auto rL = rsrc(reinterpret_cast<const char*>(pairAt(a.Alr, pair, a.batch.ws)) + co.origin, co.nr);
            const auto rT = rsrc(reinterpret_cast<const char*>(pairAt(a.Atb, pair, a.batch.ws)) + co.origin, co.nr);
            const u4v none = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};  // path not run
#pragma unroll
            for (int i = 0; i < TL::NQ; ++i) {
                const int pc = lane + 64 * i;
                const int row = pc / PPR;
                const int ob = (pc - row * PPR) * 16;
                const bool ok = row < nValid && ob / VB < nSteps;
                const uint32_t go = ok ? (uint32_t)row * volRow + (uint32_t)ob : kOffLimit;
                b.l[i] = a.numPaths >= 2 ? __builtin_amdgcn_raw_buffer_load_b128(rL, go, 0, 0) : none;
                b.t[i] = a.numPaths >= 3 ? __builtin_amdgcn_raw_buffer_load_b128(rT, go, 0, 0) : none;
                if constexpr (REFINE) {
                    const ChunkAddr<U> cw = chunk(sc, (uint32_t)W * 4, 4, 4, full);
                    const auto rW = rsrc(reinterpret_cast<const char*>(labelsV) + cw.origin, cw.nr);
                    const uint32_t wo = ok ? (uint32_t)row * (uint32_t)W * 4 + (uint32_t)(ob / VB) * 4 : kOffLimit;
                    b.w[i] = (int)__builtin_amdgcn_raw_buffer_load_b32(rW, wo, 0, 0);
                } else {
                    b.w[i] = 0;
                }
            }
        } else {
            (void)b;
            (void)k;
        }
    };
    auto wtaTile = [&](const ABuf& b, int k) {
        if constexpr (FUSE) {
            const int sc = chunkStart(k);
            const bool full = chunkFull(k);
            const int nSteps = imin(U, s1 - sc);
            const int sLast = full ? sc + U - 1 : imin(sc + U - 1, s1 - 1);
            const int col0 = W - 1 - sLast;  // the chunk's lowest column (R->L runs right to left)
            const char* const tile = s_lds + TL::OFF_T + (k & 1) * TL::TSLOT;
            int16_t* const dispV = pairAt(a.disp, pair, a.batch.ws);
            float* const confV = pairAt(a.conf, pair, a.batch.ws);
            uint32_t* const keysV = pairAt(a.keys, pair, a.batch.ws);
#pragma unroll
            for (int i = 0; i < TL::NQ; ++i) {
                const int pc = lane + 64 * i;
                const int row = pc / PPR;
                const int ob = (pc - row * PPR) * 16;
                u4v v = *reinterpret_cast<const u4v*>(tile + row * TL::ROWH + ob);
                if constexpr (MODR) {
                    const int rot = *reinterpret_cast<const uint8_t*>(s_lds + TL::OFF_R + (k & 3) * 64 + pc);
                    v = rotDown16(v, rot);
                }
                const int col = ob / VB;
                if (row < nValid && col < nSteps) {
                    const int x = col0 + col, y = line0 + row;
                    const uint32_t wr[4] = {v[0], v[1], v[2], v[3]};
                    const uint32_t wl[4] = {b.l[i][0], b.l[i][1], b.l[i][2], b.l[i][3]};
                    const uint32_t wt[4] = {b.t[i][0], b.t[i][1], b.t[i][2], b.t[i][3]};
                    const wta::Pick pk = wta::searchNarrow<1>(wr, wl, wt, 0, D);
                    wta::resolve<MODE>(a, pk, x, y, (size_t)y * W + x, b.w[i], dispV, confV, keysV);
                }
            }
        } else {
            (void)b;
            (void)k;
        }
    };
    // Tile k leaves the workgroup: stored, or (FUSE) finished from set k % 2.
    auto finish = [&](const ABuf& A0, const ABuf& A1, int k) {
        if constexpr (FUSE) {
            if (k & 1) wtaTile(A1, k);
            else wtaTile(A0, k);
        } else {
            (void)A0;
            (void)A1;
            flushTile(k);
        }
    };

    if (memWave) {
        // Chunk c is loaded into register set c % NS and scattered into LDS
        // slot c % 3 in iteration c - 2 (between barriers c - 2 and c - 1), so
        // its loads are issued NS iterations before they are needed.
        // Iteration k fills slot k+2, reloads that register set with chunk
        // k+2+NS and stores tile k-1.  FUSE: iteration k finishes chunk k-1
        // from pixel-vector set (k-1) % 2 and then loads chunk k+1's vectors
        // into that set.
        constexpr int NS = kMemSets;
        MBuf R[NS];
#pragma unroll
        for (int c = 0; c < NS; ++c)
            if (c < nChunks) mload(R[c], c);
        mstore(R[0], 0);
        if (NS < nChunks) mload(R[0], NS);
        if (nChunks > 1) mstore(R[1], 1);
        if (NS + 1 < nChunks) mload(R[1], NS + 1);
        ABuf A0, A1;
        aload(A0, 0);
        wgBarrier();
        // Iteration m with m % NS == I (a compile-time constant, so that the
        // register sets stay registers).
        auto iter = [&](auto ic, int m) {
            constexpr int I = decltype(ic)::value;
            constexpr int S = (I + 2) % NS;
            if (m + 2 < nChunks) mstore(R[S], m + 2);
            if (m + 2 + NS < nChunks) mload(R[S], m + 2 + NS);
            if (m >= 1) finish(A0, A1, m - 1);
            if (m + 1 < nChunks) {
                if constexpr (I % 2 == 0)
                    aload(A1, m + 1);
                else
                    aload(A0, m + 1);
            }
            wgBarrier();
        };
        int k = 0;
        for (; k + NS <= nChunks; k += NS) {
            auto sub = [&](auto ic) { iter(ic, k + decltype(ic)::value); };
            unrollSteps<NS>(sub);
        }
        auto tail = [&](auto ic) {
            if (k + decltype(ic)::value < nChunks) iter(ic, k + decltype(ic)::value);
        };
        unrollSteps<NS>(tail);
        finish(A0, A1, nChunks - 1);
        return;
    }

    // The recurrence wave shares its SIMD with another workgroup's memory
    // wave, which runs chunks ahead; when both have an instruction ready the
    // recurrence's dependent chain goes first.
    __builtin_amdgcn_s_setprio(3);

    // Path state of this lane's labels: complemented H (fast form) or L with
    // the reference's activity flag (EXACT).
    int Hs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) Hs[k] = 0;
    int pmaxH = 0;                 // fast form: gmaxPartial of Hs (see lastPair)
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
            // Only this wave uses s_xchg, and one wave's LDS operations complete
            // in issue order: the fences just keep the compiler from moving the
            // exchange's reads and writes across each other.
            waveLdsFence();
#pragma unroll
            for (int k = 0; k < K; ++k) s_xchg[lane * K + k] = prev[k];
            waveLdsFence();
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
    // read on moved steps only); bigMask bit u: some line shifts by more than 1;
    // mw: MODR mask words of the chunk (lane r16 of a line's row holds step
    // r16's).
    //
    // MODR: lane g holds the label whose real disparity r is congruent to g
    // modulo 16, so a window move changes no register: the previous pixel's
    // path cost at real r is this lane's own value and those at r - 1, r + 1
    // are the row rotations by one lane, each valid iff that real was in the
    // previous pixel's window (masks m0, mm, mp: bits of two words per line
    // and step that the memory wave computes).  Invalid operands read 0 ("no path"), which is what
    // sgm_kernels.cu:1237-1243 gives them (SHRT_MAX before the complement).
    auto step = [&](auto uc, const uint32_t (&cw)[(K + 1) / 2], int p2, bool moved,
                    uint32_t bigMask, int shv, char* tp, const int* mw) {
        constexpr int u = decltype(uc)::value;
        int e[K];
#pragma unroll
        for (int k = 0; k < K; ++k) e[k] = unpk<K>(cw, k);
        int Hc[K];
        if constexpr (!EXACT) {
            // Neighbours of an unmoved window; a moved one overrides them (one
            // uniform branch per step, taken on the rare moved steps).
            int h0[K], y[K];
            int lowN = 0, upN = 0;
            if constexpr (MODR) {
                (void)moved;
                (void)bigMask;
                (void)shv;
                // Step u's mask words from lane u of the line's row; bit g
                // sign-extended is this lane's 0 / -1 mask.
                const int w01 = rowBcast<u>(mw[0]), wp = rowBcast<u>(mw[1]);
                const int m0 = __builtin_amdgcn_sbfe(w01, g, 1);
                const int mm = __builtin_amdgcn_sbfe(w01, g + 16, 1);
                const int mp = __builtin_amdgcn_sbfe(wp, g, 1);
                lowN = dmov<kDppRowRor1>(Hs[0]) & mm;   // v_and_b32_dpp
                upN = dmov<kDppRowRor15>(Hs[0]) & mp;
                h0[0] = Hs[0] & m0;
                y[0] = imax(lowN, upN);
            } else if constexpr (G == 16 && K == 1) {
                // y = v_max_i32_dpp(Hs >> 1 row, Hs << 1 row); a moved step
                // forms its own neighbours.
                h0[0] = Hs[0];
                y[0] = opaqueV(imax(lowerOr0<G>(Hs[0], g), upperOr0<G>(Hs[0], g)));
            } else {
                lowN = lowerOr0<G>(Hs[K - 1], g);
                upN = upperOr0<G>(Hs[0], g);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    h0[k] = Hs[k];
                    y[k] = imax(k > 0 ? Hs[k - 1] : lowN, k < K - 1 ? Hs[k + 1] : upN);
                }
            }
            if (!MODR && REFINE && __builtin_expect(moved, 0)) {
                if (G == 16 && K == 1 && !((bigMask >> u) & 1u)) {
                    // Every line of the wave moved by -1, 0 or +1 labels: the
                    // re-indexed values are row shifts of the previous ones (zero
                    // outside the 16 labels), selected per line.
                    const int sh = rowBcast<u>(shv);
                    // (Its own moves: the unmoved step's fold into its maximum.)
                    const int hv = opaqueV(Hs[0]);
                    lowN = lowerOr0<G>(hv, g);
                    upN = upperOr0<G>(hv, g);
                    const int tm2 = dmov<kDppRowShr2>(hv), tp2 = dmov<kDppRowShl2>(hv);
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
            // H'[d] + f and max(H'[d-1], H'[d+1]) + f - P1 share the term
            // f = E - max_d H': Hc = max(max(h0, y - P1) + E - max H', E - P2'),
            // and the last level of the line maximum folds into the subtraction.
            int ma, mb;
            lastPair<G>(pmaxH, ma, mb);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int ze = imax(h0[k], y[k] - P1) + e[k];
                const int t3 = (int)__builtin_elementwise_sub_sat((uint32_t)e[k], (uint32_t)p2);
                Hc[k] = imax(imin(ze - ma, ze - mb), t3);
            }
            int m = Hc[0];
#pragma unroll
            for (int k = 1; k < K; ++k) m = imax(m, Hc[k]);
            // (Formed here, in this step's block, so that the last level's move
            // folds into its maximum.)
            if constexpr (G == 64)
                pmaxH = opaqueS(gmaxPartial<G>(m));
            else
                pmaxH = opaqueV(gmaxPartial<G>(m));
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
        tstore<K, NARROW>(Hc, tp);
    };

    // ---- compute wave --------------------------------------------------------
    struct CBuf {
        u4v q[TL::CQ];
        int w;
        int w2;  // paired lines: the second line's word
        int mw[MODR ? 2 : 1];  // MODR mask words of step r16 of line j (M0 | MM << 16, MP)
        // Formed from w by prep(), half a chunk before the chunk runs (the
        // P2' table read then completes while the previous chunk computes):
        int p2v;               // ADAPT: this lane's step's P2' (lane r16 = step r16)
        int shv;               // shifting refinement step: window shift of step r16
        uint32_t moveMask, bigMask;  // bit u: some line moves at step u / by more than one
    };
    // This lane's costs of chunk k (and its row's per-step words) from LDS.
    auto cread = [&](CBuf& c, int k) {
        const char* const p = s_lds + (k % 3) * TL::CSLOT + TL::laneOff(lane);
#pragma unroll
        for (int i = 0; i < TL::CQ; ++i) c.q[i] = *reinterpret_cast<const u4v*>(p + 16 * i);
        if constexpr (MODR) {
            const char* const pm = s_lds + TL::OFF_M + (k % 3) * TL::MSLOT + (j * 16 + r16) * 4;
            c.mw[0] = *reinterpret_cast<const int*>(pm);
            c.mw[1] = *reinterpret_cast<const int*>(pm + TL::WSLOT);
        }
        if constexpr (WORDS) {
            const char* const wbase = s_lds + TL::OFF_W + (k % 3) * TL::WSLOT;
            c.w = *reinterpret_cast<const int*>(wbase + (j * 16 + r16) * 4);
            if constexpr (PK) c.w2 = *reinterpret_cast<const int*>(wbase + ((j + LPW) * 16 + r16) * 4);
        }
    };

    // ---- paired lines (byte volumes, 16 labels): lane (j, g) holds label g of
    // line j in bits 0-15 and of line j + LPW in bits 16-31 of every packed
    // value; the step is the fast form above in 16-bit packed arithmetic.
    typedef int16_t s2v __attribute__((ext_vector_type(2)));
    typedef uint16_t u2v __attribute__((ext_vector_type(2)));
    auto S2 = [](uint32_t x) { return __builtin_bit_cast(s2v, x); };
    auto U2 = [](uint32_t x) { return __builtin_bit_cast(u2v, x); };
    auto Wd = [](auto x) { return __builtin_bit_cast(uint32_t, x); };
    auto pmax = [&](uint32_t x, uint32_t y) { return Wd(__builtin_elementwise_max(S2(x), S2(y))); };
    auto psub = [&](uint32_t x, uint32_t y) { return Wd((s2v)(S2(x) - S2(y))); };
    auto padd = [&](uint32_t x, uint32_t y) { return Wd((s2v)(S2(x) + S2(y))); };
    auto bsel = [](uint32_t m, uint32_t x, uint32_t y) { return (x & m) | (y & ~m); };
    uint32_t HsP = 0, maxHP = 0;
    int carryBase2 = a.minD, carryIm42 = 0;
    const uint32_t P1P = (uint32_t)(uint16_t)P1 | ((uint32_t)(uint16_t)P1 << 16);

    auto stepPK = [&](auto uc, uint32_t e2, uint32_t p2, bool moved, uint32_t bigMask, int shvA,
                      int shvB, char* tpA, char* tpB) {
        constexpr int u = decltype(uc)::value;
        const uint32_t lowN = (uint32_t)dmov<kDppRowShr1>((int)HsP);
        const uint32_t upN = (uint32_t)dmov<kDppRowShl1>((int)HsP);
        uint32_t h0 = HsP, y = pmax(lowN, upN);
        if (REFINE && __builtin_expect(moved, 0)) {
            const int shA = rowBcast<u>(shvA), shB = rowBcast<u>(shvB);
            if (!((bigMask >> u) & 1u)) {
                // Shifts of -1, 0 or +1: row shifts of the previous values, per half.
                const uint32_t tm2 = (uint32_t)dmov<kDppRowShr2>((int)HsP);
                const uint32_t tp2 = (uint32_t)dmov<kDppRowShl2>((int)HsP);
                const uint32_t mN = (shA < 0 ? 0xFFFFu : 0u) | (shB < 0 ? 0xFFFF0000u : 0u);
                const uint32_t mP = (shA > 0 ? 0xFFFFu : 0u) | (shB > 0 ? 0xFFFF0000u : 0u);
                h0 = bsel(mN, lowN, bsel(mP, upN, HsP));
                const uint32_t lm = bsel(mN, tm2, bsel(mP, HsP, lowN));
                const uint32_t lp = bsel(mN, HsP, bsel(mP, tp2, upN));
                y = pmax(lm, lp);
            } else {
                // Re-index each half through ds_bpermute (out-of-window labels read 0).
                const int lane0 = j * G;
                auto src = [&](int sm) { return (lane0 + imin(imax(sm, 0), D - 1)) << 2; };
                auto pick = [](uint32_t x, bool vx, uint32_t z, bool vz) {
                    return (vx ? (x & 0xFFFFu) : 0u) | (vz ? (z & 0xFFFF0000u) : 0u);
                };
                const int smA = g + shA, smB = g + shB;
                const uint32_t a0 = (uint32_t)__builtin_amdgcn_ds_bpermute(src(smA), (int)HsP);
                const uint32_t am = (uint32_t)__builtin_amdgcn_ds_bpermute(src(smA - 1), (int)HsP);
                const uint32_t ap = (uint32_t)__builtin_amdgcn_ds_bpermute(src(smA + 1), (int)HsP);
                const uint32_t b0 = (uint32_t)__builtin_amdgcn_ds_bpermute(src(smB), (int)HsP);
                const uint32_t bm = (uint32_t)__builtin_amdgcn_ds_bpermute(src(smB - 1), (int)HsP);
                const uint32_t bp = (uint32_t)__builtin_amdgcn_ds_bpermute(src(smB + 1), (int)HsP);
                h0 = pick(a0, smA >= 0 && smA < D, b0, smB >= 0 && smB < D);
                const uint32_t lm = pick(am, smA >= 1 && smA <= D, bm, smB >= 1 && smB <= D);
                const uint32_t lp = pick(ap, smA >= -1 && smA < D - 1, bp, smB >= -1 && smB < D - 1);
                y = pmax(lm, lp);
            }
        }
        const uint32_t ze = padd(pmax(h0, psub(y, P1P)), e2);
        const uint32_t t3 = Wd(__builtin_elementwise_sub_sat(U2(e2), U2(p2)));
        const uint32_t Hc = pmax(psub(ze, maxHP), t3);
        uint32_t m = Hc;
        m = pmax(m, (uint32_t)dmov<kDppQuadXor1>((int)m));
        m = pmax(m, (uint32_t)dmov<kDppQuadXor2>((int)m));
        m = pmax(m, (uint32_t)dmov<kDppRowHalfMirror>((int)m));
        m = pmax(m, (uint32_t)dmov<kDppRowMirror>((int)m));
        maxHP = m;
        HsP = Hc;
        *reinterpret_cast<uint8_t*>(tpA) = (uint8_t)Hc;
        *reinterpret_cast<uint8_t*>(tpB) = (uint8_t)(Hc >> 16);
    };

    auto processPK = [&](const CBuf& c, int k, auto fullC) {
        constexpr bool full = decltype(fullC)::value;
        const int sc = chunkStart(k);
        int shvA = 0, shvB = 0;
        uint32_t moveMask = 0, bigMask = 0, p2v = 0;
        if constexpr (REFINE) {
            const int baseA = (int)(int16_t)(c.w & 0xFFFF), baseB = (int)(int16_t)(c.w2 & 0xFFFF);
            shvA = baseA - dpp<kDppRowShr1>(carryBase, baseA);
            shvB = baseB - dpp<kDppRowShr1>(carryBase2, baseB);
            auto fold = [](uint64_t b) {
                const uint64_t f2 = b | (b >> 32);
                return (uint32_t)(f2 | (f2 >> 16)) & 0xFFFFu;
            };
            moveMask = fold(__ballot(shvA != 0 || shvB != 0));
            bigMask = fold(__ballot(shvA > 1 || shvA < -1 || shvB > 1 || shvB < -1));
            carryBase = rowBcast<U - 1>(baseA);
            carryBase2 = rowBcast<U - 1>(baseB);
        }
        if constexpr (ADAPT) {
            auto lut = [&](int w, int& carry) {
                const int im4 = REFINE ? (int)((uint32_t)w >> 22) : w << 2;
                const int prevIm4 = dpp<kDppRowShr1>(carry, im4);
                const uint32_t d4 = __builtin_amdgcn_sad_u16((uint32_t)im4, (uint32_t)prevIm4, 0u);
                carry = rowBcast<U - 1>(im4);
                return (uint32_t)*reinterpret_cast<const int*>(reinterpret_cast<const char*>(s_lut) + d4);
            };
            const uint32_t pa = lut(c.w, carryIm4), pb = lut(c.w2, carryIm42);
            p2v = (pa & 0xFFFFu) | (pb << 16);
        }
        const uint32_t p2c = (uint32_t)(uint16_t)a.P2 | ((uint32_t)(uint16_t)a.P2 << 16);
        const int nSteps = full ? U : imin(U, s1 - sc);
        char* const tile = s_lds + TL::OFF_T + (k & 1) * TL::TSLOT;
        auto run = [&](auto uc) {
            constexpr int u = decltype(uc)::value;
            if (full || u < nSteps) {
                const uint32_t p2 = ADAPT ? (uint32_t)rowBcast<u>((int)p2v) : p2c;
                const bool moved = REFINE && ((moveMask >> u) & 1u);
                // Cost bytes of step u of both lines -> 16-bit fields.
                constexpr uint32_t sel = 0x0C000C00u | ((4u + (u & 3)) << 16) | (u & 3);
                // (CQ is 2 whenever this runs; the index only keeps the
                // never-called non-paired instantiation in bounds.)
                constexpr int q1 = TL::CQ > 1 ? 1 : 0;
                const uint32_t e2 = __builtin_amdgcn_perm(c.q[q1][u >> 2], c.q[0][u >> 2], sel);
                int tbA, tbB;
                if constexpr (VERT) {
                    tbA = (u * LPWC + j) * VB + g;
                    tbB = (u * LPWC + j + LPW) * VB + g;
                } else {
                    const int su = RL ? (full ? U - 1 - u : nSteps - 1 - u) : u;
                    tbA = j * TL::ROWH + su * VB + g;
                    tbB = (j + LPW) * TL::ROWH + su * VB + g;
                }
                stepPK(uc, e2, p2, moved, bigMask, shvA, shvB, tile + tbA, tile + tbB);
            }
        };
        unrollSteps<U>(run);
    };

    // Per-chunk values from the chunk's per-step words (chunks in order: the
    // carries hold the previous chunk's last step).
    auto prep = [&](CBuf& c) {
        c.p2v = 0;
        c.shv = 0;
        c.moveMask = 0;
        c.bigMask = 0;
        if constexpr (REFINE && !MODR && !PK) {
            const int base = (int)(int16_t)(c.w & 0xFFFF);
            const int prevBase = dpp<kDppRowShr1>(carryBase, base);
            c.shv = base - prevBase;
            // Bit u: some line of the wave moves at step u (any shift / |shift| > 1).
            auto fold = [](uint64_t b) {
                const uint64_t f2 = b | (b >> 32);
                return (uint32_t)(f2 | (f2 >> 16)) & 0xFFFFu;
            };
            c.moveMask = fold(__ballot(c.shv != 0));
            c.bigMask = fold(__ballot(c.shv > 1 || c.shv < -1));
            carryBase = rowBcast<U - 1>(base);
        }
        if constexpr (ADAPT && !PK) {
            const int im4 = REFINE ? (int)((uint32_t)c.w >> 22) : c.w << 2;
            const int prevIm4 = dpp<kDppRowShr1>(carryIm4, im4);
            const uint32_t d4 = __builtin_amdgcn_sad_u16((uint32_t)im4, (uint32_t)prevIm4, 0u);
            c.p2v = *reinterpret_cast<const int*>(reinterpret_cast<const char*>(s_lut) + d4);
            carryIm4 = rowBcast<U - 1>(im4);
        }
    };

    // Process chunk k; full (a compile-time constant, so that full chunks carry
    // no per-step guard): all U steps exist.  mid() runs after the first half
    // of the steps (the next chunk's prep).
    auto process = [&](const CBuf& c, int k, auto fullC, auto&& mid) {
        constexpr bool full = decltype(fullC)::value;
        const int sc = chunkStart(k);
        const int p2v = c.p2v, shv = c.shv;
        const uint32_t moveMask = c.moveMask, bigMask = c.bigMask;
        const int nSteps = full ? U : imin(U, s1 - sc);
        char* const tile = s_lds + TL::OFF_T + (k & 1) * TL::TSLOT;
        auto run = [&](auto uc) {
            constexpr int u = decltype(uc)::value;
            if (full || u < nSteps) {
                const int p2 = ADAPT ? rowBcast<u>(p2v) : a.P2;
                const bool moved = REFINE && ((moveMask >> u) & 1u);
                uint32_t cw[W32];
                if constexpr (NARROW) {
                    cw[0] = (c.q[0][u >> 2] >> (8 * (u & 3))) & 0xFFu;
                } else if constexpr (TL::PACK16) {
                    const uint32_t x = c.q[u >> 3][(u >> 1) & 3];
                    cw[0] = (u & 1) ? (x >> 16) : (x & 0xFFFFu);
                } else {
#pragma unroll
                    for (int w = 0; w < W32; ++w) cw[w] = c.q[(u * W32 + w) >> 2][(u * W32 + w) & 3];
                }
                // Tile slot of step u: horizontal rows run in address order (R->L
                // reversed), the vertical tile is step-major.
                int tb;
                if constexpr (VERT)
                    tb = (u * LPW + j) * VB + g * K * ESZ;
                else
                    tb = j * TL::ROWH + (RL ? (full ? U - 1 - u : nSteps - 1 - u) : u) * VB + g * K * ESZ;
                step(uc, cw, p2, moved, bigMask, shv, tile + tb, c.mw);
            }
        };
        unrollSteps<U / 2>(run);
        mid();
        unrollSteps<U, U / 2>(run);
    };

    auto proc = [&](const CBuf& c, int k, auto fullC, auto&& mid) {
        if constexpr (PK) {
            processPK(c, k, fullC);
            mid();
        } else {
            process(c, k, fullC, mid);
        }
    };
    // Chunk k+1 is read from LDS while chunk k is processed, and prepared
    // half-way through it; barrier k+1 ends chunk k (its tile complete, slot k
    // free for the memory wave).
    wgBarrier();
    CBuf cA, cB;
    cread(cA, 0);
    prep(cA);
    int k = 0;
    for (; k + 2 <= nFull; k += 2) {
        cread(cB, k + 1);
        proc(cA, k, std::true_type{}, [&] { prep(cB); });
        wgBarrier();
        if (k + 2 < nChunks) cread(cA, k + 2);
        proc(cB, k + 1, std::true_type{}, [&] {
            if (k + 2 < nChunks) prep(cA);
        });
        wgBarrier();
    }
    for (; k < nChunks; ++k) {
        if (k + 1 < nChunks) cread(cB, k + 1);
        auto next = [&] {
            if (k + 1 < nChunks) prep(cB);
        };
        if (chunkFull(k))
            proc(cA, k, std::true_type{}, next);
        else
            proc(cA, k, std::false_type{}, next);
        wgBarrier();
        cA = cB;
    }
}

// T->B block b of n runs column group xcdRemap(b, n): the blocks of one XCD
// get a contiguous range of groups, so that neighbouring column groups, which
// read parts of the same 128-byte lines at every step, share one L2 instead of
// each XCD fetching the line.  (Row paths read whole lines per chunk and keep
// the dispatch order, which spreads their long lines evenly over the XCDs.)

// Line groups (recurrence + memory wave pairs) per workgroup of kPaths.  The
// byte-volume launches run two: waves 0 and 2 are the recurrences, and the
// hardware places a workgroup's even waves on different SIMDs, so no two
// recurrence waves share a SIMD (with one group per workgroup, a few percent
// of them did, ran at half speed and set the launch's length).
template <bool NARROW, bool PK>
constexpr int pathGroupsPerBlock() { return NARROW && !PK ? 2 : 1; }

template <int G, int K, int MODE, bool ADAPT, bool EXACT, bool NARROW, bool PK = false,
          bool MOD = false>
__global__ __launch_bounds__((128 * pathGroupsPerBlock<NARROW, PK>())) void kPaths(PathArgs a)
{
    constexpr int GPB = pathGroupsPerBlock<NARROW, PK>();
    using TL = Lds<G, K, NARROW, PK, modularRefine<G, K, MODE, EXACT, NARROW, PK, MOD>()>;
    static_assert(TL::BYTES % 16 == 0, "each group's LDS region starts 16-byte aligned");
    __shared__ int s_lut[256];
    __shared__ int s_xchg[GPB][MODE == kRefine ? 64 * K : 1];
    __shared__ __attribute__((aligned(16))) char s_lds[GPB][TL::BYTES];
    if constexpr (ADAPT) {
        for (int i = threadIdx.x; i < 256; i += 128 * GPB) s_lut[i] = a.lut[i];
        __syncthreads();
    }
    // Block b runs line groups b * GPB + h of one direction (h = threadIdx.x /
    // 128); a.nLRBlocks / a.nRLBlocks count blocks.
    const int b = blockIdx.x;
    // Wave-uniform, read as a scalar so that the line group's addresses (buffer
    // descriptors among them) stay scalar.
    const int h = GPB > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 7)) : 0;
    const size_t pr = blockIdx.y;  // pair
    if (b < a.nLRBlocks)
        runLine<G, K, MODE, ADAPT, EXACT, NARROW, PK, kLR, MOD>(a, pr, b * GPB + h, s_lut, s_xchg[h],
                                                                s_lds[h]);
    else if (b < a.nLRBlocks + a.nRLBlocks)
        runLine<G, K, MODE, ADAPT, EXACT, NARROW, PK, kRL, MOD>(a, pr, (b - a.nLRBlocks) * GPB + h, s_lut,
                                                                s_xchg[h], s_lds[h]);
    else {
        const int t = b - a.nLRBlocks - a.nRLBlocks, nTB = (int)gridDim.x - a.nLRBlocks - a.nRLBlocks;
        runLine<G, K, MODE, ADAPT, EXACT, NARROW, PK, kTB, MOD>(
            a, pr, (a.xcdRemap ? xcdRemap(t, nTB) : t) * GPB + h, s_lut, s_xchg[h], s_lds[h]);
    }
}

// R->L lines of a 16-label byte-volume pyramid level with the winner-take-all
// fused in (after a launch of the level's L->R and T->B lines): block b runs
// line group b of pair blockIdx.y.  PK: paired lines; MOD: modular step.
template <int MODE, bool ADAPT, bool PK, bool MOD>
__global__ __launch_bounds__(128) void kPathsRLW(PathArgs a)
{
    __shared__ int s_lut[256];
    __shared__ int s_xchg[MODE == kRefine ? 64 : 1];
    __shared__ __attribute__((aligned(16))) char
        s_lds[Lds<16, 1, true, PK, modularRefine<16, 1, MODE, false, true, PK, MOD>()>::BYTES];
    if constexpr (ADAPT) {
        for (int i = threadIdx.x; i < 256; i += 128) s_lut[i] = a.lut[i];
        __syncthreads();
    }
    runLine<16, 1, MODE, ADAPT, false, true, PK, kRL, MOD, true>(a, blockIdx.y, blockIdx.x, s_lut, s_xchg,
                                                                s_lds);
}

// Queued form: persistent workgroups take (task, block) items from a device
// counter until the items run out; a task is one pair's level with its own
// geometry, penalties and buffers, numbered blocks as in kPaths.  Tasks are
// placed longest lines first by the host, so the long lines start early and
// short ones fill the tail.
template <int G, int K, int MODE, bool ADAPT, bool EXACT, bool NARROW, bool PK = false>
__global__ __launch_bounds__(128) void kPathsQ(QueueArgs q)
{
    __shared__ int s_lut[256];
    __shared__ int s_xchg[MODE == kRefine ? 64 * K : 1];
    __shared__ __attribute__((aligned(16))) char s_lds[Lds<G, K, NARROW, PK>::BYTES];
    __shared__ int s_item;
    int cur = -1;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&q.counter[0], 1u);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= q.total) break;
        // Task t: the last with start[t] <= item.
        int lo = 0, hi = q.nTasks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (q.start[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const int t = __builtin_amdgcn_readfirstlane(lo);
        const PathArgs& a = q.tasks[t];
        if constexpr (ADAPT) {
            if (t != cur) {
                for (int i = threadIdx.x; i < 256; i += 128) s_lut[i] = a.lut[i];
                __syncthreads();
            }
        }
        cur = t;
        const int b = item - q.start[t];
        const int nLR = a.nLRBlocks, nRL = a.nRLBlocks;
        if (b < nLR)
            runLine<G, K, MODE, ADAPT, EXACT, NARROW, PK, kLR>(a, 0, b, s_lut, s_xchg, s_lds);
        else if (b < nLR + nRL)
            runLine<G, K, MODE, ADAPT, EXACT, NARROW, PK, kRL>(a, 0, b - nLR, s_lut, s_xchg, s_lds);
        else
            runLine<G, K, MODE, ADAPT, EXACT, NARROW, PK, kTB>(a, 0, b - nLR - nRL, s_lut, s_xchg, s_lds);
        // Both waves are done with the item's LDS before s_item / s_lut change.
        __syncthreads();
    }
    // Every workgroup has taken its last item once all have arrived here; the
    // last one to arrive re-arms the counters for the next launch.
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&q.counter[1], 1u) == gridDim.x - 1) {
            atomicExch(&q.counter[0], 0u);
            atomicExch(&q.counter[1], 0u);
        }
    }
}

inline int linesPerWave(int Dp) { return Dp <= 16 ? 4 : (Dp <= 32 ? 2 : 1); }


// The L->R / T->B / R->L lines of one launch (withRL: R->L lines included).
template <int MODE, bool ADAPT, bool EXACT, bool NARROW>
hipError_t launchDirs(const PathArgs& a, bool withRL, hipStream_t s)
{
    const int lpw = linesPerWave(a.Dp);
    PathArgs b = a;
    b.nLRBlocks = a.numPaths >= 2 ? (a.H + lpw - 1) / lpw : 0;
    b.nRLBlocks = withRL ? (a.H + lpw - 1) / lpw : 0;
    const int nTB = a.numPaths >= 3 ? (a.W + lpw - 1) / lpw : 0;
    const int blocks = b.nLRBlocks + b.nRLBlocks + nTB;
    if (blocks == 0) return hipSuccess;
    const dim3 grid(blocks, a.batch.n);
    if constexpr (NARROW && MODE != kSingle) {
        // Paired lines when the launch has at least four 4-line waves per SIMD
        // of the 1024: the recurrence wave is then bound by VALU issue, and two
        // lines per lane in 16-bit halves nearly halve its instructions per line.
        if (a.Dp == 16 && (a.pairedMode == 2 ||
                           (a.pairedMode == 0 && (size_t)blocks * a.batch.n >= kPairedWaves))) {
            PathArgs c = a;
            c.nLRBlocks = a.numPaths >= 2 ? (a.H + 7) / 8 : 0;
            c.nRLBlocks = withRL ? (a.H + 7) / 8 : 0;
            const int nTB8 = a.numPaths >= 3 ? (a.W + 7) / 8 : 0;
            const dim3 grid8(c.nLRBlocks + c.nRLBlocks + nTB8, a.batch.n);
            kPaths<16, 1, MODE, ADAPT, EXACT, true, true><<<grid8, 128, 0, s>>>(c);
            return hipGetLastError();
        }
    }
    if constexpr (NARROW) {
        // Two line groups per block (pathGroupsPerBlock); a direction with an
        // odd number of groups leaves its last block's second group spare.
        constexpr int GPB = pathGroupsPerBlock<true, false>();
        PathArgs b2 = b;
        b2.nLRBlocks = (b.nLRBlocks + GPB - 1) / GPB;
        b2.nRLBlocks = (b.nRLBlocks + GPB - 1) / GPB;
        const dim3 grid2(b2.nLRBlocks + b2.nRLBlocks + (nTB + GPB - 1) / GPB, a.batch.n);
        if constexpr (MODE == kRefine && !EXACT) {
            // The modular step has no per-step branch but a heavier data-moving
            // wave; it is faster up to about 1.5 recurrence waves per SIMD.
            if (a.Dp == 16 && (a.modularMode == 2 ||
                               (a.modularMode == 0 && (size_t)blocks * a.batch.n <= kModularMaxBlocks))) {
                kPaths<16, 1, MODE, ADAPT, EXACT, true, false, true><<<grid2, 128 * GPB, 0, s>>>(b2);
                return hipGetLastError();
            }
        }
        switch (a.Dp) {
        case 16: kPaths<16, 1, MODE, ADAPT, EXACT, true><<<grid2, 128 * GPB, 0, s>>>(b2); break;
        case 32: kPaths<32, 1, MODE, ADAPT, EXACT, true><<<grid2, 128 * GPB, 0, s>>>(b2); break;
        case 64: kPaths<64, 1, MODE, ADAPT, EXACT, true><<<grid2, 128 * GPB, 0, s>>>(b2); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (a.Dp) {
    case 16: kPaths<16, 1, MODE, ADAPT, EXACT, false><<<grid, 128, 0, s>>>(b); break;
    case 32: kPaths<32, 1, MODE, ADAPT, EXACT, false><<<grid, 128, 0, s>>>(b); break;
    case 64: kPaths<64, 1, MODE, ADAPT, EXACT, false><<<grid, 128, 0, s>>>(b); break;
    case 128: kPaths<64, 2, MODE, ADAPT, EXACT, false><<<grid, 128, 0, s>>>(b); break;
    case 256: kPaths<64, 4, MODE, ADAPT, EXACT, false><<<grid, 128, 0, s>>>(b); break;
    default:
        // Refinement levels hold at most 256 labels (the entry point rejects more).
        if constexpr (MODE != kRefine) {
            if (a.Dp == 512) {
                kPaths<64, 8, MODE, ADAPT, EXACT, false><<<grid, 128, 0, s>>>(b);
                break;
            }
            if (a.Dp == 1024) {
                kPaths<64, 16, MODE, ADAPT, EXACT, false><<<grid, 128, 0, s>>>(b);
                break;
            }
        }
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// R->L lines with the fused winner-take-all; paired lines and the modular
// step are chosen by this launch's own size, as in launchDirs.
template <int MODE, bool ADAPT>
hipError_t launchRLW(const PathArgs& a, hipStream_t s)
{
    const size_t rl4 = (size_t)((a.H + 3) / 4) * a.batch.n;
    if (a.pairedMode == 2 || (a.pairedMode == 0 && rl4 >= kPairedWaves)) {
        kPathsRLW<MODE, ADAPT, true, false><<<dim3((a.H + 7) / 8, a.batch.n), 128, 0, s>>>(a);
        return hipGetLastError();
    }
    const dim3 grid((a.H + 3) / 4, a.batch.n);
    if constexpr (MODE == kRefine) {
        if (a.modularMode == 2 || (a.modularMode == 0 && rl4 <= kModularMaxBlocks)) {
            kPathsRLW<MODE, ADAPT, false, true><<<grid, 128, 0, s>>>(a);
            return hipGetLastError();
        }
    }
    kPathsRLW<MODE, ADAPT, false, false><<<grid, 128, 0, s>>>(a);
    return hipGetLastError();
}

template <int MODE, bool ADAPT, bool EXACT, bool NARROW>
hipError_t dispatchPaths(const PathArgs& a, hipStream_t s)
{
    // Fused winner-take-all (the host sets fuseWTA only for 16-label
    // byte-volume pyramid levels in the fast form): the L->R and T->B lines
    // run first, then the R->L lines, which finish every pixel themselves.
    if constexpr (NARROW && MODE != kSingle && !EXACT) {
        if (a.fuseWTA) {
            if (a.Dp != 16) return hipErrorInvalidValue;
            const hipError_t e = launchDirs<MODE, ADAPT, EXACT, NARROW>(a, false, s);
            if (e != hipSuccess) return e;
            return launchRLW<MODE, ADAPT>(a, s);
        }
    }
    return launchDirs<MODE, ADAPT, EXACT, NARROW>(a, true, s);
}

// Queued launches: the kernel of one specialisation, its occupancy and launch.
template <int MODE, bool ADAPT, bool EXACT, bool NARROW>
const void* queuedKernel(int Dp, int paired)
{
    if constexpr (NARROW && MODE != kSingle) {
        if (paired && Dp == 16) return reinterpret_cast<const void*>(&kPathsQ<16, 1, MODE, ADAPT, EXACT, true, true>);
    }
    if constexpr (NARROW) {
        switch (Dp) {
        case 16: return reinterpret_cast<const void*>(&kPathsQ<16, 1, MODE, ADAPT, EXACT, true>);
        case 32: return reinterpret_cast<const void*>(&kPathsQ<32, 1, MODE, ADAPT, EXACT, true>);
        case 64: return reinterpret_cast<const void*>(&kPathsQ<64, 1, MODE, ADAPT, EXACT, true>);
        default: return nullptr;
        }
    }
    switch (Dp) {
    case 16: return reinterpret_cast<const void*>(&kPathsQ<16, 1, MODE, ADAPT, EXACT, false>);
    case 32: return reinterpret_cast<const void*>(&kPathsQ<32, 1, MODE, ADAPT, EXACT, false>);
    case 64: return reinterpret_cast<const void*>(&kPathsQ<64, 1, MODE, ADAPT, EXACT, false>);
    case 128: return reinterpret_cast<const void*>(&kPathsQ<64, 2, MODE, ADAPT, EXACT, false>);
    case 256: return reinterpret_cast<const void*>(&kPathsQ<64, 4, MODE, ADAPT, EXACT, false>);
    default:
        if constexpr (MODE != kRefine) {
            if (Dp == 512) return reinterpret_cast<const void*>(&kPathsQ<64, 8, MODE, ADAPT, EXACT, false>);
            if (Dp == 1024) return reinterpret_cast<const void*>(&kPathsQ<64, 16, MODE, ADAPT, EXACT, false>);
        }
        return nullptr;
    }
}

} // namespace path
} // namespace pyr
