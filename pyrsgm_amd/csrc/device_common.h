// device_common.h -- gfx950 building blocks shared by the pyrsgm kernels.
//
// Lane layout used by the aggregation kernels: a "line" (an image row for the
// horizontal paths, a column for the vertical path) is owned by a group of G
// lanes of one wave64 (G = 16, 32 or 64); each lane holds K consecutive labels
// of that line's current pixel (label d = g*K + k).  64/G lines share a wave.
// Cross-label operations (the per-pixel minimum and the d-1 / d+1 neighbours)
// are DPP / permlane register moves inside the wave: no LDS and no barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pyr {

constexpr int kMax = 32767;      // SHRT_MAX: saturation value and "no path" sentinel

// GFX9 DPP control words.
enum : int {
    kDppQuadXor1 = 0xB1,       // quad_perm [1,0,3,2]
    kDppQuadXor2 = 0x4E,       // quad_perm [2,3,0,1]
    kDppRowShl1 = 0x101,       // lane i <- lane i+1 inside its 16-lane row
    kDppRowShl2 = 0x102,       // lane i <- lane i+2 inside its 16-lane row
    kDppRowShr1 = 0x111,       // lane i <- lane i-1 inside its 16-lane row
    kDppRowShr2 = 0x112,       // lane i <- lane i-2 inside its 16-lane row
    kDppWaveShl1 = 0x130,      // lane i <- lane i+1 across the wave
    kDppWaveShr1 = 0x138,      // lane i <- lane i-1 across the wave
    kDppRowMirror = 0x140,     // lane i <- lane 15-i inside its row
    kDppRowRor1 = 0x121,       // lane i <- lane (i-1) mod 16 inside its row
    kDppRowRor15 = 0x12F,      // lane i <- lane (i+1) mod 16 inside its row
    kDppRowHalfMirror = 0x141, // lane i <- lane 7-i inside its half row
    kDppRowBcast15 = 0x142,    // rows 1-3 <- lane 15 of the row before (with a row mask)
    kDppRowBcast31 = 0x143,    // rows 2-3 <- lane 31 (with a row mask)
};

// DPP move; lanes whose source is outside the row/wave keep `old`.
template <int CTRL>
__device__ __forceinline__ int dpp(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, CTRL, 0xF, 0xF, false);
}

// The same move kept out of the compiler's DPP combiner.  On gfx950 a
// reversed-operand VOP2 instruction with DPP (v_subrev_u32_dpp, v_lshlrev_*_dpp,
// ...) applies the lane swizzle to its second source, while the combiner, which
// folds a move whose out-of-row lanes read 0 into its user, assumes the first;
// x - dpp(y) could become such an instruction.  The build rejects any object
// that contains one.
template <int CTRL>
__device__ __forceinline__ int dppOpaque(int old, int src)
{
    int v = dpp<CTRL>(old, src);
    asm volatile("" : "+v"(v));
    return v;
}

// Workgroups are dispatched to the 8 XCDs round-robin in linear order (x
// fastest, then y, then z), and each XCD has its own L2.  xcdRemap(b, n) is a
// bijection of [0, n) under which the blocks of one XCD (equal b % 8) get a
// contiguous range of indices.
__device__ __forceinline__ int xcdRemap(int b, int n)
{
    const int q = n >> 3, r = n & 7, x = b & 7, i = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// The tile (x, y, z) of this workgroup when tiles are handed to the XCDs in
// contiguous raster ranges: neighbouring tiles, which read parts of the same
// 128-byte lines and share halo rows, then share one L2 instead of each XCD
// fetching them.
struct Tile3 {
    int x, y, z;
};
__device__ __forceinline__ Tile3 xcdTile()
{
    const int nx = (int)gridDim.x, ny = (int)gridDim.y;
    const int n = nx * ny * (int)gridDim.z;
    const int b = (int)blockIdx.x + nx * ((int)blockIdx.y + ny * (int)blockIdx.z);
    const int t = xcdRemap(b, n);
    const int txy = t % (nx * ny);
    return Tile3{txy % nx, txy / nx, t / (nx * ny)};
}

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
__device__ __forceinline__ int sat16(int v) { return imin(32767, imax(-32768, v)); }

// Value of the next-lower lane of the group; the group's first lane gets `fill`.
template <int G>
__device__ __forceinline__ int from_lower(int v, int fill, int g)
{
    if constexpr (G == 16) {
        (void)g;
        return dpp<kDppRowShr1>(fill, v);
    } else {
        const int t = dpp<kDppWaveShr1>(fill, v);
        return g == 0 ? fill : t;
    }
}

// Value of the next-higher lane of the group; the group's last lane gets `fill`.
template <int G>
__device__ __forceinline__ int from_upper(int v, int fill, int g)
{
    if constexpr (G == 16) {
        (void)g;
        return dpp<kDppRowShl1>(fill, v);
    } else {
        const int t = dpp<kDppWaveShl1>(fill, v);
        return g == G - 1 ? fill : t;
    }
}

// Signed minimum over the G lanes of each group, returned to every lane of it.
// Row step: two quad xors and two mirrors; G>=32 adds a permlane16 swap (rows
// 0<->1, 2<->3), G==64 a permlane32 swap (halves).
template <int G>
__device__ __forceinline__ int group_min(int v)
{
    v = imin(v, dpp<kDppQuadXor1>(v, v));
    v = imin(v, dpp<kDppQuadXor2>(v, v));
    v = imin(v, dpp<kDppRowHalfMirror>(v, v));
    v = imin(v, dpp<kDppRowMirror>(v, v));
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

// Unsigned variant (packed (cost<<16)|label keys).
template <int G>
__device__ __forceinline__ uint32_t group_umin(uint32_t v)
{
    v = umin(v, (uint32_t)dpp<kDppQuadXor1>((int)v, (int)v));
    v = umin(v, (uint32_t)dpp<kDppQuadXor2>((int)v, (int)v));
    v = umin(v, (uint32_t)dpp<kDppRowHalfMirror>((int)v, (int)v));
    v = umin(v, (uint32_t)dpp<kDppRowMirror>((int)v, (int)v));
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

// Minimum over Q consecutive lanes (Q in {1,2,4}), returned to all Q lanes.
template <int Q>
__device__ __forceinline__ int quad_min(int v)
{
    if constexpr (Q >= 2) v = imin(v, dpp<kDppQuadXor1>(v, v));
    if constexpr (Q >= 4) v = imin(v, dpp<kDppQuadXor2>(v, v));
    return v;
}
template <int Q>
__device__ __forceinline__ uint32_t quad_umin(uint32_t v)
{
    if constexpr (Q >= 2) v = umin(v, (uint32_t)dpp<kDppQuadXor1>((int)v, (int)v));
    if constexpr (Q >= 4) v = umin(v, (uint32_t)dpp<kDppQuadXor2>((int)v, (int)v));
    return v;
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

// ---- K consecutive int16 per lane, kept packed two per dword ------------
template <int K>
struct Packed {
    static constexpr int kWords = (K + 1) / 2;
    uint32_t w[kWords];
};

template <int K>
__device__ __forceinline__ void load_packed(Packed<K>& out, const int16_t* p)
{
    if constexpr (K == 1) {
        out.w[0] = *reinterpret_cast<const uint16_t*>(p);
    } else if constexpr (K == 2) {
        out.w[0] = *reinterpret_cast<const uint32_t*>(p);
    } else if constexpr (K == 4) {
        const uint2 v = *reinterpret_cast<const uint2*>(p);
        out.w[0] = v.x; out.w[1] = v.y;
    } else {
        static_assert(K % 8 == 0, "K must be 1, 2, 4 or a multiple of 8");
#pragma unroll
        for (int i = 0; i < K / 8; ++i) {
            const uint4 v = reinterpret_cast<const uint4*>(p)[i];
            out.w[4 * i + 0] = v.x; out.w[4 * i + 1] = v.y;
            out.w[4 * i + 2] = v.z; out.w[4 * i + 3] = v.w;
        }
    }
}

template <int K>
__device__ __forceinline__ int unpack(const Packed<K>& v, int k)
{
    const uint32_t w = v.w[k >> 1];
    return (k & 1) ? ((int)w >> 16) : (int)(int16_t)(w & 0xFFFFu);
}

template <int K>
__device__ __forceinline__ void store_i16(int16_t* p, const int (&v)[K])
{
    if constexpr (K == 1) {
        *p = (int16_t)v[0];
    } else if constexpr (K == 2) {
        *reinterpret_cast<uint32_t*>(p) = (uint32_t)(uint16_t)v[0] | ((uint32_t)v[1] << 16);
    } else if constexpr (K == 4) {
        uint2 o;
        o.x = (uint32_t)(uint16_t)v[0] | ((uint32_t)v[1] << 16);
        o.y = (uint32_t)(uint16_t)v[2] | ((uint32_t)v[3] << 16);
        *reinterpret_cast<uint2*>(p) = o;
    } else {
#pragma unroll
        for (int i = 0; i < K / 8; ++i) {
            uint4 o;
            o.x = (uint32_t)(uint16_t)v[8 * i + 0] | ((uint32_t)v[8 * i + 1] << 16);
            o.y = (uint32_t)(uint16_t)v[8 * i + 2] | ((uint32_t)v[8 * i + 3] << 16);
            o.z = (uint32_t)(uint16_t)v[8 * i + 4] | ((uint32_t)v[8 * i + 5] << 16);
            o.w = (uint32_t)(uint16_t)v[8 * i + 6] | ((uint32_t)v[8 * i + 7] << 16);
            reinterpret_cast<uint4*>(p)[i] = o;
        }
    }
}

// C truncating division n / m for m > 0 and |n| < 2^22: a reciprocal estimate
// (relative error < 2^-22, so the truncated quotient is off by at most one)
// followed by one exact remainder correction.
__device__ __forceinline__ int div_trunc(int n, int m)
{
    const float qf = (float)n * __builtin_amdgcn_rcpf((float)m);
    int q = (int)qf;
    const int r = n - q * m;
    if (n >= 0) {
        if (r < 0) q -= 1;
        else if (r >= m) q += 1;
    } else {
        if (r > 0) q += 1;
        else if (r <= -m) q -= 1;
    }
    return q;
}

} // namespace pyr
