// internal.h -- kernel argument blocks and launchers shared by the host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pyr {

enum Mode : int { kSingle = 0, kCoarse = 1, kRefine = 2 };

constexpr int kMaxWidth = 65535;  // projection keys pack the column into 16 bits

// Labels per pixel as stored in the volumes: D rounded up to the lane layout
// (16, 32, 64, 128, 256, 512 or 1024).
inline int paddedLabels(int D)
{
    int p = 16;
    while (p < D) p *= 2;
    return p;
}

// Device workspace of one (W, H, D) geometry.  One allocation, carved up.
struct Workspace {
    int W = 0, H = 0, D = 0, Dp = 0;
    void* base = nullptr;
    size_t bytes = 0;
    uint2* desc = nullptr;      // [2][P] census words (x) or packed BT intervals (x: grad, y: intensity)
    int16_t* C = nullptr;       // [P][Dp] matching cost
    int16_t* Alr = nullptr;     // [P][Dp] L->R path costs (also the BT row-sum scratch)
    int16_t* Atb = nullptr;     // [P][Dp] T->B path costs
    int16_t* Arl = nullptr;     // [P][Dp] R->L path costs
    int32_t* labels = nullptr;  // [P] refinement label word: (base & 0xFFFF) | (valid << 16)
    int16_t* disp = nullptr;    // [P] raw disparity*16 after WTA
    float* conf = nullptr;      // [P] raw confidence after WTA
    uint32_t* keys = nullptr;   // [P] right-view projection keys
};

size_t workspaceBytes(int W, int H, int D);
int ensureWorkspace(Workspace* ws, int W, int H, int D);
void releaseWorkspace(Workspace* ws);

// Everything one level needs; filled by the C-ABI entry points.
struct LevelCall {
    const uint8_t* left; const uint8_t* right; int lstep, rstep;
    const int16_t* pdisp; const float* pconf; int pdpitch, pcpitch, PW, PH;  // pitches in elements
    int16_t* out16; float* out32; float* outc; size_t s16, s32, sc;          // pitches in bytes
    int W, H, D, Dp, minD, maxD, block, P1, P2, uniq, disp12, cap;
    int cost;        // PYRSGM_COST_*
    int post;        // PYRSGM_POST_*
    int numPaths;
    int mode;        // Mode
    float minConf;
    int adaptive;
    int32_t lut[256];
};

hipError_t launchLevel(const LevelCall& c, Workspace& ws, hipStream_t stream);

// Stage timing hook (pyrsgm_prof.h); no-ops unless a benchmark enabled it.
void profBegin(int stage, int W, int H, int D, hipStream_t s);
void profEnd(hipStream_t s);
hipError_t launchPyrDown(const uint8_t* src, int sstep, int sw, int sh,
                         uint8_t* dst, int dstep, int dw, int dh, hipStream_t stream);
hipError_t launchPyrDown2(const uint8_t* srcA, int sstepA, const uint8_t* srcB, int sstepB,
                          int sw, int sh, uint8_t* dstA, int dstepA, uint8_t* dstB, int dstepB,
                          int dw, int dh, hipStream_t stream);

// Adaptive-P2 table over |dI| (0..255) after the reference's sanitisation;
// returns whether adaptive P2 stays enabled.
int buildP2Lut(int P2, int adaptive, float alpha, float beta, float lower, float upper,
               int32_t lut[256]);

// --- per-stage launchers (kernels_*.hip) ---------------------------------
struct PathArgs {
    const int16_t* C;
    int16_t* Alr;
    int16_t* Atb;
    int16_t* Arl;
    const int32_t* labels;     // refine only
    const uint8_t* img;        // level left image (adaptive P2)
    int istep;
    int W, H, D, Dp, minD, maxD, P1, P2;
    int startX, endX;          // single-scale L->R column range
    int numPaths;
    int exact;                 // pyramid: Cmax + 2*max(P2') may reach SHRT_MAX
    int nLRBlocks;             // combined launch: blocks [0, nLR) run L->R,
    int nRLBlocks;             // [nLR, nLR+nRL) R->L, the rest T->B
    int16_t* disp;
    float* conf;
    uint32_t* keys;
    int uniq;
    float minConf;
    int32_t lut[256];
};

hipError_t launchPaths(const PathArgs& a, int mode, int adaptive, hipStream_t s);
hipError_t launchPathsSingle(const PathArgs& a, int adaptive, hipStream_t s);
hipError_t launchPathsCoarse(const PathArgs& a, int adaptive, hipStream_t s);
hipError_t launchPathsRefine(const PathArgs& a, int adaptive, hipStream_t s);
hipError_t launchWTA(const PathArgs& a, int mode, hipStream_t s);

hipError_t launchPrep(const LevelCall& c, Workspace& ws, hipStream_t s);
hipError_t launchCost(const LevelCall& c, Workspace& ws, hipStream_t s);
hipError_t launchPost(const LevelCall& c, Workspace& ws, hipStream_t s);

} // namespace pyr
