// internal.h -- kernel argument blocks and launchers shared by the host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>
#include <vector>

namespace pyr {

struct PathSink;

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

// Pairs of one batched launch.  Pair p's buffers are pair 0's offset by p times
// these strides (bytes); a per-pair call has n = 1 and zero strides.  Kernels
// take the pair from a grid dimension.
struct Batch {
    int n = 1;
    size_t ws = 0;                // workspace slice
    size_t left = 0, right = 0;   // input images
    size_t pdisp = 0, pconf = 0;  // parent level's results (refinement)
    size_t out16 = 0, out32 = 0, outc = 0;
};

template <class T>
__host__ __device__ __forceinline__ T* pairAt(T* p, size_t pair, size_t stride)
{
    // Byte arithmetic on the pointer itself (not through an integer), so that
    // the compiler keeps the address space it infers for p.
    using Byte = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return p ? reinterpret_cast<T*>(reinterpret_cast<Byte*>(p) + pair * stride) : p;
}

// Device workspace of one (W, H, D) geometry for n pairs.  One allocation,
// carved up; the pointers address pair 0's slice.
struct Workspace {
    int W = 0, H = 0, D = 0, Dp = 0;
    int n = 0;                    // pair slices
    size_t slice = 0;             // bytes per slice
    void* base = nullptr;
    size_t bytes = 0;
    uint2* desc = nullptr;      // census: uint32 [2][P] words; BT: uint2 [2][P] intervals (x: grad, y: intensity)
    int16_t* C = nullptr;       // [P][Dp] matching cost
    int16_t* Alr = nullptr;     // [P][Dp] L->R path costs (also the BT row-sum scratch)
    int16_t* Atb = nullptr;     // [P][Dp] T->B path costs
    int16_t* Arl = nullptr;     // [P][Dp] R->L path costs
    int32_t* labels = nullptr;  // [P] refinement label word: (base & 0xFFFF) | (valid << 16)
    int16_t* disp = nullptr;    // [P] raw disparity*16 after WTA
    float* conf = nullptr;      // [P] raw confidence after WTA
    uint32_t* keys = nullptr;   // [P] right-view projection keys
    bool borrowed = false;      // a view of caller-owned memory: never freed or reallocated
};

size_t workspaceBytes(int W, int H, int D);
// A one-slice workspace view of `bytes` bytes at `base` (256-byte aligned
// device memory that the caller owns) for (W, H, D); -1 if it does not fit.
int borrowWorkspace(Workspace* ws, void* base, size_t bytes, int W, int H, int D);
// (Re)allocates n slices of one level's buffers unless the geometry matches
// and at least n slices are there (a borrowed view is never reallocated: -1).
int ensureWorkspace(Workspace* ws, int W, int H, int D, int n = 1);
void releaseWorkspace(Workspace* ws);

// Launch shapes a matcher may choose at run time (every choice gives the same
// bits).  0 everywhere is the default shape.
struct TileConfig {
    int localRows = 0;  // rows per block of the refinement census cost: 0 (= 4), 2, 4 or 8
    int paired = 0;     // paired lines for 16-label byte volumes: 0 by launch size, 1 off, 2 on
    int xcdRemap = 0;   // XCD-contiguous T->B column groups: 0 default, 1 off, 2 on
    int modular = 0;    // modular refinement step: 0 by launch size, 1 off, 2 on
    int fused = 0;      // R->L lines finish the winner-take-all: 0 default (off), 1 off, 2 on
};

// The parent level's results a refinement cost kernel reads (pitches in elements).
struct RefineIn {
    const int16_t* pdisp;
    const float* pconf;
    int pdpitch, pcpitch, PW, PH;
    float minConf;
};

struct StageSink;

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
    int narrow;      // byte volumes (narrowVolumes)
    int stages;      // kStage* bits
    PathSink* sink;  // non-null: path launches are recorded here instead of issued
    StageSink* stageSink;  // non-null: pyrDown / prep / cost / post launches are recorded here
    TileConfig tiles;
    Batch batch;
    int32_t lut[256];
};

// Byte volumes: census levels of at most 64 labels whose every cost and path
// cost fits in a byte.  A path cost is at most its pixel's cost plus the
// largest penalty (Cmax + max(P2') with Cmax = 24), and the fast path step is
// exact below a ceiling M when Cmax + 2 * max(P2') < M; with M = 255 the
// volumes store 255 - C and 255 - L (0: SHRT_MAX) and every value is unchanged.
int narrowVolumes(const LevelCall& c);

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
    int narrow;                // byte volumes (see narrowVolumes)
    int xcdRemap;              // XCD-contiguous line groups (on unless PYRSGM_NO_XCD_REMAP is set)
    int pairedMode;            // TileConfig::paired
    int modularMode;           // TileConfig::modular
    int fuseWTA;               // the R->L lines finish the winner-take-all (see fusedWTA)
    int nLRBlocks;             // combined launch: blocks [0, nLR) run L->R,
    int nRLBlocks;             // [nLR, nLR+nRL) R->L, the rest T->B
    int16_t* disp;
    float* conf;
    uint32_t* keys;
    int uniq;
    float minConf;
    double invW;               // 1/W rounded up by one ulp (pixel rows in the winner-take-all)
    Batch batch;
    int32_t lut[256];
};

hipError_t launchPaths(const PathArgs& a, int mode, int adaptive, hipStream_t s);

// Path kernels of one (mode, adaptive P2, volume form) specialisation; each is
// compiled in a unit of its own (kernels_path.hip) and registers itself here
// when the library loads.  Form: exact pyramid step (path costs may reach
// SHRT_MAX), byte volumes, or int16 volumes.
constexpr int kFormExact = 0, kFormNarrow = 1, kFormWide = 2;
struct PathPart {
    hipError_t (*launch)(const PathArgs& a, hipStream_t s) = nullptr;  // per-launch kPaths
    const void* (*queued)(int Dp, int paired) = nullptr;               // kPathsQ kernel
};
PathPart& pathPart(int mode, int adaptive, int form);
inline int pathForm(int mode, int exact, int narrow)
{
    return (mode != kSingle && exact) ? kFormExact : (narrow ? kFormNarrow : kFormWide);
}
hipError_t launchWTA(const PathArgs& a, int mode, hipStream_t s);

// Parts of a level a call enqueues: the descriptor pass (census words / BT
// intervals and the key reset, which read only the level's images), the cost
// volume, the path recurrences, and the finish (winner-take-all and
// post-processing).
constexpr int kStagePrep = 1, kStageCost = 2, kStagePaths = 4, kStageWTA = 8, kStagePost = 16;
constexpr int kStageFinish = kStageWTA | kStagePost;
constexpr int kStageRest = kStageCost | kStagePaths | kStageFinish;
constexpr int kStagesAll = kStagePrep | kStageRest;

// A path launch recorded instead of issued (LevelCall::sink): its arguments
// for one pair (pair strides unused) and its kernel specialisation.
struct PathTask {
    PathArgs a;
    int mode, adaptive;
};
struct PathSink {
    std::vector<PathTask> tasks;
};

// ---- queued path launches (many pairs' levels of one specialisation) -----
// Work item = (task, block) where a task's blocks are numbered as in its own
// launch (L->R rows, R->L rows, T->B column groups); items are numbered task by
// task, `start[t]` being task t's first.  Workgroups take items from
// counter[0] with an atomic add until it passes the total; the last workgroup
// to finish puts both counters back to zero, so consecutive launches on one
// stream reuse them (launches that share counters must not overlap).
struct QueueArgs {
    const PathArgs* tasks;  // device table [nTasks] (nLRBlocks / nRLBlocks set per task)
    const int* start;       // device [nTasks + 1]
    uint32_t* counter;      // device [2], zero before the launch
    int nTasks, total;
};
// Lines per block of a queued launch of labels Dp (paired: 16-label byte
// volumes with two lines per lane).
int queuedLinesPerBlock(int Dp, int paired);
// Whether a launch of `blocks4` four-line blocks of one 16-label byte-volume
// pyramid specialisation pairs lines (the per-launch rule of the path kernels).
int pairedLinesAuto(size_t blocks4);
// Path launches of at least this many 4-line waves use paired lines (byte
// volumes, 16 labels, pyramid levels): four per SIMD.
constexpr size_t kPairedWaves = 4096;
// Fused winner-take-all (TileConfig::fused = 2, 16-label byte-volume pyramid
// levels in the fast form): the L->R and T->B lines run in one launch, then
// the R->L lines finish each pixel from its three path-cost vectors, so the
// R->L volume is neither written nor read again.  It is a tuning choice, off
// unless asked for: in launches large enough to be bound by VALU issue the
// winner-take-all's instructions move into the R->L launch unchanged, and
// smaller launches are bound by the recurrence's latency, which running R->L
// after the other two directions lengthens.
// Whether a level call runs the fused winner-take-all (TileConfig::fused,
// volume form and stages).
int fusedWTA(const LevelCall& c, int exact);
// 16-label byte-volume refinement path launches of at most this many line
// groups use the modular step (two groups per workgroup; measured faster at
// 1020 groups: 1920x1080 124 against 155 us, and slower at 2040: two 1920x1080
// pairs 231 against 193 us).
constexpr size_t kModularMaxBlocks = 1536;
// Workgroups of one queued launch that are resident at once per CU (occupancy).
int queuedBlocksPerCU(int mode, int Dp, int adaptive, int exact, int narrow, int paired);
hipError_t launchPathsQueued(const QueueArgs& q, int mode, int Dp, int adaptive, int exact,
                             int narrow, int paired, int grid, hipStream_t s);
// Winner-take-all of many pairs' levels of one specialisation (mode, Dp,
// byte volumes) in one launch: block b belongs to task t with
// start[t] <= b < start[t + 1], task t having wtaBlocks(tasks[t]) blocks.
struct WtaQueueArgs {
    const PathArgs* tasks;
    const int* start;  // device [nTasks + 1]
    int nTasks, total;
};
int wtaBlocks(const PathArgs& a);
hipError_t launchWTAQueued(const WtaQueueArgs& q, int mode, int Dp, int narrow, hipStream_t s);
// Sets tasks[idx[i]].img / .istep (the per-call left image of single-scale
// tasks with adaptive P2) on the device.
constexpr int kMaxImgPatches = 64;
struct ImgPatch {
    PathArgs* tasks;
    int n;
    int idx[kMaxImgPatches];
    int istep[kMaxImgPatches];
    const uint8_t* img[kMaxImgPatches];
};
hipError_t launchImgPatch(const ImgPatch& p, hipStream_t s);

// The C-ABI level / single-scale computations for b.n pairs at once (the
// extern "C" entry points call these with one pair).
int computeSingleScale(void* workspace, const unsigned char* d_left_gray,
                       const unsigned char* d_right_gray, short* d_disp, float* d_disp_float,
                       float* d_conf_float, size_t left_step_bytes, size_t right_step_bytes,
                       size_t disp_step_bytes, size_t disp_float_step_bytes,
                       size_t conf_float_step_bytes, int width, int height, int numDisparities,
                       int minDisparity, int blockSize, int P1, int P2, int uniquenessRatio,
                       int disp12MaxDiff, int preFilterCap, int costMode, int postFlags,
                       int numPaths, int adaptiveP2, float p2Alpha, float p2Beta,
                       float p2LowerMult, float p2UpperMult, const Batch& b, hipStream_t stream,
                       int stages = kStagesAll, PathSink* sink = nullptr,
                       const TileConfig* tiles = nullptr, StageSink* stageSink = nullptr);
int computeLevel(void* workspace, const unsigned char* d_left_gray,
                 const unsigned char* d_right_gray, const short* d_parent_disp,
                 const float* d_parent_conf, short* d_disp, float* d_disp_float,
                 float* d_conf_float, size_t left_step_bytes, size_t right_step_bytes,
                 size_t parent_disp_step_bytes, size_t parent_conf_step_bytes,
                 size_t disp_step_bytes, size_t disp_float_step_bytes,
                 size_t conf_float_step_bytes, int width, int height, int parent_width,
                 int parent_height, int numDisparities, int minDisparity, int maxDisparity,
                 int blockSize, int P1, int P2, int uniquenessRatio, int disp12MaxDiff,
                 int preFilterCap, int costMode, int postFlags, int numPaths, int refineLevel,
                 float minConfidence, int adaptiveP2, float p2Alpha, float p2Beta,
                 float p2LowerMult, float p2UpperMult, const Batch& b, hipStream_t stream,
                 int stages = kStagesAll, PathSink* sink = nullptr,
                 const TileConfig* tiles = nullptr, StageSink* stageSink = nullptr);
// pyrDown of n pairs of (one or two) images; src / dst pair strides in bytes.
// (sink: the launch is recorded instead of issued.)
hipError_t launchPyrDownBatch(const uint8_t* srcA, int sstepA, const uint8_t* srcB, int sstepB,
                              size_t sStride, int sw, int sh, uint8_t* dstA, int dstepA,
                              uint8_t* dstB, int dstepB, size_t dStride, int dw, int dh, int n,
                              hipStream_t stream, StageSink* sink = nullptr);

// One 2-D copy of a batched copy launch (pitches and row length in bytes).
struct CopyDesc {
    const uint8_t* src;
    uint8_t* dst;
    size_t spitch, dpitch;
    int rowBytes, rows;
};
// `count` copies read from the device array `descs`, in one launch.
hipError_t launchCopy2D(const CopyDesc* descs, int count, int maxRowBytes, int maxRows,
                        hipStream_t stream);

hipError_t launchPrep(const LevelCall& c, Workspace& ws, hipStream_t s);
hipError_t launchCost(const LevelCall& c, Workspace& ws, hipStream_t s);
hipError_t launchPost(const LevelCall& c, Workspace& ws, hipStream_t s);

// ---- stage launches (pyrDown, descriptor pass, cost volume, post) --------
// Every stage kernel runs either as a launch of its own or as one task of a
// queued launch that runs all tasks of one kernel specialisation (the
// pairs of a mixed batch at one pyramid step): block b of a queued launch
// belongs to the task whose block range holds it and runs tile
// (x, y, z) = (local % gx, local / gx % gy, local / (gx gy)) of that task.
enum StageKind : int {
    kStPrepCensus = 0, kStPrepBT, kStCensusCost, kStCensusLocal, kStBTRow, kStBoxV, kStBTLocal,
    kStBTLocalTile, kStPyrDown, kStPost, kStKinds
};
// Arguments of every stage kernel (each reads the fields it uses).
struct StageArgs {
    const uint8_t* left;  // level images (pyrDown: sources); refinement costs read left only
    const uint8_t* right;
    int lstep, rstep;
    uint8_t* dstA;  // pyrDown outputs
    uint8_t* dstB;
    int dstepA, dstepB, sw, sh, nv;
    size_t sStride, dStride;  // pyrDown pair strides
    uint2* desc;
    uint32_t* keys;
    int32_t* labels;
    int16_t* C;
    int16_t* Hs;  // BT row sums
    RefineIn r;
    int W, H, D, Dp, minD, maxD, half, cap, geom;
    const int16_t* disp;  // post
    const float* conf;
    int maxDiff, lrcheck, median, strict;
    int16_t* out16;
    float* out32;
    float* outc;
    size_t s16, s32, sc;
    Batch batch;
};
// Per-call fields a queued task takes from its pair's I/O entry: the input
// images (left, right and their pitches) and the outputs of level 0.
constexpr int kIoIn = 1, kIoOut = 2;
struct StageLaunch {
    int kind, variant;  // kernel and template variant
    dim3 grid;          // per-launch grid (one task: grid.z counts this pair's slices)
    size_t lds;         // dynamic LDS bytes
    StageArgs a;
    int io;             // kIoIn / kIoOut fields taken from the caller's I/O when queued
};
// Stage launches recorded instead of issued (LevelCall::stageSink).
struct StageSink {
    std::vector<StageLaunch> launches;
};
hipError_t launchStage(const StageLaunch& l, hipStream_t s);
// The caller's buffers of one pair, as the queued tasks read them; outc is
// already routed (float disparities without a confidence output write the
// confidence to the pair's scratch).
struct StageIo {
    const uint8_t* left;
    const uint8_t* right;
    int16_t* out16;
    float* out32;
    float* outc;
    size_t lstep, rstep, s16, s32, sc;
};
struct StageTask {
    StageArgs a;
    int gx, gy;  // tile grid of the task (x, y; z is the rest)
    int io, pair;
};
struct StageQueue {
    const StageTask* tasks;  // device [nTasks]
    const int* start;        // device [nTasks + 1] first block of each task
    const StageIo* io;       // device [pairs]
    int nTasks;
};
hipError_t launchStageQueued(int kind, int variant, const StageQueue& q, int blocks, size_t lds,
                             hipStream_t s);
// Writes n entries of the device I/O table (table[idx[i]] = io[i]) in one launch.
constexpr int kMaxIoPatch = 32;
struct IoPatch {
    StageIo* table;
    int n;
    int idx[kMaxIoPatch];
    StageIo io[kMaxIoPatch];
};
hipError_t launchIoPatch(const IoPatch& p, hipStream_t s);

} // namespace pyr
