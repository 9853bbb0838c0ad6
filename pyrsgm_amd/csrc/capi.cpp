// capi.cpp -- the pyrsgm_c.h entry points: argument checks, workspace
// management, the adaptive-P2 table, and the per-level launch sequence.
//
// Per level, five launches on the caller's stream and no host synchronisation:
//   kPrep      census words / BT intervals of both views, projection keys reset
//   cost       full-range census | BT + window sums | refinement-local costs
//   kPaths     L->R, R->L and T->B path costs, all three in one launch
//   kWTA       path sums + winner-take-all + uniqueness + sub-pixel + projection
//   kPost      LR check + median + confidence clear + output conversions
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <new>

#include "internal.h"
#include "pyrsgm/pyrsgm_c.h"
#include "pyrsgm/pyrsgm_prof.h"

namespace pyr {

namespace {
constexpr size_t kAlign = 256;
size_t alignUp(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Layout {
    size_t desc, C, Alr, Atb, Arl, labels, disp, conf, keys, total;
};

Layout layoutFor(int W, int H, int D)
{
    const size_t P = (size_t)W * H;
    const size_t Dp = (size_t)paddedLabels(D);
    Layout l{};
    size_t o = 0;
    l.desc = o; o += alignUp(P * 2 * sizeof(uint2));
    l.C = o; o += alignUp(P * Dp * 2);
    l.Alr = o; o += alignUp(P * Dp * 2);
    l.Atb = o; o += alignUp(P * Dp * 2);
    l.Arl = o; o += alignUp(P * Dp * 2);
    l.labels = o; o += alignUp(P * 4);
    l.disp = o; o += alignUp(P * 2);
    l.conf = o; o += alignUp(P * 4);
    l.keys = o; o += alignUp(P * 4);
    l.total = o;
    return l;
}
} // namespace

size_t workspaceBytes(int W, int H, int D) { return layoutFor(W, H, D).total; }

void releaseWorkspace(Workspace* ws)
{
    if (ws->base && !ws->borrowed) (void)hipFree(ws->base);
    *ws = Workspace{};
}

namespace {
// Points ws's buffers into n slices of layout l starting at base.
void carve(Workspace* ws, void* base, const Layout& l, int n, int W, int H, int D)
{
    char* b = static_cast<char*>(base);
    ws->base = base;
    ws->bytes = l.total * (size_t)n;
    ws->n = n;
    ws->slice = l.total;
    ws->W = W;
    ws->H = H;
    ws->D = D;
    ws->Dp = paddedLabels(D);
    ws->desc = reinterpret_cast<uint2*>(b + l.desc);
    ws->C = reinterpret_cast<int16_t*>(b + l.C);
    ws->Alr = reinterpret_cast<int16_t*>(b + l.Alr);
    ws->Atb = reinterpret_cast<int16_t*>(b + l.Atb);
    ws->Arl = reinterpret_cast<int16_t*>(b + l.Arl);
    ws->labels = reinterpret_cast<int32_t*>(b + l.labels);
    ws->disp = reinterpret_cast<int16_t*>(b + l.disp);
    ws->conf = reinterpret_cast<float*>(b + l.conf);
    ws->keys = reinterpret_cast<uint32_t*>(b + l.keys);
}
} // namespace

int ensureWorkspace(Workspace* ws, int W, int H, int D, int n)
{
    if (!ws || n < 1) return -1;
    // A workspace holding at least n slices of this geometry serves n pairs.
    if (ws->base && ws->W == W && ws->H == H && ws->D == D && ws->n >= n) return 0;
    if (ws->borrowed) return -1;
    releaseWorkspace(ws);
    const Layout l = layoutFor(W, H, D);
    void* base = nullptr;
    if (hipMalloc(&base, l.total * (size_t)n) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    carve(ws, base, l, n, W, H, D);
    return 0;
}

int borrowWorkspace(Workspace* ws, void* base, size_t bytes, int W, int H, int D)
{
    if (!ws || !base || (reinterpret_cast<uintptr_t>(base) % kAlign) != 0) return -1;
    const Layout l = layoutFor(W, H, D);
    if (l.total > bytes) return -1;
    *ws = Workspace{};
    carve(ws, base, l, 1, W, H, D);
    ws->borrowed = true;
    return 0;
}

// sgm_kernels.cu:63-74 tabulated over |dI|, after the sanitisation of
// sgm_kernels.cu:1671-1675 / :1891-1895.  glibc expf stands in for CUDA __expf;
// this file is compiled with -ffp-contract=off so every float operation rounds
// on its own, in the reference's order.
int buildP2Lut(int P2, int adaptive, float alpha, float beta, float lower, float upper,
               int32_t lut[256])
{
    if (alpha < 0.0f) alpha = 0.0f;
    beta = fmaxf(beta, 0.5f);
    upper = fmaxf(upper, 1.0f);
    lower = fminf(fmaxf(lower, 0.0f), upper);
    if (alpha <= 0.0f && lower >= 1.0f) adaptive = 0;
    for (int i = 0; i < 256; ++i) {
        if (!adaptive) {
            lut[i] = P2;
            continue;
        }
        const float diff = (float)i;
        const float e = expf(-diff / beta);
        const float ratio = lower + (1.0f + alpha - lower) * e;
        float v = (float)P2 * ratio;
        const float lo = fmaxf(1.0f, (float)P2 * lower);
        const float hi = (float)P2 * upper;
        v = fminf(hi, fmaxf(lo, v));
        lut[i] = (int32_t)(v + 0.5f);
    }
    return adaptive;
}

// The pyramid path kernels may drop the reference's SHRT_MAX tests when an
// active line's jump term minPrev + P2' provably stays below SHRT_MAX:
// minPrev <= Cmax + max(P2') (a path cost never exceeds its pixel's cost plus
// the largest penalty), so Cmax + 2 * max(P2') < SHRT_MAX suffices.
int exactPathsNeeded(const LevelCall& c)
{
    if (c.mode == kSingle) return 0;
    long cmax;
    if (c.cost == 1) {
        cmax = 24;
    } else {
        cmax = 255L * c.block * c.block;
        if (cmax > 32766) cmax = 32766;
    }
    long p2max = 0;
    for (int i = 0; i < 256; ++i) p2max = std::max<long>(p2max, c.lut[i]);
    p2max = std::max<long>(p2max, c.P2);
    return cmax + 2 * p2max >= 32767 ? 1 : 0;
}

int narrowVolumes(const LevelCall& c)
{
    if (c.cost != 1 || c.Dp > 64) return 0;
    long p2max = c.P2;
    for (int i = 0; i < 256; ++i) p2max = std::max<long>(p2max, c.lut[i]);
    return 24 + 2 * p2max < 255 ? 1 : 0;
}

PathPart& pathPart(int mode, int adaptive, int form)
{
    // Filled by the kernels_path.hip units' static initialisers (construct on
    // first use, so their order does not matter).
    static PathPart table[3][2][3];
    return table[mode][adaptive ? 1 : 0][form];
}

hipError_t launchPaths(const PathArgs& a, int mode, int adaptive, hipStream_t s)
{
    const PathPart& p = pathPart(mode, adaptive, pathForm(mode, a.exact, a.narrow));
    return p.launch ? p.launch(a, s) : hipErrorInvalidValue;
}

int queuedLinesPerBlock(int Dp, int paired)
{
    if (paired && Dp == 16) return 8;
    return Dp <= 16 ? 4 : (Dp <= 32 ? 2 : 1);
}

int pairedLinesAuto(size_t blocks4) { return blocks4 >= kPairedWaves ? 1 : 0; }

namespace {
const void* queuedKernelOf(int mode, int Dp, int adaptive, int exact, int narrow, int paired)
{
    const PathPart& p = pathPart(mode, adaptive, pathForm(mode, exact, narrow));
    return p.queued ? p.queued(Dp, paired) : nullptr;
}
} // namespace

int queuedBlocksPerCU(int mode, int Dp, int adaptive, int exact, int narrow, int paired)
{
    const void* k = queuedKernelOf(mode, Dp, adaptive, exact, narrow, paired);
    int n = 0;
    if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, 128, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

hipError_t launchPathsQueued(const QueueArgs& q, int mode, int Dp, int adaptive, int exact,
                             int narrow, int paired, int grid, hipStream_t s)
{
    if (q.total <= 0) return hipSuccess;
    const void* k = queuedKernelOf(mode, Dp, adaptive, exact, narrow, paired);
    if (!k || grid < 1) return hipErrorInvalidValue;
    QueueArgs args = q;
    void* params[] = {&args};
    return hipLaunchKernel(k, dim3(grid), dim3(128), params, 0, s);
}

int fusedWTA(const LevelCall& c, int exact)
{
    if (c.sink || !(c.stages & kStagePaths) || !(c.stages & kStageWTA)) return 0;
    if (c.mode == kSingle || !c.narrow || exact || c.Dp != 16) return 0;
    return c.tiles.fused == 2 ? 1 : 0;
}

hipError_t launchLevel(const LevelCall& c, Workspace& ws, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if (c.stages & kStagePrep) {
        profBegin(PYRSGM_STAGE_PREP, c.W, c.H, c.D, s);
        e = launchPrep(c, ws, s);
        profEnd(s);
        if (e != hipSuccess) return e;
    }
    if (c.stages & kStageCost) {
        profBegin(PYRSGM_STAGE_COST, c.W, c.H, c.D, s);
        e = launchCost(c, ws, s);
        profEnd(s);
        if (e != hipSuccess) return e;
    }
    if (!(c.stages & (kStagePaths | kStageFinish))) return hipSuccess;
    if (!(c.stages & (kStagePaths | kStageWTA))) {
        profBegin(PYRSGM_STAGE_POST, c.W, c.H, c.D, s);
        e = launchPost(c, ws, s);
        profEnd(s);
        return e;
    }

    PathArgs a;
    memset(&a, 0, sizeof(a));
    a.C = ws.C;
    a.Alr = ws.Alr;
    a.Atb = ws.Atb;
    a.Arl = ws.Arl;
    a.labels = c.mode == kRefine ? ws.labels : nullptr;
    a.img = c.left;
    a.istep = c.lstep;
    a.W = c.W;
    a.H = c.H;
    a.D = c.D;
    a.Dp = c.Dp;
    a.minD = c.minD;
    a.maxD = c.maxD;
    a.P1 = c.P1;
    a.P2 = c.P2;
    a.startX = std::max(c.minD + c.D, 0);
    a.endX = c.W + std::min(c.minD, 0);
    a.numPaths = c.numPaths;
    a.exact = exactPathsNeeded(c);
    a.narrow = c.narrow;
    static const int remap = getenv("PYRSGM_NO_XCD_REMAP") ? 0 : 1;
    a.xcdRemap = c.tiles.xcdRemap == 0 ? remap : (c.tiles.xcdRemap == 2 ? 1 : 0);
    a.pairedMode = c.tiles.paired;
    // PYRSGM_MODULAR=1 / 2 turns the modular refinement step off / on where the
    // launch shape leaves it to the launch size (A/B timing; same results).
    static const int modularEnv = [] {
        const char* e = getenv("PYRSGM_MODULAR");
        const int v = e ? atoi(e) : 0;
        return (v == 1 || v == 2) ? v : 0;
    }();
    a.modularMode = c.tiles.modular ? c.tiles.modular : modularEnv;
    a.fuseWTA = fusedWTA(c, a.exact);
    a.disp = ws.disp;
    a.conf = ws.conf;
    a.keys = ws.keys;
    a.uniq = c.uniq;
    a.minConf = c.minConf;
    a.invW = std::nextafter(1.0 / c.W, 2.0);
    a.batch = c.batch;
    memcpy(a.lut, c.lut, sizeof(a.lut));

    if (c.stages & kStagePaths) {
        if (c.sink) {
            c.sink->tasks.push_back(PathTask{a, c.mode, c.adaptive});
        } else {
            profBegin(PYRSGM_STAGE_PATHS, c.W, c.H, c.D, s);
            e = launchPaths(a, c.mode, c.adaptive, s);
            profEnd(s);
            if (e != hipSuccess) return e;
        }
    }
    if ((c.stages & kStageWTA) && !a.fuseWTA) {
        profBegin(PYRSGM_STAGE_WTA, c.W, c.H, c.D, s);
        e = launchWTA(a, c.mode, s);
        profEnd(s);
        if (e != hipSuccess) return e;
    }
    if (!(c.stages & kStagePost)) return hipSuccess;
    profBegin(PYRSGM_STAGE_POST, c.W, c.H, c.D, s);
    e = launchPost(c, ws, s);
    profEnd(s);
    return e;
}

} // namespace pyr

using namespace pyr;

namespace {

int fail(hipError_t e, const char* what)
{
    fprintf(stderr, "pyrsgm HIP error in %s: %s\n", what, hipGetErrorString(e));
    return -1;
}

bool geometryOk(int W, int H, int D)
{
    return W > 0 && H > 0 && W <= kMaxWidth && D >= 16 && D <= 1024 && (D % 16) == 0;
}

// The kernels rely on non-negative penalties (the reference C++ API requires
// 0 <= P1 <= P2); window sums are sized for blocks up to 255.
bool paramsOk(int block, int P1, int P2)
{
    return block >= 1 && block <= 255 && P1 >= 0 && P2 >= 0;
}

} // namespace

namespace pyr {

// Reference: sgm_kernels.cu:1864-2076 (b.n pairs).
int computeSingleScale(
    void* workspace, const unsigned char* d_left_gray, const unsigned char* d_right_gray,
    short* d_disp, float* d_disp_float, float* d_conf_float, size_t left_step_bytes,
    size_t right_step_bytes, size_t disp_step_bytes, size_t disp_float_step_bytes,
    size_t conf_float_step_bytes, int width, int height, int numDisparities, int minDisparity,
    int blockSize, int P1, int P2, int uniquenessRatio, int disp12MaxDiff, int preFilterCap,
    int costMode, int postFlags, int numPaths, int adaptiveP2, float p2Alpha, float p2Beta,
    float p2LowerMult, float p2UpperMult, const Batch& b, hipStream_t stream, int stages,
    PathSink* sink, const TileConfig* tiles, StageSink* stageSink)
{
    if (!workspace || !d_left_gray || !d_right_gray || !geometryOk(width, height, numDisparities) ||
        !paramsOk(blockSize, P1, P2))
        return -1;
    Workspace* ws = static_cast<Workspace*>(workspace);
    if (ensureWorkspace(ws, width, height, numDisparities, b.n) != 0) {
        fprintf(stderr, "pyrsgm: workspace allocation failed\n");
        return -1;
    }
    LevelCall c;
    memset(&c, 0, sizeof(c));
    c.left = d_left_gray;
    c.right = d_right_gray;
    c.lstep = (int)left_step_bytes;
    c.rstep = (int)right_step_bytes;
    c.out16 = d_disp;
    c.out32 = d_disp_float;
    c.outc = d_conf_float;
    c.s16 = disp_step_bytes;
    c.s32 = disp_float_step_bytes;
    c.sc = conf_float_step_bytes;
    c.W = width;
    c.H = height;
    c.D = numDisparities;
    c.Dp = ws->Dp;
    c.minD = minDisparity;
    c.maxD = minDisparity + numDisparities - 1;
    c.block = blockSize;
    c.P1 = P1;
    c.P2 = P2;
    c.uniq = uniquenessRatio;
    c.disp12 = disp12MaxDiff;
    c.cap = preFilterCap;
    c.cost = costMode == PYRSGM_COST_CENSUS ? 1 : 0;
    c.post = postFlags;
    c.numPaths = std::min(3, std::max(1, numPaths));
    c.mode = kSingle;
    c.adaptive = buildP2Lut(P2, adaptiveP2, p2Alpha, p2Beta, p2LowerMult, p2UpperMult, c.lut);
    c.narrow = narrowVolumes(c);
    c.stages = stages;
    c.sink = sink;
    c.stageSink = stageSink;
    if (tiles) c.tiles = *tiles;
    c.batch = b;
    c.batch.ws = ws->slice;
    const hipError_t e = launchLevel(c, *ws, stream);
    return e == hipSuccess ? 0 : fail(e, "pyrsgm_compute_single_scale");
}

// Reference: sgm_kernels.cu:1617-1862 (b.n pairs).
int computeLevel(
    void* workspace, const unsigned char* d_left_gray, const unsigned char* d_right_gray,
    const short* d_parent_disp, const float* d_parent_conf, short* d_disp, float* d_disp_float,
    float* d_conf_float, size_t left_step_bytes, size_t right_step_bytes,
    size_t parent_disp_step_bytes, size_t parent_conf_step_bytes, size_t disp_step_bytes,
    size_t disp_float_step_bytes, size_t conf_float_step_bytes, int width, int height,
    int parent_width, int parent_height, int numDisparities, int minDisparity,
    int maxDisparity, int blockSize, int P1, int P2, int uniquenessRatio, int disp12MaxDiff,
    int preFilterCap, int costMode, int postFlags, int numPaths, int refineLevel,
    float minConfidence, int adaptiveP2, float p2Alpha, float p2Beta, float p2LowerMult,
    float p2UpperMult, const Batch& b, hipStream_t stream, int stages, PathSink* sink,
    const TileConfig* tiles, StageSink* stageSink)
{
    if (!workspace || !d_left_gray || !d_right_gray || !geometryOk(width, height, numDisparities) ||
        (refineLevel && numDisparities > 256) || !paramsOk(blockSize, P1, P2))
        return -1;
    if (refineLevel && (!d_parent_disp || !d_parent_conf || parent_width <= 0 || parent_height <= 0))
        return -1;
    maxDisparity = std::min(maxDisparity, width - 1);
    maxDisparity = std::min(maxDisparity, SHRT_MAX / 16);
    if (!refineLevel) maxDisparity = std::min(maxDisparity, minDisparity + numDisparities - 1);
    if (maxDisparity < minDisparity) return -1;
    Workspace* ws = static_cast<Workspace*>(workspace);
    if (ensureWorkspace(ws, width, height, numDisparities, b.n) != 0) {
        fprintf(stderr, "pyrsgm: workspace allocation failed\n");
        return -1;
    }
    LevelCall c;
    memset(&c, 0, sizeof(c));
    c.left = d_left_gray;
    c.right = d_right_gray;
    c.lstep = (int)left_step_bytes;
    c.rstep = (int)right_step_bytes;
    c.pdisp = d_parent_disp;
    c.pconf = d_parent_conf;
    c.pdpitch = (int)(parent_disp_step_bytes / sizeof(short));
    c.pcpitch = (int)(parent_conf_step_bytes / sizeof(float));
    c.PW = parent_width;
    c.PH = parent_height;
    c.out16 = d_disp;
    c.out32 = d_disp_float;
    c.outc = d_conf_float;
    c.s16 = disp_step_bytes;
    c.s32 = disp_float_step_bytes;
    c.sc = conf_float_step_bytes;
    c.W = width;
    c.H = height;
    c.D = numDisparities;
    c.Dp = ws->Dp;
    c.minD = minDisparity;
    c.maxD = maxDisparity;
    c.block = blockSize;
    c.P1 = P1;
    c.P2 = P2;
    c.uniq = uniquenessRatio;
    c.disp12 = disp12MaxDiff;
    c.cap = preFilterCap;
    c.cost = costMode == PYRSGM_COST_CENSUS ? 1 : 0;
    c.post = postFlags;
    c.numPaths = std::min(3, std::max(1, numPaths));
    c.mode = refineLevel ? kRefine : kCoarse;
    c.minConf = fminf(fmaxf(minConfidence, 0.0f), 1.0f);
    c.adaptive = buildP2Lut(P2, adaptiveP2, p2Alpha, p2Beta, p2LowerMult, p2UpperMult, c.lut);
    c.narrow = narrowVolumes(c);
    c.stages = stages;
    c.sink = sink;
    c.stageSink = stageSink;
    if (tiles) c.tiles = *tiles;
    c.batch = b;
    c.batch.ws = ws->slice;
    const hipError_t e = launchLevel(c, *ws, stream);
    return e == hipSuccess ? 0 : fail(e, "pyrsgm_compute_level");
}

} // namespace pyr

extern "C" {

void* pyrsgm_create_workspace(void)
{
    return new (std::nothrow) Workspace();
}

void pyrsgm_destroy_workspace(void* workspace)
{
    Workspace* ws = static_cast<Workspace*>(workspace);
    if (!ws) return;
    releaseWorkspace(ws);
    delete ws;
}

int pyrsgm_workspace_reserve(void* workspace, int width, int height, int numDisparities)
{
    if (!workspace || !geometryOk(width, height, numDisparities)) return -1;
    return ensureWorkspace(static_cast<Workspace*>(workspace), width, height, numDisparities);
}

// Reference: sgm_kernels.cu:1864-2076.
int pyrsgm_compute_single_scale(
    void* workspace, const unsigned char* d_left_gray, const unsigned char* d_right_gray,
    short* d_disp, float* d_disp_float, float* d_conf_float, size_t left_step_bytes,
    size_t right_step_bytes, size_t disp_step_bytes, size_t disp_float_step_bytes,
    size_t conf_float_step_bytes, int width, int height, int numDisparities, int minDisparity,
    int blockSize, int P1, int P2, int uniquenessRatio, int disp12MaxDiff, int preFilterCap,
    int costMode, int postFlags, int numPaths, int adaptiveP2, float p2Alpha, float p2Beta,
    float p2LowerMult, float p2UpperMult, hipStream_t stream)
{
    return computeSingleScale(workspace, d_left_gray, d_right_gray, d_disp, d_disp_float,
                              d_conf_float, left_step_bytes, right_step_bytes, disp_step_bytes,
                              disp_float_step_bytes, conf_float_step_bytes, width, height,
                              numDisparities, minDisparity, blockSize, P1, P2, uniquenessRatio,
                              disp12MaxDiff, preFilterCap, costMode, postFlags, numPaths,
                              adaptiveP2, p2Alpha, p2Beta, p2LowerMult, p2UpperMult, Batch{},
                              stream);
}

// Reference: sgm_kernels.cu:1617-1862.
int pyrsgm_compute_level(
    void* workspace, const unsigned char* d_left_gray, const unsigned char* d_right_gray,
    const short* d_parent_disp, const float* d_parent_conf, short* d_disp, float* d_disp_float,
    float* d_conf_float, size_t left_step_bytes, size_t right_step_bytes,
    size_t parent_disp_step_bytes, size_t parent_conf_step_bytes, size_t disp_step_bytes,
    size_t disp_float_step_bytes, size_t conf_float_step_bytes, int width, int height,
    int parent_width, int parent_height, int numDisparities, int minDisparity,
    int maxDisparity, int blockSize, int P1, int P2, int uniquenessRatio, int disp12MaxDiff,
    int preFilterCap, int costMode, int postFlags, int numPaths, int refineLevel,
    float minConfidence, int adaptiveP2, float p2Alpha, float p2Beta, float p2LowerMult,
    float p2UpperMult, hipStream_t stream)
{
    return computeLevel(workspace, d_left_gray, d_right_gray, d_parent_disp, d_parent_conf,
                        d_disp, d_disp_float, d_conf_float, left_step_bytes, right_step_bytes,
                        parent_disp_step_bytes, parent_conf_step_bytes, disp_step_bytes,
                        disp_float_step_bytes, conf_float_step_bytes, width, height,
                        parent_width, parent_height, numDisparities, minDisparity, maxDisparity,
                        blockSize, P1, P2, uniquenessRatio, disp12MaxDiff, preFilterCap, costMode,
                        postFlags, numPaths, refineLevel, minConfidence, adaptiveP2, p2Alpha,
                        p2Beta, p2LowerMult, p2UpperMult, Batch{}, stream);
}

// Reference: pyrdown.cu:56-74.
int pyrsgm_pyrdown_u8(const unsigned char* d_src, size_t src_step_bytes, int src_width,
                      int src_height, unsigned char* d_dst, size_t dst_step_bytes, int dst_width,
                      int dst_height, hipStream_t stream)
{
    if (!d_src || !d_dst || src_width < 2 || src_height < 2 ||
        dst_width != (src_width + 1) / 2 || dst_height != (src_height + 1) / 2)
        return -1;
    const hipError_t e = launchPyrDown(d_src, (int)src_step_bytes, src_width, src_height, d_dst,
                                       (int)dst_step_bytes, dst_width, dst_height, stream);
    return e == hipSuccess ? 0 : fail(e, "pyrsgm_pyrdown_u8");
}

} // extern "C"
