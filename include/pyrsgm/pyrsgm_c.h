/*
 * pyrsgm_c.h -- C-ABI of the MI355X (gfx950) pyrsgm engine.
 *
 * Drop-in for gegeges/pyrsgm include/pyrsgm/pyrsgm_c.h: same symbol names,
 * argument order, types, flag values and return convention; the stream type
 * is hipStream_t.  Every entry point enqueues device work on `stream` and
 * returns without synchronising the host.  0 = success, -1 = argument or
 * launch failure (launch failures also print the HIP error to stderr);
 * asynchronous device errors surface through hipGetLastError() /
 * hipStreamSynchronize() as usual.
 *
 * Images are device pointers with row pitches in bytes.  The 16-bit
 * disparity is OpenCV StereoSGBM fixed point: value = disparity * 16,
 * invalid = (minDisparity - 1) * 16.
 */
#ifndef PYRSGM_C_H
#define PYRSGM_C_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* costMode -- reference pyrsgm_c.h:25-26 */
#define PYRSGM_COST_BT       0   /* Birchfield-Tomasi on gradient + intensity, SAD window */
#define PYRSGM_COST_CENSUS   1   /* 5x5 Census + Hamming distance                         */

/* postFlags bits -- reference pyrsgm_c.h:29-30 */
#define PYRSGM_POST_LRCHECK  1   /* left-right consistency check (pseudo / strict)        */
#define PYRSGM_POST_MEDIAN   2   /* 3x3 median of the final disparity                     */

/*
 * Device workspace for one (width, height, numDisparities) geometry.
 * Replaces reference pyrsgm_c.h:38-39.  It is (re)allocated lazily when the
 * geometry changes; use one workspace per pyramid level and per stream that
 * runs concurrently.
 */
void* pyrsgm_create_workspace(void);
void  pyrsgm_destroy_workspace(void* workspace);

/*
 * Allocate the workspace for a geometry up front so that later compute calls
 * with the same geometry never allocate.  Replaces reference pyrsgm_c.h:46.
 * numDisparities: multiple of 16 in [16, 1024]; width <= 65535.
 */
int pyrsgm_workspace_reserve(void* workspace, int width, int height, int numDisparities);

/*
 * Full-resolution single-scale SGM.  Replaces reference pyrsgm_c.h:61-83.
 * d_disp, d_disp_float and d_conf_float are optional (NULL skips them); the
 * float outputs are produced only when both float pointers are given.
 * numDisparities: multiple of 16 in [16, 1024].  numPaths: 1..3
 * (3 = left->right, top->bottom, right->left; 2 drops top->bottom; 1 keeps
 * right->left).  adaptiveP2 != 0 enables
 *   P2' = P2 * (lower + (1 + alpha - lower) * exp(-|dI| / beta))
 * clamped to [max(1, P2*lower), P2*upper].
 */
int pyrsgm_compute_single_scale(
    void* workspace,
    const unsigned char* d_left_gray,
    const unsigned char* d_right_gray,
    short* d_disp,
    float* d_disp_float,
    float* d_conf_float,
    size_t left_step_bytes,
    size_t right_step_bytes,
    size_t disp_step_bytes,
    size_t disp_float_step_bytes,
    size_t conf_float_step_bytes,
    int width, int height,
    int numDisparities, int minDisparity,
    int blockSize, int P1, int P2,
    int uniquenessRatio, int disp12MaxDiff,
    int preFilterCap,
    int costMode,
    int postFlags,
    int numPaths,
    int adaptiveP2,
    float p2Alpha, float p2Beta, float p2LowerMult, float p2UpperMult,
    hipStream_t stream);

/*
 * Size of the fixed device scratch array behind
 * pyrsgm_compute_single_scale_noalloc: 192 MiB, held by every device the
 * function runs on until the process ends.  It fits a workspace of
 * W*H*(8*Dp + 30) bytes (Dp = numDisparities rounded up to 16 * 2^k), e.g.
 * 640x480 with 64 disparities.
 */
#define PYRSGM_NOALLOC_SCRATCH_BYTES (192u << 20)

/*
 * pyrsgm_compute_single_scale without a workspace argument and without any
 * device allocation: the same kernels run on a fixed device array of
 * PYRSGM_NOALLOC_SCRATCH_BYTES per device.  No counterpart in the reference;
 * it is for callers that cannot allocate (work recorded under stream
 * capture, serving loops and harnesses that forbid hipMalloc).  Arguments,
 * outputs and error behaviour are those of pyrsgm_compute_single_scale;
 * returns -1 (with a message on stderr) when the workspace for
 * (width, height, numDisparities) exceeds the array.  All calls on one device
 * share the array: calls that overlap in time on one device race, so
 * serialise them (one stream, or events between streams).
 */
int pyrsgm_compute_single_scale_noalloc(
    const unsigned char* d_left_gray,
    const unsigned char* d_right_gray,
    short* d_disp,
    float* d_disp_float,
    float* d_conf_float,
    size_t left_step_bytes,
    size_t right_step_bytes,
    size_t disp_step_bytes,
    size_t disp_float_step_bytes,
    size_t conf_float_step_bytes,
    int width, int height,
    int numDisparities, int minDisparity,
    int blockSize, int P1, int P2,
    int uniquenessRatio, int disp12MaxDiff,
    int preFilterCap,
    int costMode,
    int postFlags,
    int numPaths,
    int adaptiveP2,
    float p2Alpha, float p2Beta, float p2LowerMult, float p2UpperMult,
    hipStream_t stream);

/*
 * One level of the coarse-to-fine pyramid.  Replaces reference
 * pyrsgm_c.h:99-128.
 *   refineLevel == 0: full search over [minDisparity, minDisparity+numDisparities)
 *                     capped by maxDisparity; parent inputs are ignored.
 *   refineLevel != 0: numDisparities (16; at most 256) contiguous local labels
 *                     around the up-scaled parent disparity (parent is the
 *                     half-resolution level); parent pixels that are invalid
 *                     or below minConfidence stay invalid.
 * Intermediate levels should pass postFlags = PYRSGM_POST_LRCHECK.
 */
int pyrsgm_compute_level(
    void* workspace,
    const unsigned char* d_left_gray,
    const unsigned char* d_right_gray,
    const short* d_parent_disp,
    const float* d_parent_conf,
    short* d_disp,
    float* d_disp_float,
    float* d_conf_float,
    size_t left_step_bytes,
    size_t right_step_bytes,
    size_t parent_disp_step_bytes,
    size_t parent_conf_step_bytes,
    size_t disp_step_bytes,
    size_t disp_float_step_bytes,
    size_t conf_float_step_bytes,
    int width, int height,
    int parent_width, int parent_height,
    int numDisparities, int minDisparity, int maxDisparity,
    int blockSize, int P1, int P2,
    int uniquenessRatio, int disp12MaxDiff,
    int preFilterCap,
    int costMode,
    int postFlags,
    int numPaths,
    int refineLevel,
    float minConfidence,
    int adaptiveP2,
    float p2Alpha, float p2Beta, float p2LowerMult, float p2UpperMult,
    hipStream_t stream);

/*
 * 5x5 Gaussian ([1 4 6 4 1]^2 / 256) down-sampling by two, BORDER_REFLECT_101,
 * rounding (sum + 128) >> 8 (cv::pyrDown for 8UC1).  dst must be
 * ((src_width + 1) / 2, (src_height + 1) / 2); src at least 2x2.
 * Replaces reference pyrsgm_c.h:135-140.
 */
int pyrsgm_pyrdown_u8(
    const unsigned char* d_src, size_t src_step_bytes,
    int src_width, int src_height,
    unsigned char* d_dst, size_t dst_step_bytes,
    int dst_width, int dst_height,
    hipStream_t stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PYRSGM_C_H */
