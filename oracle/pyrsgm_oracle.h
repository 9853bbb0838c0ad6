/*
 * pyrsgm_oracle.h -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Single-threaded C restatement of the reference's per-pair SGM path
 * (gegeges/pyrsgm: src/sgm_kernels.cu, src/pyrdown.cu, src/stereo_matcher.cpp).
 * Each function cites the reference lines whose semantics it restates.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline.  The product
 * (pyrsgm_amd/) never links, loads or calls it.
 *
 * All buffers are host memory.  Images are pitched (steps in bytes, as in the
 * reference C-ABI); internal volumes are dense int16 [y][x][D], D innermost.
 */
#ifndef PYRSGM_ORACLE_H
#define PYRSGM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_COST_BT 0
#define ORC_COST_CENSUS 1
#define ORC_POST_LRCHECK 1
#define ORC_POST_MEDIAN 2

/* ---- stage functions (exported for stage-level pins) ------------------ */
int  orc_pyrdown_u8(const uint8_t* src, size_t sstep, int sw, int sh,
                    uint8_t* dst, size_t dstep, int dw, int dh);
void orc_census5x5(const uint8_t* img, size_t step, int w, int h, uint32_t* out);
void orc_bt_grad(const uint8_t* img, size_t step, int w, int h, int cap, uint8_t* out);
void orc_bt_cost(const uint8_t* g1, const uint8_t* g2,
                 const uint8_t* i1, size_t s1, const uint8_t* i2, size_t s2,
                 int w, int h, int D, int minD, int16_t* cost);
void orc_box_h(const int16_t* src, int16_t* dst, int w, int h, int D, int block);
void orc_box_v(const int16_t* src, int16_t* dst, int w, int h, int D, int block);
void orc_census_cost(const uint32_t* cl, const uint32_t* cr, int w, int h,
                     int D, int minD, int16_t* cost);
/* Returns 1 if adaptive P2 stays enabled after sanitisation (lut filled),
 * 0 if it degenerates to the constant P2 (lut filled with P2). */
int  orc_p2_lut(int P2, int adaptive, float alpha, float beta, float lower,
                float upper, int32_t lut[256]);

void orc_path_lr(int16_t* acc, const int16_t* C, const uint8_t* img, size_t istep,
                 int w, int h, int D, int startX, int endX, int P1, int P2,
                 int adaptive, const int32_t* lut);
void orc_path_tb(int16_t* acc, const int16_t* C, const uint8_t* img, size_t istep,
                 int w, int h, int D, int P1, int P2, int adaptive, const int32_t* lut);
void orc_path_rl_wta(int16_t* acc, const int16_t* C, const uint8_t* img, size_t istep,
                     int16_t* disp, int16_t* disp2, int16_t* disp2_cost, float* conf,
                     int w, int h, int D, int minD, int P1, int P2, int uniq,
                     int accumulate, int adaptive, const int32_t* lut);
void orc_pseudo_lr_check(int16_t* disp, const int16_t* disp2, float* conf,
                         int w, int h, int minD, int maxDiff);
void orc_strict_lr_check(int16_t* disp, const int16_t* disp2, float* conf,
                         int w, int h, int minD, int maxDiff);
void orc_median3x3(const int16_t* src, int16_t* dst, int w, int h);

void orc_prepare_refine_labels(const int16_t* pdisp, const float* pconf, int pw, int ph,
                               size_t pdstep, size_t pcstep, int16_t* base, uint8_t* valid,
                               int w, int h, int minD, int maxD, int localD, float minConf);
void orc_census_cost_local(const uint32_t* cl, const uint32_t* cr, const int16_t* base,
                           const uint8_t* valid, int16_t* cost, int w, int h, int D,
                           int minD, int maxD);
void orc_bt_window_cost_local(const uint8_t* g1, const uint8_t* g2,
                              const uint8_t* i1, size_t s1, const uint8_t* i2, size_t s2,
                              const int16_t* base, const uint8_t* valid, int16_t* cost,
                              int w, int h, int D, int minD, int maxD, int block);
void orc_path_lr_pyr(int16_t* acc, const int16_t* C, const int16_t* base, const uint8_t* valid,
                     const uint8_t* img, size_t istep, int w, int h, int D, int minD, int maxD,
                     int P1, int P2, int adaptive, const int32_t* lut);
void orc_path_tb_pyr(int16_t* acc, const int16_t* C, const int16_t* base, const uint8_t* valid,
                     const uint8_t* img, size_t istep, int w, int h, int D, int minD, int maxD,
                     int P1, int P2, int adaptive, const int32_t* lut);
void orc_path_rl_pyr_wta(int16_t* acc, const int16_t* C, const int16_t* base,
                         const uint8_t* valid, const uint8_t* img, size_t istep,
                         int16_t* disp, int16_t* disp2, int16_t* disp2_cost, float* conf,
                         int w, int h, int D, int minD, int maxD, int P1, int P2, int uniq,
                         float minConf, int accumulate, int adaptive, const int32_t* lut);

/* ---- C-ABI restatements (host pointers, same argument meaning) -------- */
int orc_compute_single_scale(
    const uint8_t* left, const uint8_t* right, int16_t* disp, float* dispF, float* confF,
    size_t lstep, size_t rstep, size_t dstep, size_t dfstep, size_t cfstep,
    int w, int h, int D, int minD, int block, int P1, int P2, int uniq, int disp12,
    int cap, int costMode, int postFlags, int numPaths, int adaptive,
    float alpha, float beta, float lower, float upper);

int orc_compute_level(
    const uint8_t* left, const uint8_t* right, const int16_t* pdisp, const float* pconf,
    int16_t* disp, float* dispF, float* confF,
    size_t lstep, size_t rstep, size_t pdstep, size_t pcstep,
    size_t dstep, size_t dfstep, size_t cfstep,
    int w, int h, int pw, int ph, int D, int minD, int maxD, int block, int P1, int P2,
    int uniq, int disp12, int cap, int costMode, int postFlags, int numPaths,
    int refine, float minConf, int adaptive, float alpha, float beta, float lower,
    float upper);

/* ---- StereoMatcher restatement (validated params assumed) ------------- */
typedef struct {
    int num_disparities, min_disparity, block_size, cost_mode, p1, p2;
    int uniqueness_ratio, disp12_max_diff, lr_check, median_filter, num_paths;
    int pre_filter_cap, pyramid_levels, pyramid_coarse_disparities, pyramid_max_disparity;
    float pyramid_min_confidence;
    int adaptive_p2;
    float p2_alpha, p2_beta, p2_lower_mult, p2_upper_mult;
} orc_params;

typedef struct { int level, width, height, num_labels, max_disparity, block_size, refine; } orc_level_info;

/* Fills info[0..levels] (levels = p->pyramid_levels); returns number of levels. */
int orc_level_geometry(int w, int h, const orc_params* p, orc_level_info info[5]);

/* Dense host images; any output may be NULL.  Returns 0 on success. */
int orc_matcher_compute(int w, int h, const orc_params* p,
                        const uint8_t* left, const uint8_t* right,
                        int16_t* disp16, float* disp32, float* conf);

#ifdef __cplusplus
}
#endif
#endif
