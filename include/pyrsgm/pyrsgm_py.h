/*
 * pyrsgm_py.h -- plain-C handles on pyrsgm::StereoMatcher for FFI callers
 * (the Python package binds these with ctypes).
 *
 * These have no counterpart in gegeges/pyrsgm, which exposes StereoMatcher
 * only as a C++ class; they exist so that a C-only FFI can drive the C++
 * matcher.  Errors are reported as -1 / NULL with the exception text copied
 * into `err` (truncated to errlen).
 */
#ifndef PYRSGM_PY_H
#define PYRSGM_PY_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field-for-field copy of pyrsgm::Params (bools as int, cost_mode as int). */
typedef struct {
    int num_disparities, min_disparity, block_size, cost_mode, p1, p2;
    int uniqueness_ratio, disp12_max_diff, lr_check, median_filter, num_paths;
    int pre_filter_cap, pyramid_levels, pyramid_coarse_disparities, pyramid_max_disparity;
    float pyramid_min_confidence;
    int adaptive_p2;
    float p2_alpha, p2_beta, p2_lower_mult, p2_upper_mult;
} pyrsgm_params_c;

/* pyrsgm::validateParams; returns 0 and fills *out, or -1 (std::invalid_argument). */
int pyrsgm_validate_params(const pyrsgm_params_c* in, pyrsgm_params_c* out, char* err, int errlen);

void* pyrsgm_matcher_create(int width, int height, const pyrsgm_params_c* params,
                            char* err, int errlen);
void pyrsgm_matcher_destroy(void* matcher);
int pyrsgm_matcher_compute(void* matcher,
                           const uint8_t* d_left, size_t left_pitch,
                           const uint8_t* d_right, size_t right_pitch,
                           int16_t* d_disp16, size_t disp16_pitch,
                           float* d_disp32, size_t disp32_pitch,
                           float* d_conf, size_t conf_pitch,
                           hipStream_t stream, char* err, int errlen);
/* out[7] = {level, width, height, num_labels, max_disparity, block_size, refine} */
int pyrsgm_matcher_level_info(void* matcher, int level, int out[7]);
int pyrsgm_matcher_num_levels(void* matcher);
int pyrsgm_matcher_params(void* matcher, pyrsgm_params_c* out);
size_t pyrsgm_matcher_device_bytes(void* matcher);
const char* pyrsgm_version(void);

#ifdef __cplusplus
}
#endif
#endif
