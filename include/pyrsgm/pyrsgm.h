/**
 * @file pyrsgm.h
 * @brief C++ API of the MI355X (gfx950) coarse-to-fine SGM engine.
 *
 * Drop-in for gegeges/pyrsgm include/pyrsgm/pyrsgm.h (reference lines 27-146):
 * same namespace, types, field names, defaults and member functions; the
 * stream type is hipStream_t.
 *
 * @code
 *   pyrsgm::Params p;
 *   p.num_disparities = 256;
 *   p.pyramid_levels  = 2;          // full search at 1/4, refine x2 with 16 labels
 *   p.cost_mode = pyrsgm::CostMode::Census;
 *   pyrsgm::StereoMatcher sgm(width, height, p);
 *   sgm.compute(d_left, left_pitch, d_right, right_pitch,
 *               d_disp16, disp_pitch, nullptr, 0, nullptr, 0, stream);
 * @endcode
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>

#include <hip/hip_runtime_api.h>

namespace pyrsgm {

enum class CostMode {
    BirchfieldTomasi = 0,  ///< BT sampling-insensitive cost + gradient, SAD block window
    Census = 1             ///< 5x5 Census transform + Hamming distance
};

struct Params {
    // ---- matching (reference pyrsgm.h:33-58) ------------------------------
    int num_disparities = 128;    ///< full-resolution range; aligned up to 16, <= 1024
    int min_disparity = 0;        ///< pyramid mode requires 0
    int block_size = 5;           ///< BT SAD window (odd, >= 3); levels >= 2 use 3
    CostMode cost_mode = CostMode::BirchfieldTomasi;
    int p1 = 300;                 ///< small-jump penalty
    int p2 = 2400;                ///< large-jump penalty
    int uniqueness_ratio = 10;    ///< percent, 0 disables
    int disp12_max_diff = 1;      ///< LR tolerance in pixels (pyramid requires >= 0)
    bool lr_check = true;
    bool median_filter = true;    ///< 3x3 median on the final level
    int num_paths = 3;            ///< 3 = L->R, T->B, R->L; 2 drops T->B; 1 keeps R->L
    int pre_filter_cap = 63;      ///< BT gradient clipping

    // ---- coarse-to-fine pyramid (reference pyrsgm.h:61-72) ----------------
    int pyramid_levels = 0;                 ///< 0 = single scale; 1..4 extra levels
    int pyramid_coarse_disparities = 128;   ///< cap on coarsest-level labels
    int pyramid_max_disparity = 0;          ///< <= 0 selects num_disparities - 1
    float pyramid_min_confidence = 0.05f;   ///< parents below this are not propagated

    // ---- adaptive P2 (reference pyrsgm.h:75-81) ---------------------------
    bool adaptive_p2 = false;
    float p2_alpha = 3.041f;
    float p2_beta = 16.137f;
    float p2_lower_mult = 0.601f;
    float p2_upper_mult = 1.77f;
};

/// Resolves derived values and throws std::invalid_argument for unsupported
/// combinations (reference stereo_matcher.cpp:100-148).
Params validateParams(const Params& params);

/// Fixed-point invalid marker of the 16-bit disparity (reference pyrsgm.h:89-92).
inline int16_t invalidDisparity16(const Params& params)
{
    return static_cast<int16_t>((params.min_disparity - 1) * 16);
}

/// Geometry of one pyramid level (reference pyrsgm.h:95-103).
struct LevelInfo {
    int level = 0;
    int width = 0;
    int height = 0;
    int num_labels = 0;      ///< labels evaluated per pixel at this level
    int max_disparity = 0;   ///< real disparity cap at this level's scale
    int block_size = 0;
    bool refine = false;
};

/**
 * Matcher for one image size (reference pyrsgm.h:116-144).  All device
 * memory is allocated in the constructor; compute() allocates nothing and
 * only enqueues work on `stream`.
 *
 * Inputs: 8-bit rectified grayscale device images.  Outputs (nullptr skips):
 *   disp16: int16 disparity*16, invalid = invalidDisparity16();
 *   disp32: float pixels, invalid = 0;   conf: float in [0,1], 0 if invalid.
 */
class StereoMatcher {
public:
    StereoMatcher(int width, int height, const Params& params);
    ~StereoMatcher();
    StereoMatcher(StereoMatcher&&) noexcept;
    StereoMatcher& operator=(StereoMatcher&&) noexcept;
    StereoMatcher(const StereoMatcher&) = delete;
    StereoMatcher& operator=(const StereoMatcher&) = delete;

    /// Enqueue one frame.  Throws std::runtime_error on launch failure.
    void compute(const uint8_t* d_left, size_t left_pitch,
                 const uint8_t* d_right, size_t right_pitch,
                 int16_t* d_disp16, size_t disp16_pitch,
                 float* d_disp32, size_t disp32_pitch,
                 float* d_conf, size_t conf_pitch,
                 hipStream_t stream = nullptr);

    int width() const;
    int height() const;
    const Params& params() const;
    int numLevels() const;               ///< pyramid_levels + 1
    LevelInfo levelInfo(int level) const;
    size_t deviceMemoryBytes() const;    ///< device memory owned by the matcher

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

const char* version();

} // namespace pyrsgm
