/*
 * pyrsgm_batch.h -- batched form: many stereo pairs of one size per launch.
 *
 * No counterpart in gegeges/pyrsgm (whose only multi-pair form is several
 * StereoMatchers on several streams).  A BatchMatcher holds N pairs' worth of
 * device state for one image size and one Params, and runs each stage of the
 * pyramid (pyrDown, costs, paths, winner-take-all, post-processing) as ONE
 * launch over all N pairs; the pair index is a grid dimension.  Results are
 * bit-identical to running the pairs one at a time through StereoMatcher.
 *
 * Pair i of a batched buffer lives at base + i * pair_stride (bytes), each pair
 * pitched by the given row pitch.  compute() allocates nothing and may run
 * fewer pairs than the matcher holds (the first `count`).  capture() records
 * the whole pyramid for one set of buffers into a HIP graph, which launch()
 * replays with one call.
 */
#ifndef PYRSGM_BATCH_H
#define PYRSGM_BATCH_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#include "pyrsgm/pyrsgm_py.h"

#ifdef __cplusplus
#include <memory>

#include "pyrsgm/pyrsgm.h"

namespace pyrsgm {

/// A batched image buffer: pair i at data + i * pair_stride, rows `pitch` bytes apart.
template <class T>
struct BatchBuffer {
    T* data = nullptr;
    size_t pitch = 0;
    size_t pair_stride = 0;
};

/// Launch shapes of a BatchMatcher; every choice gives the same bits.
struct BatchTuning {
    int census_rows = 4;   ///< rows per block of the refinement census cost: 2, 4 or 8
    int paired = -1;       ///< paired lines for 16-label byte volumes: -1 by launch size, 0 off, 1 on
    int xcd_remap = -1;    ///< XCD-contiguous T->B column groups: -1 default, 0 off, 1 on
    int side_stream = 1;   ///< refinement descriptor passes on a side stream (batches of 2+)
    int modular = -1;      ///< refinement path step for 16-label byte volumes: -1 by launch
                           ///< size, 0 shifting (per-step window re-index), 1 modular (lane g
                           ///< holds disparity g mod 16, no branch)
    int fused = -1;        ///< R->L lines finish the winner-take-all (16-label byte-volume
                           ///< pyramid levels): -1 default (off), 0 off, 1 on
};

class BatchMatcher {
public:
    BatchMatcher(int width, int height, int batch, const Params& params);
    ~BatchMatcher();
    BatchMatcher(BatchMatcher&&) noexcept;
    BatchMatcher& operator=(BatchMatcher&&) noexcept;
    BatchMatcher(const BatchMatcher&) = delete;
    BatchMatcher& operator=(const BatchMatcher&) = delete;

    /// Enqueue the first `count` pairs (-1: all) on `stream`.  Null outputs are
    /// skipped as in StereoMatcher.
    void compute(BatchBuffer<const uint8_t> left, BatchBuffer<const uint8_t> right,
                 BatchBuffer<int16_t> disp16, BatchBuffer<float> disp32, BatchBuffer<float> conf,
                 hipStream_t stream = nullptr, int count = -1);
    /// Record compute() on these buffers into a graph (replaces a previous one).
    void capture(BatchBuffer<const uint8_t> left, BatchBuffer<const uint8_t> right,
                 BatchBuffer<int16_t> disp16, BatchBuffer<float> disp32, BatchBuffer<float> conf,
                 int count = -1);
    /// Replay the captured graph on `stream`.  Throws if nothing was captured.
    void launch(hipStream_t stream = nullptr);

    void setTuning(const BatchTuning& t);  ///< drops a captured graph
    BatchTuning tuning() const;
    /// Time candidate launch shapes on these buffers (reps passes each after a
    /// warm-up, coordinate descent), keep and return the fastest; drops a
    /// captured graph.  The choice is remembered per (size, batch, Params) for
    /// the life of the process and used by later matchers of that key.
    BatchTuning tune(BatchBuffer<const uint8_t> left, BatchBuffer<const uint8_t> right,
                     BatchBuffer<int16_t> disp16, BatchBuffer<float> disp32,
                     BatchBuffer<float> conf, hipStream_t stream, int reps, double* best_ms);

    int batch() const;
    int width() const;
    int height() const;
    const Params& params() const;
    int numLevels() const;
    LevelInfo levelInfo(int level) const;
    size_t deviceMemoryBytes() const;    ///< device memory owned by the matcher
    /// Device memory a matcher of these arguments would allocate (validates them).
    static size_t requiredBytes(int width, int height, int batch, const Params& params);

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

} // namespace pyrsgm

extern "C" {
#endif

/* Plain-C handles for FFI callers.  Errors: NULL / -1 with the message copied
 * into err (truncated to errlen). */
void* pyrsgm_batch_create(int width, int height, int batch, const pyrsgm_params_c* params,
                          char* err, int errlen);
void pyrsgm_batch_destroy(void* matcher);
int pyrsgm_batch_compute(void* matcher,
                         const unsigned char* d_left, size_t left_pitch, size_t left_stride,
                         const unsigned char* d_right, size_t right_pitch, size_t right_stride,
                         short* d_disp16, size_t disp16_pitch, size_t disp16_stride,
                         float* d_disp32, size_t disp32_pitch, size_t disp32_stride,
                         float* d_conf, size_t conf_pitch, size_t conf_stride,
                         int count, hipStream_t stream, char* err, int errlen);
int pyrsgm_batch_capture(void* matcher,
                         const unsigned char* d_left, size_t left_pitch, size_t left_stride,
                         const unsigned char* d_right, size_t right_pitch, size_t right_stride,
                         short* d_disp16, size_t disp16_pitch, size_t disp16_stride,
                         float* d_disp32, size_t disp32_pitch, size_t disp32_stride,
                         float* d_conf, size_t conf_pitch, size_t conf_stride,
                         int count, char* err, int errlen);
int pyrsgm_batch_launch(void* matcher, hipStream_t stream, char* err, int errlen);
/* BatchTuning as {census_rows, paired, xcd_remap, side_stream, fused, modular}. */
int pyrsgm_batch_set_tuning(void* matcher, const int tuning[6], char* err, int errlen);
int pyrsgm_batch_get_tuning(void* matcher, int tuning[6]);
int pyrsgm_batch_tune(void* matcher,
                      const unsigned char* d_left, size_t left_pitch, size_t left_stride,
                      const unsigned char* d_right, size_t right_pitch, size_t right_stride,
                      short* d_disp16, size_t disp16_pitch, size_t disp16_stride,
                      float* d_disp32, size_t disp32_pitch, size_t disp32_stride,
                      float* d_conf, size_t conf_pitch, size_t conf_stride,
                      hipStream_t stream, int reps, int best[6], double* best_ms, char* err,
                      int errlen);
size_t pyrsgm_batch_device_bytes(void* matcher);
/* BatchMatcher::requiredBytes; 0 on invalid arguments (message in err). */
size_t pyrsgm_batch_required_bytes(int width, int height, int batch,
                                   const pyrsgm_params_c* params, char* err, int errlen);

#ifdef __cplusplus
}
#endif
#endif
