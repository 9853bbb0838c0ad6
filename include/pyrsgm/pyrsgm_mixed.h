/*
 * pyrsgm_mixed.h -- mixed batch: many stereo pairs of different sizes,
 * disparity ranges, pyramid depths, cost modes and post-processing in one
 * fused pass.
 *
 * No counterpart in gegeges/pyrsgm (whose only multi-pair form is several
 * StereoMatchers on several streams).  A MixedBatch holds every pair's
 * resident device state.  Its pyramids run step by step from the coarse end
 * (step s is level levels_i - s of pair i):
 *   - the path recurrences of all pairs at a step that share a kernel
 *     specialisation run as ONE persistent launch whose workgroups take
 *     (pair level, direction, line group) items from a device queue, and
 *     their winner-take-all as ONE launch over the same task table;
 *   - the per-pair stages around them (pyrDown, descriptors, costs,
 *     post-processing) run on parallel streams that join before and fork
 *     after each step's queued launches.
 * capture() records the whole pass into one HIP graph.  tune() times
 * candidate launch shapes (streams for the per-pair stages, resident
 * workgroups per CU of the queued launches, paired lines) on the device and
 * keeps the fastest.  Every result is bit-identical to the same pair matched
 * alone by StereoMatcher, whatever the tuning.
 *
 * Calls on one MixedBatch must not overlap (they share its queues and
 * buffers); different MixedBatches may run concurrently.
 */
#ifndef PYRSGM_MIXED_H
#define PYRSGM_MIXED_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#include "pyrsgm/pyrsgm_py.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One pair's device buffers (row pitches in bytes).  Null outputs are skipped
 * as in StereoMatcher::compute; at least one output is required. */
typedef struct {
    const unsigned char* left;
    size_t left_pitch;
    const unsigned char* right;
    size_t right_pitch;
    short* disp16;
    size_t disp16_pitch;
    float* disp32;
    size_t disp32_pitch;
    float* conf;
    size_t conf_pitch;
} pyrsgm_pair_io_c;

/* Launch shape.  branches: streams for the per-pair stages (1..4);
 * blocks_per_cu: resident workgroups per CU of each queued path launch
 * (0: the kernel's occupancy); paired: two lines per lane for 16-label byte
 * volumes (-1: by launch size, 0: off, 1: on). */
typedef struct {
    int branches, blocks_per_cu, paired, queued_stages;
} pyrsgm_mixed_tuning_c;

void* pyrsgm_mixed_create(int n, const int* widths, const int* heights,
                          const pyrsgm_params_c* params, char* err, int errlen);
void pyrsgm_mixed_destroy(void* batch);
int pyrsgm_mixed_compute(void* batch, const pyrsgm_pair_io_c* io, hipStream_t stream,
                         char* err, int errlen);
int pyrsgm_mixed_capture(void* batch, const pyrsgm_pair_io_c* io, char* err, int errlen);
int pyrsgm_mixed_launch(void* batch, hipStream_t stream, char* err, int errlen);
int pyrsgm_mixed_set_tuning(void* batch, const pyrsgm_mixed_tuning_c* t, char* err, int errlen);
int pyrsgm_mixed_get_tuning(void* batch, pyrsgm_mixed_tuning_c* out);
/* Times the candidates (reps runs each, after one warm-up) on `stream`, keeps
 * the fastest and returns it with its mean milliseconds per pass.  Drops a
 * captured graph. */
int pyrsgm_mixed_tune(void* batch, const pyrsgm_pair_io_c* io, hipStream_t stream, int reps,
                      pyrsgm_mixed_tuning_c* best, double* best_ms, char* err, int errlen);
/* out[6] = {queued path launches (each followed by one queued winner-take-all
 * launch), tasks in them, work items in them, per-pair stage launches, queued
 * stage launches (pyrDown, descriptors, costs, post), tasks in them} */
int pyrsgm_mixed_plan(void* batch, long long out[6]);
size_t pyrsgm_mixed_device_bytes(void* batch);

#ifdef __cplusplus
}

#include <memory>
#include <vector>

#include "pyrsgm/pyrsgm.h"

namespace pyrsgm {

struct PairSpec {
    int width = 0, height = 0;
    Params params;
};

struct MixedTuning {
    int branches = 4;       ///< streams of the per-pair stage launches (queued_stages = 0)
    int blocks_per_cu = 0;  ///< resident workgroups per CU of the queued path launches (0: occupancy)
    int paired = -1;        ///< paired lines: -1 by launch size, 0 off, 1 on
    int queued_stages = 1;  ///< pyrDown, descriptors, costs and post as queued launches (1) or per pair (0)
};

class MixedBatch {
public:
    explicit MixedBatch(const std::vector<PairSpec>& pairs, const MixedTuning& tuning = {});
    ~MixedBatch();
    MixedBatch(const MixedBatch&) = delete;
    MixedBatch& operator=(const MixedBatch&) = delete;

    /// Enqueue every pair on `stream` (io has one entry per pair).
    void compute(const pyrsgm_pair_io_c* io, hipStream_t stream = nullptr);
    /// Record compute() on these buffers into a graph (replaces a previous one).
    void capture(const pyrsgm_pair_io_c* io);
    /// Replay the captured graph on `stream`.  Throws if nothing was captured.
    void launch(hipStream_t stream = nullptr);

    void setTuning(const MixedTuning& t);
    MixedTuning tuning() const;
    /// Time the candidate tunings on these buffers, keep and return the fastest.
    MixedTuning tune(const pyrsgm_pair_io_c* io, hipStream_t stream, int reps, double* best_ms);

    int pairs() const;
    const Params& params(int i) const;
    size_t deviceMemoryBytes() const;
    /// {queued path launches, tasks, work items, per-pair stage launches,
    ///  queued stage launches, stage tasks}
    std::vector<long long> plan() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

} // namespace pyrsgm
#endif
#endif
