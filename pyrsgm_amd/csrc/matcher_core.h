// matcher_core.h -- level geometry, device state and the coarse-to-fine
// sequence shared by StereoMatcher (one pair), BatchMatcher (n pairs of one
// size) and the serving lanes.  Internal to the library.
#ifndef PYRSGM_MATCHER_CORE_H
#define PYRSGM_MATCHER_CORE_H

#include <algorithm>
#include <array>
#include <sstream>
#include <stdexcept>
#include <string>

#include "internal.h"
#include "pyrsgm/pyrsgm.h"
#include "pyrsgm/pyrsgm_batch.h"
#include "pyrsgm/pyrsgm_c.h"
#include "pyrsgm/pyrsgm_prof.h"

namespace pyrsgm {
namespace detail {

constexpr int kMaxPyramidLevels = 4;
constexpr int kRefineLabels = 16;
constexpr int kMaxCoarseLabels = 1024;

inline int align16(int v) { return std::max(16, ((v + 15) / 16) * 16); }

[[noreturn]] inline void throwHip(const std::string& what, hipError_t err)
{
    std::ostringstream oss;
    oss << "pyrsgm: " << what << ": " << hipGetErrorString(err);
    throw std::runtime_error(oss.str());
}

inline void check(hipError_t err, const char* what)
{
    if (err != hipSuccess) throwHip(what, err);
}

template <class T>
struct DeviceImage {
    T* data = nullptr;
    size_t pitch = 0;  // bytes
    int width = 0, height = 0;

    DeviceImage() = default;
    DeviceImage(const DeviceImage&) = delete;
    DeviceImage& operator=(const DeviceImage&) = delete;
    ~DeviceImage() { release(); }

    void create(int w, int h)
    {
        release();
        if (w <= 0 || h <= 0) return;
        void* p = nullptr;
        check(hipMallocPitch(&p, &pitch, (size_t)w * sizeof(T), (size_t)h), "hipMallocPitch");
        data = static_cast<T*>(p);
        width = w;
        height = h;
    }
    void release()
    {
        if (data) (void)hipFree(data);
        data = nullptr;
        pitch = 0;
        width = height = 0;
    }
    size_t bytes() const { return pitch * (size_t)height; }
};


} // namespace detail

using detail::check;
using detail::throwHip;
using detail::DeviceImage;
using detail::align16;
using detail::kMaxPyramidLevels;
using detail::kRefineLabels;
using detail::kMaxCoarseLabels;

// Level geometry, device state and the coarse-to-fine sequence for n pairs of
// one size; StereoMatcher is n = 1, BatchMatcher any n.  Per-pair images of
// one level are stacked: pair i's rows follow pair i-1's.
struct MatcherCore {
    int width = 0, height = 0, n = 1;
    Params params;
    int levels = 0;
    std::array<void*, kMaxPyramidLevels + 1> workspaces{};
    std::array<DeviceImage<uint8_t>, kMaxPyramidLevels + 1> pyrLeft, pyrRight;  // [0] unused
    std::array<DeviceImage<int16_t>, kMaxPyramidLevels + 1> disp16;            // [0] unused
    std::array<DeviceImage<float>, kMaxPyramidLevels + 1> conf;                // [0] scratch
    std::array<LevelInfo, kMaxPyramidLevels + 1> info{};
    size_t workspaceBytes = 0;
    // The refinement levels' descriptor passes depend only on the pyramid images,
    // so for several pairs compute() runs them on `side` while the coarse level
    // runs on the caller's stream (fork / join through two events; graph-capture
    // safe).  For one pair the cross-stream synchronisation costs more than the
    // overlap saves (measured), and everything stays on the caller's stream.
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    bool useSide = true;      // run the descriptor passes on `side` when it exists
    pyr::TileConfig tiles{};  // launch shapes (all give the same bits)

    // Reference: stereo_matcher.cpp:169-193.  Fills info[0..levels].
    static void geometry(int w, int h, const Params& params,
                         std::array<LevelInfo, kMaxPyramidLevels + 1>& info)
    {
        const int levels = params.pyramid_levels;
        int lw = w, lh = h;
        for (int level = 0; level <= levels; ++level) {
            if (level > 0) {
                lw = (lw + 1) / 2;
                lh = (lh + 1) / 2;
            }
            LevelInfo& li = info[level];
            li.level = level;
            li.width = lw;
            li.height = lh;
            if (levels == 0) {
                li.refine = false;
                li.num_labels = params.num_disparities;
                li.max_disparity = params.min_disparity + params.num_disparities - 1;
                li.block_size = params.block_size;
            } else {
                li.refine = level != levels;
                li.max_disparity = std::min(params.pyramid_max_disparity >> level, lw - 1);
                li.num_labels = li.refine ? kRefineLabels
                    : align16(std::min(params.pyramid_coarse_disparities, li.max_disparity + 1));
                li.block_size = level >= 2 ? 3 : params.block_size;
            }
        }
    }

    static void checkSize(int w, int h, int pairs)
    {
        if (w <= 0 || h <= 0) throw std::invalid_argument("pyrsgm: empty image size");
        if (w > pyr::kMaxWidth) throw std::invalid_argument("pyrsgm: width must be <= 65535");
        if (pairs < 1) throw std::invalid_argument("pyrsgm: batch must be >= 1");
    }

    // Device bytes a core of n pairs allocates (row pitches rounded to 256 B,
    // the allocator's pitch granule).
    static size_t requiredBytes(int w, int h, int pairs, const Params& p)
    {
        const Params params = validateParams(p);
        checkSize(w, h, pairs);
        std::array<LevelInfo, kMaxPyramidLevels + 1> info{};
        geometry(w, h, params, info);
        auto img = [&](int lw, int lh, size_t elem) {
            return ((size_t)lw * elem + 255) / 256 * 256 * (size_t)lh * (size_t)pairs;
        };
        size_t total = img(w, h, sizeof(float));  // level-0 confidence scratch
        for (int level = 0; level <= params.pyramid_levels; ++level) {
            const LevelInfo& li = info[level];
            total += pyr::workspaceBytes(li.width, li.height, li.num_labels) * (size_t)pairs;
            if (level > 0)
                total += img(li.width, li.height, 1) * 2 + img(li.width, li.height, 2) +
                         img(li.width, li.height, 4);
        }
        return total;
    }

    MatcherCore(int w, int h, int pairs, const Params& p)
        : width(w), height(h), n(pairs), params(validateParams(p)), levels(params.pyramid_levels)
    {
        checkSize(w, h, pairs);
        geometry(w, h, params, info);
        try {
            for (int level = 0; level <= levels; ++level) {
                workspaces[level] = pyrsgm_create_workspace();
                if (!workspaces[level]) throw std::runtime_error("pyrsgm: workspace allocation failed");
                const LevelInfo& li = info[level];
                if (pyr::ensureWorkspace(static_cast<pyr::Workspace*>(workspaces[level]), li.width,
                                         li.height, li.num_labels, n) != 0)
                    throwHip("workspace reserve", hipErrorOutOfMemory);
                workspaceBytes += pyr::workspaceBytes(li.width, li.height, li.num_labels) * n;
                if (level > 0) {
                    pyrLeft[level].create(li.width, li.height * n);
                    pyrRight[level].create(li.width, li.height * n);
                    disp16[level].create(li.width, li.height * n);
                    conf[level].create(li.width, li.height * n);
                }
            }
            conf[0].create(w, h * n);
            if (levels > 0 && n > 1) {
                check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
                check(hipEventCreateWithFlags(&evFork, hipEventDisableTiming), "hipEventCreateWithFlags");
                check(hipEventCreateWithFlags(&evJoin, hipEventDisableTiming), "hipEventCreateWithFlags");
            }
        } catch (...) {
            destroy();
            throw;
        }
    }

    ~MatcherCore() { destroy(); }

    void destroy()
    {
        if (side) (void)hipStreamSynchronize(side);
        if (evFork) (void)hipEventDestroy(evFork);
        if (evJoin) (void)hipEventDestroy(evJoin);
        if (side) (void)hipStreamDestroy(side);
        side = nullptr;
        evFork = evJoin = nullptr;
        for (auto& ws : workspaces) {
            if (ws) pyrsgm_destroy_workspace(ws);
            ws = nullptr;
        }
    }

    // Reference: stereo_matcher.cpp:229-236.
    int postFlags(int level) const
    {
        if (level != 0) return PYRSGM_POST_LRCHECK;
        return (params.lr_check ? PYRSGM_POST_LRCHECK : 0) | (params.median_filter ? PYRSGM_POST_MEDIAN : 0);
    }

    // Pair stride of a stacked per-level image.
    template <class T>
    size_t stride(const DeviceImage<T>& im, int level) const
    {
        return im.pitch * (size_t)info[level].height;
    }

    // Reference: stereo_matcher.cpp:238-338.
    // The first `count` pairs of the buffers (count <= n; -1 = all n).
    void compute(BatchBuffer<const uint8_t> L, BatchBuffer<const uint8_t> R,
                 BatchBuffer<int16_t> d16, BatchBuffer<float> d32, BatchBuffer<float> cf,
                 hipStream_t stream, int count = -1)
    {
        if (count < 0) count = n;
        if (count < 1 || count > n) throw std::invalid_argument("pyrsgm: pair count outside [1, batch]");
        if (!L.data || !R.data) throw std::invalid_argument("pyrsgm: null input image");
        if (!d16.data && !d32.data && !cf.data) return;
        const BatchBuffer<float> oc = confRoute(d32, cf);
        if (levels == 0) {
            singleCall(L, R, d16, d32, oc, stream, pyr::kStagesAll, count);
            return;
        }
        pyrDownChain(L, R, stream, count);
        // Fork: the refinement levels' descriptor passes on the side stream.
        hipStream_t side = useSide ? this->side : nullptr;
        if (side) {
            check(hipEventRecord(evFork, stream), "hipEventRecord");
            check(hipStreamWaitEvent(side, evFork, 0), "hipStreamWaitEvent");
            for (int level = levels - 1; level >= 0; --level) levelCall(level, L, R, d16, d32, oc, side, pyr::kStagePrep, count);
            check(hipEventRecord(evJoin, side), "hipEventRecord");
        }
        for (int level = levels; level >= 0; --level) {
            if (side && level == levels - 1) check(hipStreamWaitEvent(stream, evJoin, 0), "hipStreamWaitEvent");
            levelCall(level, L, R, d16, d32, oc, stream,
                      (side && level < levels) ? pyr::kStageRest : pyr::kStagesAll, count);
        }
    }

    // Output routing (stereo_matcher.cpp:248-253): float disparities without a
    // confidence output still need one, so the level-0 scratch takes it.
    BatchBuffer<float> confRoute(BatchBuffer<float> d32, BatchBuffer<float> cf) const
    {
        BatchBuffer<float> oc = cf;
        if (!oc.data && d32.data) oc = {conf[0].data, conf[0].pitch, stride(conf[0], 0)};
        return oc;
    }

    // Single-scale matcher (levels == 0): the given stages of the one level.
    void singleCall(BatchBuffer<const uint8_t> L, BatchBuffer<const uint8_t> R,
                    BatchBuffer<int16_t> d16, BatchBuffer<float> d32, BatchBuffer<float> oc,
                    hipStream_t stream, int stages, int count, pyr::PathSink* sink = nullptr,
                    pyr::StageSink* stageSink = nullptr)
    {
        const int costMode = params.cost_mode == CostMode::Census ? PYRSGM_COST_CENSUS : PYRSGM_COST_BT;
        pyr::Batch b;
        b.n = count;
        b.left = L.pair_stride;
        b.right = R.pair_stride;
        b.out16 = d16.pair_stride;
        b.out32 = d32.pair_stride;
        b.outc = oc.pair_stride;
        const int ret = pyr::computeSingleScale(
            workspaces[0], L.data, R.data, d16.data, d32.data, oc.data, L.pitch, R.pitch,
            d16.pitch, d32.pitch, oc.pitch, width, height, params.num_disparities,
            params.min_disparity, params.block_size, params.p1, params.p2,
            params.uniqueness_ratio, params.disp12_max_diff, params.pre_filter_cap, costMode,
            postFlags(0), params.num_paths, params.adaptive_p2 ? 1 : 0, params.p2_alpha,
            params.p2_beta, params.p2_lower_mult, params.p2_upper_mult, b, stream, stages, sink,
            &tiles, stageSink);
        if (ret != 0) throwHip("single-scale SGM", hipGetLastError());
    }

    // The pyramid images of levels 1..levels: both views in one launch per
    // level when their pair strides agree.
    void pyrDownChain(BatchBuffer<const uint8_t> L, BatchBuffer<const uint8_t> R,
                      hipStream_t stream, int count, pyr::StageSink* stageSink = nullptr)
    {
        for (int level = 1; level <= levels; ++level) {
            const LevelInfo& parent = info[level - 1];
            const LevelInfo& li = info[level];
            const bool top = level == 1;
            const uint8_t* sl = top ? L.data : pyrLeft[level - 1].data;
            const uint8_t* sr = top ? R.data : pyrRight[level - 1].data;
            const size_t slp = top ? L.pitch : pyrLeft[level - 1].pitch;
            const size_t srp = top ? R.pitch : pyrRight[level - 1].pitch;
            const size_t sls = top ? L.pair_stride : stride(pyrLeft[level - 1], level - 1);
            const size_t srs = top ? R.pair_stride : stride(pyrRight[level - 1], level - 1);
            const size_t ds = stride(pyrLeft[level], level);
            pyr::profBegin(PYRSGM_STAGE_PYRDOWN, li.width, li.height, 0, stream);
            hipError_t e;
            if (sls == srs) {
                e = pyr::launchPyrDownBatch(sl, (int)slp, sr, (int)srp, sls, parent.width,
                                            parent.height, pyrLeft[level].data,
                                            (int)pyrLeft[level].pitch, pyrRight[level].data,
                                            (int)pyrRight[level].pitch, ds, li.width, li.height,
                                            count, stream, stageSink);
            } else {
                e = pyr::launchPyrDownBatch(sl, (int)slp, nullptr, 0, sls, parent.width,
                                            parent.height, pyrLeft[level].data,
                                            (int)pyrLeft[level].pitch, nullptr, 0, ds, li.width,
                                            li.height, count, stream, stageSink);
                if (e == hipSuccess)
                    e = pyr::launchPyrDownBatch(sr, (int)srp, nullptr, 0, srs, parent.width,
                                                parent.height, pyrRight[level].data,
                                                (int)pyrRight[level].pitch, nullptr, 0, ds,
                                                li.width, li.height, count, stream, stageSink);
            }
            pyr::profEnd(stream);
            if (e != hipSuccess) throwHip("pyrDown", e);
        }
    }

    // One pyramid level (or the given stages of it) for the first `count` pairs.
    void levelCall(int level, BatchBuffer<const uint8_t> L, BatchBuffer<const uint8_t> R,
                   BatchBuffer<int16_t> d16, BatchBuffer<float> d32, BatchBuffer<float> oc,
                   hipStream_t stream, int stages, int count, pyr::PathSink* sink = nullptr,
                   pyr::StageSink* stageSink = nullptr)
    {
        const int costMode = params.cost_mode == CostMode::Census ? PYRSGM_COST_CENSUS : PYRSGM_COST_BT;
        {
            const LevelInfo& li = info[level];
            const bool refine = li.refine;
            const bool fine = level == 0;
            const uint8_t* Lp = fine ? L.data : pyrLeft[level].data;
            const uint8_t* Rp = fine ? R.data : pyrRight[level].data;
            const size_t lp = fine ? L.pitch : pyrLeft[level].pitch;
            const size_t rp = fine ? R.pitch : pyrRight[level].pitch;
            const DeviceImage<int16_t>* pd = refine ? &disp16[level + 1] : nullptr;
            const DeviceImage<float>* pc = refine ? &conf[level + 1] : nullptr;
            int16_t* o16 = fine ? d16.data : disp16[level].data;
            const size_t o16p = fine ? d16.pitch : disp16[level].pitch;
            float* o32 = fine ? d32.data : nullptr;
            const size_t o32p = fine ? d32.pitch : 0;
            float* oc2 = fine ? oc.data : conf[level].data;
            const size_t ocp = fine ? oc.pitch : conf[level].pitch;
            pyr::Batch b;
            b.n = count;
            b.left = fine ? L.pair_stride : stride(pyrLeft[level], level);
            b.right = fine ? R.pair_stride : stride(pyrRight[level], level);
            b.pdisp = refine ? stride(*pd, level + 1) : 0;
            b.pconf = refine ? stride(*pc, level + 1) : 0;
            b.out16 = fine ? d16.pair_stride : stride(disp16[level], level);
            b.out32 = fine ? d32.pair_stride : 0;
            b.outc = fine ? oc.pair_stride : stride(conf[level], level);
            const int ret = pyr::computeLevel(
                workspaces[level], Lp, Rp, pd ? pd->data : nullptr, pc ? pc->data : nullptr, o16,
                o32, oc2, lp, rp, pd ? pd->pitch : 0, pc ? pc->pitch : 0, o16p, o32p, ocp,
                li.width, li.height, pd ? info[level + 1].width : 0,
                pd ? info[level + 1].height : 0, li.num_labels, params.min_disparity,
                li.max_disparity, li.block_size, params.p1, params.p2, params.uniqueness_ratio,
                params.disp12_max_diff, params.pre_filter_cap, costMode, postFlags(level),
                params.num_paths, refine ? 1 : 0, params.pyramid_min_confidence,
                params.adaptive_p2 ? 1 : 0, params.p2_alpha, params.p2_beta, params.p2_lower_mult,
                params.p2_upper_mult, b, stream, stages, sink, &tiles, stageSink);
            if (ret != 0) {
                std::ostringstream oss;
                oss << "pyramid level " << level << " (" << li.width << "x" << li.height
                    << ", D=" << li.num_labels << ", refine=" << refine << ")";
                throwHip(oss.str(), hipGetLastError());
            }
        }
    }

    size_t deviceBytes() const
    {
        size_t total = workspaceBytes;
        for (int level = 0; level <= levels; ++level)
            total += pyrLeft[level].bytes() + pyrRight[level].bytes() + disp16[level].bytes() +
                     conf[level].bytes();
        return total;
    }
};

} // namespace pyrsgm
#endif
