// stereo_matcher.cpp -- pyrsgm::StereoMatcher and pyrsgm::BatchMatcher.
//
// Same orchestration contract as the reference (stereo_matcher.cpp): level
// geometry and parameter validation at construction, every device buffer
// allocated there, and compute() = pyrDown chain + coarse-to-fine level calls
// on the caller's stream.  The pyrDown chain processes both views in one
// launch per level.
#include "pyrsgm/pyrsgm.h"
#include "pyrsgm/pyrsgm_batch.h"

#include <algorithm>
#include <array>
#include <sstream>
#include <stdexcept>
#include <string>

#include "internal.h"
#include "pyrsgm/pyrsgm_c.h"
#include "pyrsgm/pyrsgm_prof.h"

namespace pyrsgm {
namespace {

constexpr int kMaxPyramidLevels = 4;
constexpr int kRefineLabels = 16;
constexpr int kMaxCoarseLabels = 1024;

int align16(int v) { return std::max(16, ((v + 15) / 16) * 16); }

[[noreturn]] void throwHip(const std::string& what, hipError_t err)
{
    std::ostringstream oss;
    oss << "pyrsgm: " << what << ": " << hipGetErrorString(err);
    throw std::runtime_error(oss.str());
}

void check(hipError_t err, const char* what)
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

} // namespace

// Reference: stereo_matcher.cpp:100-148.
Params validateParams(const Params& in)
{
    Params p = in;
    if (p.num_disparities <= 0 || p.num_disparities > kMaxCoarseLabels)
        throw std::invalid_argument("pyrsgm: num_disparities must be in (0, 1024]");
    p.num_disparities = align16(p.num_disparities);
    if (p.num_disparities > kMaxCoarseLabels)
        throw std::invalid_argument("pyrsgm: num_disparities must be <= 1024 after 16-alignment");
    if (p.block_size < 1) throw std::invalid_argument("pyrsgm: block_size must be >= 1");
    p.block_size = std::max(3, p.block_size | 1);
    if (p.block_size > 255) throw std::invalid_argument("pyrsgm: block_size must be <= 255");
    if (p.p1 < 0 || p.p2 < p.p1) throw std::invalid_argument("pyrsgm: require 0 <= p1 <= p2");
    if (p.uniqueness_ratio < 0 || p.uniqueness_ratio >= 100)
        throw std::invalid_argument("pyrsgm: uniqueness_ratio must be in [0, 100)");
    if (p.pre_filter_cap <= 0) throw std::invalid_argument("pyrsgm: pre_filter_cap must be > 0");
    p.num_paths = std::clamp(p.num_paths, 1, 3);
    if (p.pyramid_levels < 0 || p.pyramid_levels > kMaxPyramidLevels)
        throw std::invalid_argument("pyrsgm: pyramid_levels must be in [0, 4]");
    if (p.pyramid_levels > 0) {
        if (p.min_disparity != 0)
            throw std::invalid_argument("pyrsgm: pyramid mode requires min_disparity = 0");
        if (p.disp12_max_diff < 0)
            throw std::invalid_argument("pyrsgm: pyramid mode requires disp12_max_diff >= 0");
        if (p.pyramid_coarse_disparities <= 0)
            throw std::invalid_argument("pyrsgm: pyramid_coarse_disparities must be > 0");
        p.pyramid_coarse_disparities = std::min(kMaxCoarseLabels, align16(p.pyramid_coarse_disparities));
        if (p.pyramid_max_disparity <= 0) p.pyramid_max_disparity = p.min_disparity + p.num_disparities - 1;
        if (p.pyramid_max_disparity < p.min_disparity)
            throw std::invalid_argument("pyrsgm: pyramid_max_disparity < min_disparity");
        p.pyramid_min_confidence = std::clamp(p.pyramid_min_confidence, 0.0f, 1.0f);
    }
    if (p.adaptive_p2 && p.p2_beta <= 0.0f)
        throw std::invalid_argument("pyrsgm: adaptive_p2 requires p2_beta > 0");
    return p;
}

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

    MatcherCore(int w, int h, int pairs, const Params& p)
        : width(w), height(h), n(pairs), params(validateParams(p)), levels(params.pyramid_levels)
    {
        if (w <= 0 || h <= 0) throw std::invalid_argument("pyrsgm: empty image size");
        if (w > pyr::kMaxWidth) throw std::invalid_argument("pyrsgm: width must be <= 65535");
        if (pairs < 1) throw std::invalid_argument("pyrsgm: batch must be >= 1");
        // Reference: stereo_matcher.cpp:169-193.
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
        } catch (...) {
            destroy();
            throw;
        }
    }

    ~MatcherCore() { destroy(); }

    void destroy()
    {
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
    void compute(BatchBuffer<const uint8_t> L, BatchBuffer<const uint8_t> R,
                 BatchBuffer<int16_t> d16, BatchBuffer<float> d32, BatchBuffer<float> cf,
                 hipStream_t stream)
    {
        if (!L.data || !R.data) throw std::invalid_argument("pyrsgm: null input image");
        if (!d16.data && !d32.data && !cf.data) return;
        BatchBuffer<float> oc = cf;
        if (!oc.data && d32.data) oc = {conf[0].data, conf[0].pitch, stride(conf[0], 0)};
        const int costMode = params.cost_mode == CostMode::Census ? PYRSGM_COST_CENSUS : PYRSGM_COST_BT;
        if (levels == 0) {
            pyr::Batch b;
            b.n = n;
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
                params.p2_beta, params.p2_lower_mult, params.p2_upper_mult, b, stream);
            if (ret != 0) throwHip("single-scale SGM", hipGetLastError());
            return;
        }
        // pyrDown chain: both views in one launch when their pair strides agree.
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
                                            n, stream);
            } else {
                e = pyr::launchPyrDownBatch(sl, (int)slp, nullptr, 0, sls, parent.width,
                                            parent.height, pyrLeft[level].data,
                                            (int)pyrLeft[level].pitch, nullptr, 0, ds, li.width,
                                            li.height, n, stream);
                if (e == hipSuccess)
                    e = pyr::launchPyrDownBatch(sr, (int)srp, nullptr, 0, srs, parent.width,
                                                parent.height, pyrRight[level].data,
                                                (int)pyrRight[level].pitch, nullptr, 0, ds,
                                                li.width, li.height, n, stream);
            }
            pyr::profEnd(stream);
            if (e != hipSuccess) throwHip("pyrDown", e);
        }
        for (int level = levels; level >= 0; --level) {
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
            b.n = n;
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
                params.p2_upper_mult, b, stream);
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

struct StereoMatcher::Impl : MatcherCore {
    Impl(int w, int h, const Params& p) : MatcherCore(w, h, 1, p) {}

    void compute(const uint8_t* d_left, size_t left_pitch, const uint8_t* d_right,
                 size_t right_pitch, int16_t* d_disp16, size_t disp16_pitch, float* d_disp32,
                 size_t disp32_pitch, float* d_conf, size_t conf_pitch, hipStream_t stream)
    {
        MatcherCore::compute({d_left, left_pitch, 0}, {d_right, right_pitch, 0},
                             {d_disp16, disp16_pitch, 0}, {d_disp32, disp32_pitch, 0},
                             {d_conf, conf_pitch, 0}, stream);
    }
};

StereoMatcher::StereoMatcher(int width, int height, const Params& params)
    : impl_(new Impl(width, height, params)) {}
StereoMatcher::~StereoMatcher() = default;
StereoMatcher::StereoMatcher(StereoMatcher&&) noexcept = default;
StereoMatcher& StereoMatcher::operator=(StereoMatcher&&) noexcept = default;

void StereoMatcher::compute(const uint8_t* d_left, size_t left_pitch, const uint8_t* d_right,
                            size_t right_pitch, int16_t* d_disp16, size_t disp16_pitch,
                            float* d_disp32, size_t disp32_pitch, float* d_conf, size_t conf_pitch,
                            hipStream_t stream)
{
    impl_->compute(d_left, left_pitch, d_right, right_pitch, d_disp16, disp16_pitch, d_disp32,
                   disp32_pitch, d_conf, conf_pitch, stream);
}

int StereoMatcher::width() const { return impl_->width; }
int StereoMatcher::height() const { return impl_->height; }
const Params& StereoMatcher::params() const { return impl_->params; }
int StereoMatcher::numLevels() const { return impl_->levels + 1; }
LevelInfo StereoMatcher::levelInfo(int level) const
{
    if (level < 0 || level > impl_->levels) throw std::out_of_range("pyrsgm: level");
    return impl_->info[level];
}
size_t StereoMatcher::deviceMemoryBytes() const { return impl_->deviceBytes(); }

// ---- BatchMatcher ---------------------------------------------------------
struct BatchMatcher::Impl : MatcherCore {
    hipStream_t captureStream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    Impl(int w, int h, int n, const Params& p) : MatcherCore(w, h, n, p) {}
    ~Impl() { dropGraph(); if (captureStream) (void)hipStreamDestroy(captureStream); }

    void dropGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    // The sequence is recorded on a stream of the matcher's own (capture needs
    // a stream other than the legacy default one); launch() replays it on any.
    void capture(BatchBuffer<const uint8_t> L, BatchBuffer<const uint8_t> R,
                 BatchBuffer<int16_t> d16, BatchBuffer<float> d32, BatchBuffer<float> cf)
    {
        dropGraph();
        if (!captureStream) check(hipStreamCreateWithFlags(&captureStream, hipStreamNonBlocking),
                                  "hipStreamCreateWithFlags");
        check(hipStreamBeginCapture(captureStream, hipStreamCaptureModeThreadLocal),
              "hipStreamBeginCapture");
        try {
            compute(L, R, d16, d32, cf, captureStream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(captureStream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(captureStream, &graph), "hipStreamEndCapture");
        check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
};

BatchMatcher::BatchMatcher(int width, int height, int batch, const Params& params)
    : impl_(new Impl(width, height, batch, params)) {}
BatchMatcher::~BatchMatcher() = default;
BatchMatcher::BatchMatcher(BatchMatcher&&) noexcept = default;
BatchMatcher& BatchMatcher::operator=(BatchMatcher&&) noexcept = default;

void BatchMatcher::compute(BatchBuffer<const uint8_t> left, BatchBuffer<const uint8_t> right,
                           BatchBuffer<int16_t> disp16, BatchBuffer<float> disp32,
                           BatchBuffer<float> conf, hipStream_t stream)
{
    impl_->compute(left, right, disp16, disp32, conf, stream);
}

void BatchMatcher::capture(BatchBuffer<const uint8_t> left, BatchBuffer<const uint8_t> right,
                           BatchBuffer<int16_t> disp16, BatchBuffer<float> disp32,
                           BatchBuffer<float> conf)
{
    impl_->capture(left, right, disp16, disp32, conf);
}

void BatchMatcher::launch(hipStream_t stream)
{
    if (!impl_->exec) throw std::logic_error("pyrsgm: BatchMatcher::launch before capture");
    check(hipGraphLaunch(impl_->exec, stream), "hipGraphLaunch");
}

int BatchMatcher::batch() const { return impl_->n; }
int BatchMatcher::width() const { return impl_->width; }
int BatchMatcher::height() const { return impl_->height; }
const Params& BatchMatcher::params() const { return impl_->params; }
int BatchMatcher::numLevels() const { return impl_->levels + 1; }
LevelInfo BatchMatcher::levelInfo(int level) const
{
    if (level < 0 || level > impl_->levels) throw std::out_of_range("pyrsgm: level");
    return impl_->info[level];
}
size_t BatchMatcher::deviceMemoryBytes() const { return impl_->deviceBytes(); }

const char* version() { return "0.1.0"; }

} // namespace pyrsgm
