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
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>

#include "matcher_core.h"

namespace pyrsgm {

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
namespace {

// Tunings found by BatchMatcher::tune, per (size, batch, Params), for the life
// of the process.
std::mutex g_batchTuneMutex;
std::map<std::string, BatchTuning> g_batchTuned;

std::string batchKey(int w, int h, int n, const Params& p)
{
    std::ostringstream o;
    o << w << 'x' << h << 'x' << n << ':' << p.num_disparities << ':' << p.min_disparity << ':'
      << p.block_size << ':' << (int)p.cost_mode << ':' << p.p1 << ':' << p.p2 << ':'
      << p.uniqueness_ratio << ':' << p.disp12_max_diff << ':' << p.lr_check << ':'
      << p.median_filter << ':' << p.num_paths << ':' << p.pre_filter_cap << ':'
      << p.pyramid_levels << ':' << p.pyramid_coarse_disparities << ':'
      << p.pyramid_max_disparity << ':' << p.pyramid_min_confidence << ':' << p.adaptive_p2
      << ':' << p.p2_alpha << ':' << p.p2_beta << ':' << p.p2_lower_mult << ':'
      << p.p2_upper_mult;
    return o.str();
}

void checkBatchTuning(const BatchTuning& t)
{
    if (t.census_rows != 2 && t.census_rows != 4 && t.census_rows != 8)
        throw std::invalid_argument("pyrsgm: BatchTuning.census_rows must be 2, 4 or 8");
    if (t.paired < -1 || t.paired > 1 || t.xcd_remap < -1 || t.xcd_remap > 1)
        throw std::invalid_argument("pyrsgm: BatchTuning.paired / xcd_remap must be -1, 0 or 1");
    if (t.side_stream != 0 && t.side_stream != 1)
        throw std::invalid_argument("pyrsgm: BatchTuning.side_stream must be 0 or 1");
    if (t.fused < -1 || t.fused > 1)
        throw std::invalid_argument("pyrsgm: BatchTuning.fused must be -1, 0 or 1");
    if (t.modular < -1 || t.modular > 1)
        throw std::invalid_argument("pyrsgm: BatchTuning.modular must be -1, 0 or 1");
}

} // namespace

struct BatchMatcher::Impl : MatcherCore {
    hipStream_t captureStream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    BatchTuning tun;

    Impl(int w, int h, int n, const Params& p) : MatcherCore(w, h, n, p)
    {
        std::lock_guard<std::mutex> lock(g_batchTuneMutex);
        auto it = g_batchTuned.find(batchKey(width, height, this->n, params));
        if (it != g_batchTuned.end()) apply(it->second);
    }
    ~Impl() { dropGraph(); if (captureStream) (void)hipStreamDestroy(captureStream); }

    void apply(const BatchTuning& t)
    {
        checkBatchTuning(t);
        tun = t;
        tiles.localRows = t.census_rows;
        tiles.paired = t.paired < 0 ? 0 : (t.paired ? 2 : 1);
        tiles.xcdRemap = t.xcd_remap < 0 ? 0 : (t.xcd_remap ? 2 : 1);
        useSide = t.side_stream != 0;
        tiles.fused = t.fused < 0 ? 0 : (t.fused ? 2 : 1);
        tiles.modular = t.modular < 0 ? 0 : (t.modular ? 2 : 1);
    }

    double timePasses(BatchBuffer<const uint8_t> L, BatchBuffer<const uint8_t> R,
                      BatchBuffer<int16_t> d16, BatchBuffer<float> d32, BatchBuffer<float> cf,
                      hipStream_t s, int reps)
    {
        hipEvent_t a = nullptr, b = nullptr;
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventCreate(&b), "hipEventCreate");
        compute(L, R, d16, d32, cf, s, -1);  // warm-up
        check(hipEventRecord(a, s), "hipEventRecord");
        for (int r = 0; r < reps; ++r) compute(L, R, d16, d32, cf, s, -1);
        check(hipEventRecord(b, s), "hipEventRecord");
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return (double)ms / reps;
    }

    // Coordinate descent from the current tuning over each knob in turn.
    BatchTuning tune(BatchBuffer<const uint8_t> L, BatchBuffer<const uint8_t> R,
                     BatchBuffer<int16_t> d16, BatchBuffer<float> d32, BatchBuffer<float> cf,
                     hipStream_t s, int reps, double* bestMs)
    {
        if (reps < 1) reps = 1;
        dropGraph();
        BatchTuning best = tun;
        double bestT = timePasses(L, R, d16, d32, cf, s, reps);
        auto consider = [&](const BatchTuning& t) {
            apply(t);
            const double ms = timePasses(L, R, d16, d32, cf, s, reps);
            if (ms < bestT) {
                bestT = ms;
                best = t;
            }
        };
        for (int rows : {2, 4, 8})
            if (rows != best.census_rows) { BatchTuning t = best; t.census_rows = rows; consider(t); }
        for (int pk : {-1, 0, 1})
            if (pk != best.paired) { BatchTuning t = best; t.paired = pk; consider(t); }
        for (int xr : {-1, 0, 1})
            if (xr != best.xcd_remap) { BatchTuning t = best; t.xcd_remap = xr; consider(t); }
        for (int fu : {-1, 0, 1})
            if (fu != best.fused) { BatchTuning t = best; t.fused = fu; consider(t); }
        for (int md : {-1, 0, 1})
            if (md != best.modular) { BatchTuning t = best; t.modular = md; consider(t); }
        if (side)
            for (int sd : {0, 1})
                if (sd != best.side_stream) { BatchTuning t = best; t.side_stream = sd; consider(t); }
        apply(best);
        {
            std::lock_guard<std::mutex> lock(g_batchTuneMutex);
            g_batchTuned[batchKey(width, height, n, params)] = best;
        }
        if (bestMs) *bestMs = bestT;
        return best;
    }

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
                 BatchBuffer<int16_t> d16, BatchBuffer<float> d32, BatchBuffer<float> cf,
                 int count)
    {
        if (count < -1 || count == 0 || count > n)
            throw std::invalid_argument("pyrsgm: pair count outside [1, batch]");
        dropGraph();
        if (!captureStream) check(hipStreamCreateWithFlags(&captureStream, hipStreamNonBlocking),
                                  "hipStreamCreateWithFlags");
        check(hipStreamBeginCapture(captureStream, hipStreamCaptureModeThreadLocal),
              "hipStreamBeginCapture");
        try {
            compute(L, R, d16, d32, cf, captureStream, count);
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
                           BatchBuffer<float> conf, hipStream_t stream, int count)
{
    impl_->compute(left, right, disp16, disp32, conf, stream, count);
}

void BatchMatcher::capture(BatchBuffer<const uint8_t> left, BatchBuffer<const uint8_t> right,
                           BatchBuffer<int16_t> disp16, BatchBuffer<float> disp32,
                           BatchBuffer<float> conf, int count)
{
    impl_->capture(left, right, disp16, disp32, conf, count);
}

void BatchMatcher::launch(hipStream_t stream)
{
    if (!impl_->exec) throw std::logic_error("pyrsgm: BatchMatcher::launch before capture");
    check(hipGraphLaunch(impl_->exec, stream), "hipGraphLaunch");
}

void BatchMatcher::setTuning(const BatchTuning& t)
{
    impl_->dropGraph();
    impl_->apply(t);
}

BatchTuning BatchMatcher::tuning() const { return impl_->tun; }

BatchTuning BatchMatcher::tune(BatchBuffer<const uint8_t> left, BatchBuffer<const uint8_t> right,
                               BatchBuffer<int16_t> disp16, BatchBuffer<float> disp32,
                               BatchBuffer<float> conf, hipStream_t stream, int reps,
                               double* best_ms)
{
    if (!left.data || !right.data) throw std::invalid_argument("pyrsgm: null input image");
    return impl_->tune(left, right, disp16, disp32, conf, stream, reps, best_ms);
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
size_t BatchMatcher::requiredBytes(int width, int height, int batch, const Params& params)
{
    return MatcherCore::requiredBytes(width, height, batch, params);
}

const char* version() { return "0.1.0"; }

} // namespace pyrsgm
