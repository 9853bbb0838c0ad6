// matcher_capi.cpp -- plain-C handles on pyrsgm::StereoMatcher (pyrsgm_py.h),
// pyrsgm::BatchMatcher (pyrsgm_batch.h), pyrsgm::Server (pyrsgm_serve.h) and
// pyrsgm::MixedBatch (pyrsgm_mixed.h).
#include <stdio.h>
#include <string.h>

#include <exception>
#include <new>

#include "pyrsgm/pyrsgm.h"
#include "pyrsgm/pyrsgm_batch.h"
#include "pyrsgm/pyrsgm_mixed.h"
#include "pyrsgm/pyrsgm_py.h"
#include "pyrsgm/pyrsgm_serve.h"

namespace {

void copyErr(const char* what, char* err, int errlen)
{
    if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", what);
}

pyrsgm::Params toParams(const pyrsgm_params_c& c)
{
    pyrsgm::Params p;
    p.num_disparities = c.num_disparities;
    p.min_disparity = c.min_disparity;
    p.block_size = c.block_size;
    p.cost_mode = c.cost_mode ? pyrsgm::CostMode::Census : pyrsgm::CostMode::BirchfieldTomasi;
    p.p1 = c.p1;
    p.p2 = c.p2;
    p.uniqueness_ratio = c.uniqueness_ratio;
    p.disp12_max_diff = c.disp12_max_diff;
    p.lr_check = c.lr_check != 0;
    p.median_filter = c.median_filter != 0;
    p.num_paths = c.num_paths;
    p.pre_filter_cap = c.pre_filter_cap;
    p.pyramid_levels = c.pyramid_levels;
    p.pyramid_coarse_disparities = c.pyramid_coarse_disparities;
    p.pyramid_max_disparity = c.pyramid_max_disparity;
    p.pyramid_min_confidence = c.pyramid_min_confidence;
    p.adaptive_p2 = c.adaptive_p2 != 0;
    p.p2_alpha = c.p2_alpha;
    p.p2_beta = c.p2_beta;
    p.p2_lower_mult = c.p2_lower_mult;
    p.p2_upper_mult = c.p2_upper_mult;
    return p;
}

pyrsgm_params_c fromParams(const pyrsgm::Params& p)
{
    pyrsgm_params_c c;
    c.num_disparities = p.num_disparities;
    c.min_disparity = p.min_disparity;
    c.block_size = p.block_size;
    c.cost_mode = p.cost_mode == pyrsgm::CostMode::Census ? 1 : 0;
    c.p1 = p.p1;
    c.p2 = p.p2;
    c.uniqueness_ratio = p.uniqueness_ratio;
    c.disp12_max_diff = p.disp12_max_diff;
    c.lr_check = p.lr_check ? 1 : 0;
    c.median_filter = p.median_filter ? 1 : 0;
    c.num_paths = p.num_paths;
    c.pre_filter_cap = p.pre_filter_cap;
    c.pyramid_levels = p.pyramid_levels;
    c.pyramid_coarse_disparities = p.pyramid_coarse_disparities;
    c.pyramid_max_disparity = p.pyramid_max_disparity;
    c.pyramid_min_confidence = p.pyramid_min_confidence;
    c.adaptive_p2 = p.adaptive_p2 ? 1 : 0;
    c.p2_alpha = p.p2_alpha;
    c.p2_beta = p.p2_beta;
    c.p2_lower_mult = p.p2_lower_mult;
    c.p2_upper_mult = p.p2_upper_mult;
    return c;
}

} // namespace

namespace {
pyrsgm::BatchTuning batchTuning(const int t[6])
{
    pyrsgm::BatchTuning b;
    b.census_rows = t[0];
    b.paired = t[1];
    b.xcd_remap = t[2];
    b.side_stream = t[3];
    b.fused = t[4];
    b.modular = t[5];
    return b;
}

void batchTuningOut(const pyrsgm::BatchTuning& b, int t[6])
{
    t[0] = b.census_rows;
    t[1] = b.paired;
    t[2] = b.xcd_remap;
    t[3] = b.side_stream;
    t[4] = b.fused;
    t[5] = b.modular;
}
} // namespace

extern "C" {

int pyrsgm_validate_params(const pyrsgm_params_c* in, pyrsgm_params_c* out, char* err, int errlen)
{
    if (!in || !out) return -1;
    try {
        *out = fromParams(pyrsgm::validateParams(toParams(*in)));
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
        return -1;
    }
}

void* pyrsgm_matcher_create(int width, int height, const pyrsgm_params_c* params, char* err,
                            int errlen)
{
    if (!params) return nullptr;
    try {
        return new pyrsgm::StereoMatcher(width, height, toParams(*params));
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
        return nullptr;
    }
}

void pyrsgm_matcher_destroy(void* matcher)
{
    delete static_cast<pyrsgm::StereoMatcher*>(matcher);
}

int pyrsgm_matcher_compute(void* matcher, const uint8_t* d_left, size_t left_pitch,
                           const uint8_t* d_right, size_t right_pitch, int16_t* d_disp16,
                           size_t disp16_pitch, float* d_disp32, size_t disp32_pitch,
                           float* d_conf, size_t conf_pitch, hipStream_t stream, char* err,
                           int errlen)
{
    if (!matcher) return -1;
    try {
        static_cast<pyrsgm::StereoMatcher*>(matcher)->compute(
            d_left, left_pitch, d_right, right_pitch, d_disp16, disp16_pitch, d_disp32,
            disp32_pitch, d_conf, conf_pitch, stream);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
        return -1;
    }
}

int pyrsgm_matcher_level_info(void* matcher, int level, int out[7])
{
    if (!matcher || !out) return -1;
    try {
        const pyrsgm::LevelInfo li = static_cast<pyrsgm::StereoMatcher*>(matcher)->levelInfo(level);
        out[0] = li.level;
        out[1] = li.width;
        out[2] = li.height;
        out[3] = li.num_labels;
        out[4] = li.max_disparity;
        out[5] = li.block_size;
        out[6] = li.refine ? 1 : 0;
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int pyrsgm_matcher_num_levels(void* matcher)
{
    return matcher ? static_cast<pyrsgm::StereoMatcher*>(matcher)->numLevels() : -1;
}

int pyrsgm_matcher_params(void* matcher, pyrsgm_params_c* out)
{
    if (!matcher || !out) return -1;
    *out = fromParams(static_cast<pyrsgm::StereoMatcher*>(matcher)->params());
    return 0;
}

size_t pyrsgm_matcher_device_bytes(void* matcher)
{
    return matcher ? static_cast<pyrsgm::StereoMatcher*>(matcher)->deviceMemoryBytes() : 0;
}

const char* pyrsgm_version(void) { return pyrsgm::version(); }


// ---- BatchMatcher handles (pyrsgm_batch.h) --------------------------------
void* pyrsgm_batch_create(int width, int height, int batch, const pyrsgm_params_c* params,
                          char* err, int errlen)
{
    if (!params) {
        copyErr("pyrsgm: null params", err, errlen);
        return nullptr;
    }
    try {
        return new pyrsgm::BatchMatcher(width, height, batch, toParams(*params));
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    } catch (...) {
        copyErr("pyrsgm: unknown error", err, errlen);
    }
    return nullptr;
}

void pyrsgm_batch_destroy(void* matcher) { delete static_cast<pyrsgm::BatchMatcher*>(matcher); }

int pyrsgm_batch_compute(void* matcher,
                         const unsigned char* d_left, size_t left_pitch, size_t left_stride,
                         const unsigned char* d_right, size_t right_pitch, size_t right_stride,
                         short* d_disp16, size_t disp16_pitch, size_t disp16_stride,
                         float* d_disp32, size_t disp32_pitch, size_t disp32_stride,
                         float* d_conf, size_t conf_pitch, size_t conf_stride,
                         int count, hipStream_t stream, char* err, int errlen)
{
    if (!matcher) return -1;
    try {
        static_cast<pyrsgm::BatchMatcher*>(matcher)->compute(
            {d_left, left_pitch, left_stride}, {d_right, right_pitch, right_stride},
            {d_disp16, disp16_pitch, disp16_stride}, {d_disp32, disp32_pitch, disp32_stride},
            {d_conf, conf_pitch, conf_stride}, stream, count);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_batch_capture(void* matcher,
                         const unsigned char* d_left, size_t left_pitch, size_t left_stride,
                         const unsigned char* d_right, size_t right_pitch, size_t right_stride,
                         short* d_disp16, size_t disp16_pitch, size_t disp16_stride,
                         float* d_disp32, size_t disp32_pitch, size_t disp32_stride,
                         float* d_conf, size_t conf_pitch, size_t conf_stride,
                         int count, char* err, int errlen)
{
    if (!matcher) return -1;
    try {
        static_cast<pyrsgm::BatchMatcher*>(matcher)->capture(
            {d_left, left_pitch, left_stride}, {d_right, right_pitch, right_stride},
            {d_disp16, disp16_pitch, disp16_stride}, {d_disp32, disp32_pitch, disp32_stride},
            {d_conf, conf_pitch, conf_stride}, count);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_batch_launch(void* matcher, hipStream_t stream, char* err, int errlen)
{
    if (!matcher) return -1;
    try {
        static_cast<pyrsgm::BatchMatcher*>(matcher)->launch(stream);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}


int pyrsgm_batch_set_tuning(void* matcher, const int tuning[6], char* err, int errlen)
{
    if (!matcher || !tuning) return -1;
    try {
        static_cast<pyrsgm::BatchMatcher*>(matcher)->setTuning(batchTuning(tuning));
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_batch_get_tuning(void* matcher, int tuning[6])
{
    if (!matcher || !tuning) return -1;
    batchTuningOut(static_cast<pyrsgm::BatchMatcher*>(matcher)->tuning(), tuning);
    return 0;
}

int pyrsgm_batch_tune(void* matcher,
                      const unsigned char* d_left, size_t left_pitch, size_t left_stride,
                      const unsigned char* d_right, size_t right_pitch, size_t right_stride,
                      short* d_disp16, size_t disp16_pitch, size_t disp16_stride,
                      float* d_disp32, size_t disp32_pitch, size_t disp32_stride,
                      float* d_conf, size_t conf_pitch, size_t conf_stride,
                      hipStream_t stream, int reps, int best[6], double* best_ms, char* err,
                      int errlen)
{
    if (!matcher) return -1;
    try {
        const pyrsgm::BatchTuning b = static_cast<pyrsgm::BatchMatcher*>(matcher)->tune(
            {d_left, left_pitch, left_stride}, {d_right, right_pitch, right_stride},
            {d_disp16, disp16_pitch, disp16_stride}, {d_disp32, disp32_pitch, disp32_stride},
            {d_conf, conf_pitch, conf_stride}, stream, reps, best_ms);
        if (best) batchTuningOut(b, best);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

size_t pyrsgm_batch_device_bytes(void* matcher)
{
    return matcher ? static_cast<pyrsgm::BatchMatcher*>(matcher)->deviceMemoryBytes() : 0;
}

size_t pyrsgm_batch_required_bytes(int width, int height, int batch,
                                   const pyrsgm_params_c* params, char* err, int errlen)
{
    if (!params) return 0;
    try {
        return pyrsgm::BatchMatcher::requiredBytes(width, height, batch, toParams(*params));
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return 0;
}

// ---- Server ----------------------------------------------------------------
void* pyrsgm_serve_create(size_t memory_budget, int max_batch, int lanes_per_key, int use_graphs,
                          char* err, int errlen)
{
    try {
        pyrsgm::ServeOptions o;
        o.memory_budget = memory_budget;
        o.max_batch = max_batch;
        o.lanes_per_key = lanes_per_key;
        o.use_graphs = use_graphs != 0;
        return new pyrsgm::Server(o);
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return nullptr;
}

void pyrsgm_serve_destroy(void* server) { delete static_cast<pyrsgm::Server*>(server); }

long long pyrsgm_serve_submit(void* server, const pyrsgm_serve_request_c* request,
                              const pyrsgm_params_c* params, long long arrival_ns, char* err,
                              int errlen)
{
    if (!server || !request || !params) return -1;
    try {
        pyrsgm::ServeRequest r;
        r.left = request->left;
        r.left_pitch = request->left_pitch;
        r.right = request->right;
        r.right_pitch = request->right_pitch;
        r.width = request->width;
        r.height = request->height;
        r.disp16 = request->disp16;
        r.disp16_pitch = request->disp16_pitch;
        r.disp32 = request->disp32;
        r.disp32_pitch = request->disp32_pitch;
        r.conf = request->conf;
        r.conf_pitch = request->conf_pitch;
        r.params = toParams(*params);
        return (long long)static_cast<pyrsgm::Server*>(server)->submit(r, arrival_ns);
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_serve_poll(void* server, unsigned long long ticket, pyrsgm_serve_result_c* out)
{
    if (!server) return -1;
    return static_cast<pyrsgm::Server*>(server)->poll(ticket, out) ? 1 : 0;
}

int pyrsgm_serve_wait(void* server, unsigned long long ticket, pyrsgm_serve_result_c* out,
                      char* err, int errlen)
{
    if (!server || !out) return -1;
    try {
        *out = static_cast<pyrsgm::Server*>(server)->wait(ticket);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_serve_drain(void* server)
{
    if (!server) return -1;
    static_cast<pyrsgm::Server*>(server)->drain();
    return 0;
}

int pyrsgm_serve_take(void* server, pyrsgm_serve_result_c* out, int max)
{
    if (!server || !out || max < 0) return -1;
    return (int)static_cast<pyrsgm::Server*>(server)->takeResults(out, (size_t)max);
}

int pyrsgm_serve_stats(void* server, pyrsgm_serve_stats_c* out)
{
    if (!server || !out) return -1;
    *out = static_cast<pyrsgm::Server*>(server)->stats();
    return 0;
}

void pyrsgm_serve_reset_stats(void* server)
{
    if (server) static_cast<pyrsgm::Server*>(server)->resetStats();
}

long long pyrsgm_serve_now_ns(void) { return (long long)pyrsgm::Server::nowNs(); }

// ---- MixedBatch -------------------------------------------------------------

void* pyrsgm_mixed_create(int n, const int* widths, const int* heights,
                          const pyrsgm_params_c* params, char* err, int errlen)
{
    if (n < 1 || !widths || !heights || !params) {
        copyErr("pyrsgm: MixedBatch needs n >= 1 sizes and parameter sets", err, errlen);
        return nullptr;
    }
    try {
        std::vector<pyrsgm::PairSpec> specs((size_t)n);
        for (int i = 0; i < n; ++i) {
            specs[(size_t)i].width = widths[i];
            specs[(size_t)i].height = heights[i];
            specs[(size_t)i].params = toParams(params[i]);
        }
        return new pyrsgm::MixedBatch(specs);
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    } catch (...) {
        copyErr("pyrsgm: unknown error", err, errlen);
    }
    return nullptr;
}

void pyrsgm_mixed_destroy(void* batch) { delete static_cast<pyrsgm::MixedBatch*>(batch); }

int pyrsgm_mixed_compute(void* batch, const pyrsgm_pair_io_c* io, hipStream_t stream, char* err,
                         int errlen)
{
    if (!batch) return -1;
    try {
        static_cast<pyrsgm::MixedBatch*>(batch)->compute(io, stream);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_mixed_capture(void* batch, const pyrsgm_pair_io_c* io, char* err, int errlen)
{
    if (!batch) return -1;
    try {
        static_cast<pyrsgm::MixedBatch*>(batch)->capture(io);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_mixed_launch(void* batch, hipStream_t stream, char* err, int errlen)
{
    if (!batch) return -1;
    try {
        static_cast<pyrsgm::MixedBatch*>(batch)->launch(stream);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_mixed_set_tuning(void* batch, const pyrsgm_mixed_tuning_c* t, char* err, int errlen)
{
    if (!batch || !t) return -1;
    try {
        pyrsgm::MixedTuning m;
        m.branches = t->branches;
        m.blocks_per_cu = t->blocks_per_cu;
        m.paired = t->paired;
        m.queued_stages = t->queued_stages;
        static_cast<pyrsgm::MixedBatch*>(batch)->setTuning(m);
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_mixed_get_tuning(void* batch, pyrsgm_mixed_tuning_c* out)
{
    if (!batch || !out) return -1;
    const pyrsgm::MixedTuning m = static_cast<pyrsgm::MixedBatch*>(batch)->tuning();
    out->branches = m.branches;
    out->blocks_per_cu = m.blocks_per_cu;
    out->paired = m.paired;
    out->queued_stages = m.queued_stages;
    return 0;
}

int pyrsgm_mixed_tune(void* batch, const pyrsgm_pair_io_c* io, hipStream_t stream, int reps,
                      pyrsgm_mixed_tuning_c* best, double* best_ms, char* err, int errlen)
{
    if (!batch) return -1;
    try {
        const pyrsgm::MixedTuning m = static_cast<pyrsgm::MixedBatch*>(batch)->tune(io, stream, reps, best_ms);
        if (best) {
            best->branches = m.branches;
            best->blocks_per_cu = m.blocks_per_cu;
            best->paired = m.paired;
            best->queued_stages = m.queued_stages;
        }
        return 0;
    } catch (const std::exception& e) {
        copyErr(e.what(), err, errlen);
    }
    return -1;
}

int pyrsgm_mixed_plan(void* batch, long long out[6])
{
    if (!batch || !out) return -1;
    const std::vector<long long> p = static_cast<pyrsgm::MixedBatch*>(batch)->plan();
    for (int i = 0; i < 6; ++i) out[i] = p[(size_t)i];
    return 0;
}

size_t pyrsgm_mixed_device_bytes(void* batch)
{
    return batch ? static_cast<pyrsgm::MixedBatch*>(batch)->deviceMemoryBytes() : 0;
}

} // extern "C"
