/*
 * pyrsgm_serve.h -- serving loop: stereo pairs of varying size arrive one at a
 * time and are matched on resident per-size lanes with continuous batching.
 *
 * No counterpart in gegeges/pyrsgm (whose multi-request form is one
 * StereoMatcher per stream, driven by the caller).  A Server owns a scheduler
 * thread and a set of lanes.  A lane is the device state for one key
 * (width, height, validated Params, requested outputs): a batched matcher for
 * up to `cap` pairs, stacked staging images, its own stream and one HIP graph
 * per batch size.  Lanes stay resident between requests, so a size seen before
 * costs no allocation.
 *
 * Scheduling: a request joins the next launch of an idle lane of its key,
 * together with every other pending request of that key (up to cap).  While
 * all lanes of a key are busy, its requests wait and form the next batch
 * (continuous batching); lanes of different keys run on their own streams and
 * overlap on the device.  A new lane is admitted only if its device footprint
 * fits the memory budget, after evicting idle lanes in least-recently-used
 * order; otherwise its requests wait until memory frees up.  The footprint is
 * the matcher's own allocation formula before a lane of that key exists and
 * the bytes the lane actually allocated after.
 *
 * Each result is bit-identical to matching the pair alone with StereoMatcher:
 * a launch of n pairs runs the same kernels on n independent workspaces.
 *
 * Requests carry device buffers owned by the caller; they must stay valid and
 * unchanged until the request completes (poll/wait/take report it).  Times are
 * steady-clock nanoseconds (Server::nowNs); a request's latency is done_ns -
 * arrival_ns, where arrival_ns defaults to the submit time.
 */
#ifndef PYRSGM_SERVE_H
#define PYRSGM_SERVE_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "pyrsgm/pyrsgm_py.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One stereo pair and its outputs (device buffers, pitches in bytes; null
 * outputs are not written, as in StereoMatcher::compute). */
typedef struct {
    const unsigned char* left;
    size_t left_pitch;
    const unsigned char* right;
    size_t right_pitch;
    int width, height;
    short* disp16;
    size_t disp16_pitch;
    float* disp32;
    size_t disp32_pitch;
    float* conf;
    size_t conf_pitch;
} pyrsgm_serve_request_c;

typedef struct {
    unsigned long long ticket;
    int status;      /* 0 done, -1 failed */
    int batch_size;  /* pairs in the launch that carried the request */
    long long arrival_ns, launch_ns, done_ns;
} pyrsgm_serve_result_c;

typedef struct {
    unsigned long long submitted, completed, failed;
    double pairs_per_s;  /* completed / (last done - first arrival), since the last reset */
    double p50_ms, p99_ms, max_ms, mean_ms, mean_batch;
    size_t resident_bytes, peak_bytes, budget_bytes;
    int lanes, lanes_created, lanes_evicted, launches;
} pyrsgm_serve_stats_c;

#ifdef __cplusplus
}

#include <memory>

#include "pyrsgm/pyrsgm.h"

namespace pyrsgm {

struct ServeOptions {
    size_t memory_budget = 0;  ///< device bytes for lanes; 0 = 80 % of the free memory at construction
    int max_batch = 16;        ///< pairs per launch of one lane
    int lanes_per_key = 2;     ///< lanes one key may hold (one runs while the next batch forms)
    bool use_graphs = true;    ///< replay each lane launch as a captured HIP graph
};

struct ServeRequest {
    const uint8_t* left = nullptr;
    size_t left_pitch = 0;
    const uint8_t* right = nullptr;
    size_t right_pitch = 0;
    int width = 0, height = 0;
    int16_t* disp16 = nullptr;
    size_t disp16_pitch = 0;
    float* disp32 = nullptr;
    size_t disp32_pitch = 0;
    float* conf = nullptr;
    size_t conf_pitch = 0;
    Params params;
};

using ServeResult = pyrsgm_serve_result_c;
using ServeStats = pyrsgm_serve_stats_c;

class Server {
public:
    explicit Server(const ServeOptions& options = ServeOptions());
    ~Server();  ///< finishes every submitted request, then frees all lanes
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    /// Queue a request; returns its ticket (>= 1).  Throws std::invalid_argument
    /// for bad sizes, null images, no outputs, invalid params or a pair whose
    /// one-pair lane exceeds the memory budget.  Thread-safe.
    uint64_t submit(const ServeRequest& request, int64_t arrival_ns = 0);
    /// True (and the result, which is then forgotten) once the ticket is done.
    bool poll(uint64_t ticket, ServeResult* result);
    /// Block until the ticket is done.  Throws std::invalid_argument for a ticket
    /// never issued ("unknown ticket"), and for one that another call (poll,
    /// takeResults, a second wait) retrieved, before or while this call waits.
    ServeResult wait(uint64_t ticket);
    /// Block until every submitted request is done.
    void drain();
    /// Move up to `max` finished results (not yet polled) into `out`.
    size_t takeResults(ServeResult* out, size_t max);
    ServeStats stats() const;
    void resetStats();  ///< latency samples and counters; resident lanes stay
    const ServeOptions& options() const;

    static int64_t nowNs();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

} // namespace pyrsgm

extern "C" {
#endif

/* Plain-C handles.  Errors: NULL / -1 with the message copied into err. */
void* pyrsgm_serve_create(size_t memory_budget, int max_batch, int lanes_per_key, int use_graphs,
                          char* err, int errlen);
void pyrsgm_serve_destroy(void* server);
long long pyrsgm_serve_submit(void* server, const pyrsgm_serve_request_c* request,
                              const pyrsgm_params_c* params, long long arrival_ns, char* err,
                              int errlen);
/* 1 done (result filled), 0 not yet, -1 bad arguments. */
int pyrsgm_serve_poll(void* server, unsigned long long ticket, pyrsgm_serve_result_c* out);
int pyrsgm_serve_wait(void* server, unsigned long long ticket, pyrsgm_serve_result_c* out,
                      char* err, int errlen);
int pyrsgm_serve_drain(void* server);
int pyrsgm_serve_take(void* server, pyrsgm_serve_result_c* out, int max);
int pyrsgm_serve_stats(void* server, pyrsgm_serve_stats_c* out);
void pyrsgm_serve_reset_stats(void* server);
long long pyrsgm_serve_now_ns(void);

#ifdef __cplusplus
}
#endif
#endif
