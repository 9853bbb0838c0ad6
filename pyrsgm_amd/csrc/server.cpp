// server.cpp -- pyrsgm::Server: resident per-key lanes, continuous batching,
// admission by device footprint (the contract is in pyrsgm_serve.h).
//
// Threads: callers submit / poll / wait under one mutex; one scheduler thread
// owns the lanes and is the only thread that issues device work.  It reaps
// finished launches (event queries), admits pending requests to lanes and
// launches every idle lane that has requests, releasing the mutex while it
// enqueues device work.  With nothing to do it sleeps on a doorbell that a
// submit, the destructor, or a lane's completion rings; a lane's completion
// rings it from a host function enqueued behind the lane's work.
#include "pyrsgm/pyrsgm_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "matcher_core.h"

namespace pyrsgm {
namespace {

constexpr int kOut16 = 1, kOut32 = 2, kOutConf = 4;

// Lane key: size, requested outputs and every field of the validated Params.
std::vector<int32_t> makeKey(int w, int h, int outputs, const Params& p)
{
    auto f = [](float v) {
        int32_t i;
        std::memcpy(&i, &v, sizeof i);
        return i;
    };
    return {w, h, outputs, p.num_disparities, p.min_disparity, p.block_size,
            (int32_t)p.cost_mode, p.p1, p.p2, p.uniqueness_ratio, p.disp12_max_diff,
            p.lr_check, p.median_filter, p.num_paths, p.pre_filter_cap, p.pyramid_levels,
            p.pyramid_coarse_disparities, p.pyramid_max_disparity, f(p.pyramid_min_confidence),
            p.adaptive_p2, f(p.p2_alpha), f(p.p2_beta), f(p.p2_lower_mult), f(p.p2_upper_mult)};
}

// Wake-up counter of the scheduler.  It has a mutex of its own, which nothing
// holds across a HIP call: the host functions that ring it run on the HIP
// runtime's callback thread and must never wait for the scheduler's mutex
// (a hipFree under that mutex waits for the streams, whose host functions
// would wait for the mutex).
struct Doorbell {
    std::mutex m;
    std::condition_variable cv;
    uint64_t rings = 0;

    void ring()
    {
        {
            std::lock_guard<std::mutex> lk(m);
            ++rings;
        }
        cv.notify_all();
    }
    uint64_t count()
    {
        std::lock_guard<std::mutex> lk(m);
        return rings;
    }
    // Returns once the counter has moved past `seen`.
    void waitPast(uint64_t seen)
    {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return rings != seen; });
    }
};

// Host function enqueued behind a lane's work: rings the doorbell and drops
// the reference it was given (the doorbell outlives every pending ring).
void ringFromStream(void* user)
{
    auto* ref = static_cast<std::shared_ptr<Doorbell>*>(user);
    (*ref)->ring();
    delete ref;
}

struct Req {
    uint64_t ticket = 0;
    ServeRequest r;
    std::vector<int32_t> key;
    int outputs = 0;
    int64_t arrival = 0, launch = 0, done = 0;
    int batch = 0;
    int status = 0;
};

// Staging bytes of a lane of n pairs (stacked images; pitches rounded to 256).
size_t stagingBytes(int w, int h, int n, int outputs)
{
    auto img = [&](size_t elem) { return ((size_t)w * elem + 255) / 256 * 256 * (size_t)h * n; };
    size_t b = 2 * img(1);
    if (outputs & kOut16) b += img(2);
    if (outputs & kOut32) b += img(4);
    if (outputs & kOutConf) b += img(4);
    return b + (size_t)5 * n * sizeof(pyr::CopyDesc);
}

struct Lane {
    std::vector<int32_t> key;
    int w = 0, h = 0, outputs = 0, cap = 1;
    std::unique_ptr<MatcherCore> core;
    DeviceImage<uint8_t> inL, inR;
    DeviceImage<int16_t> o16;
    DeviceImage<float> o32, oc;
    pyr::CopyDesc* dDesc = nullptr;  // device: [0, 2 cap) gathers, [2 cap, 5 cap) scatters
    pyr::CopyDesc* hDesc = nullptr;  // pinned host mirror
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<hipGraphExec_t> graphs;  // by pair count; null until first used
    std::vector<Req*> running;
    bool busy = false;
    int64_t lastUsed = 0;
    size_t bytes = 0;

    Lane(const std::vector<int32_t>& k, int width, int height, int outs, int n, const Params& p)
        : key(k), w(width), h(height), outputs(outs), cap(n)
    {
        try {
            core.reset(new MatcherCore(w, h, cap, p));
            inL.create(w, h * cap);
            inR.create(w, h * cap);
            if (outputs & kOut16) o16.create(w, h * cap);
            if (outputs & kOut32) o32.create(w, h * cap);
            if (outputs & kOutConf) oc.create(w, h * cap);
            const size_t db = (size_t)5 * cap * sizeof(pyr::CopyDesc);
            check(hipMalloc(reinterpret_cast<void**>(&dDesc), db), "hipMalloc");
            check(hipHostMalloc(reinterpret_cast<void**>(&hDesc), db, hipHostMallocDefault),
                  "hipHostMalloc");
            check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            check(hipEventCreateWithFlags(&done, hipEventDisableTiming), "hipEventCreateWithFlags");
            graphs.assign((size_t)cap + 1, nullptr);
            bytes = core->deviceBytes() + inL.bytes() + inR.bytes() + o16.bytes() + o32.bytes() +
                    oc.bytes() + db;
        } catch (...) {
            release();
            throw;
        }
    }
    ~Lane() { release(); }

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        graphs.clear();
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
        if (dDesc) (void)hipFree(dDesc);
        if (hDesc) (void)hipHostFree(hDesc);
        done = nullptr;
        stream = nullptr;
        dDesc = nullptr;
        hDesc = nullptr;
        core.reset();
    }

    int numOutputs() const
    {
        return ((outputs & kOut16) != 0) + ((outputs & kOut32) != 0) + ((outputs & kOutConf) != 0);
    }

    template <class T>
    static BatchBuffer<T> buf(DeviceImage<T>& im, int h)
    {
        return {im.data, im.pitch, im.pitch * (size_t)h};
    }

    // Gather, match and scatter `count` pairs whose descriptors are in dDesc.
    void enqueue(int count)
    {
        const int maxRow = w * (int)sizeof(float);
        check(pyr::launchCopy2D(dDesc, 2 * count, w, h, stream), "gather");
        core->compute({inL.data, inL.pitch, inL.pitch * (size_t)h},
                      {inR.data, inR.pitch, inR.pitch * (size_t)h}, buf(o16, h), buf(o32, h),
                      buf(oc, h), stream, count);
        check(pyr::launchCopy2D(dDesc + 2 * cap, numOutputs() * count, maxRow, h, stream),
              "scatter");
    }

    void launch(const std::vector<Req*>& batch, bool useGraphs,
                const std::shared_ptr<Doorbell>& bell)
    {
        const int n = (int)batch.size();
        const size_t sL = inL.pitch * (size_t)h, sR = inR.pitch * (size_t)h;
        for (int i = 0; i < n; ++i) {
            const ServeRequest& r = batch[i]->r;
            hDesc[2 * i] = {r.left, inL.data + i * sL, r.left_pitch, inL.pitch, w, h};
            hDesc[2 * i + 1] = {r.right, inR.data + i * sR, r.right_pitch, inR.pitch, w, h};
        }
        int k = 2 * cap;
        for (int i = 0; i < n; ++i) {
            const ServeRequest& r = batch[i]->r;
            if (outputs & kOut16)
                hDesc[k++] = {reinterpret_cast<const uint8_t*>(o16.data) + i * o16.pitch * (size_t)h,
                              reinterpret_cast<uint8_t*>(r.disp16), o16.pitch, r.disp16_pitch,
                              w * 2, h};
            if (outputs & kOut32)
                hDesc[k++] = {reinterpret_cast<const uint8_t*>(o32.data) + i * o32.pitch * (size_t)h,
                              reinterpret_cast<uint8_t*>(r.disp32), o32.pitch, r.disp32_pitch,
                              w * 4, h};
            if (outputs & kOutConf)
                hDesc[k++] = {reinterpret_cast<const uint8_t*>(oc.data) + i * oc.pitch * (size_t)h,
                              reinterpret_cast<uint8_t*>(r.conf), oc.pitch, r.conf_pitch, w * 4, h};
        }
        check(hipMemcpyAsync(dDesc, hDesc, (size_t)k * sizeof(pyr::CopyDesc), hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync");
        if (useGraphs) {
            if (!graphs[n]) {
                hipGraph_t g = nullptr;
                check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal),
                      "hipStreamBeginCapture");
                try {
                    enqueue(n);
                } catch (...) {
                    (void)hipStreamEndCapture(stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
                const hipError_t e = hipGraphInstantiate(&graphs[n], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                check(e, "hipGraphInstantiate");
            }
            check(hipGraphLaunch(graphs[n], stream), "hipGraphLaunch");
        } else {
            enqueue(n);
        }
        check(hipEventRecord(done, stream), "hipEventRecord");
        auto* ref = new std::shared_ptr<Doorbell>(bell);
        const hipError_t e = hipLaunchHostFunc(stream, ringFromStream, ref);
        if (e != hipSuccess) {
            delete ref;
            check(e, "hipLaunchHostFunc");
        }
    }
};

} // namespace

struct Server::Impl {
    ServeOptions opt;
    int device = 0;
    mutable std::mutex mu;
    std::condition_variable cvDone;
    std::shared_ptr<Doorbell> bell = std::make_shared<Doorbell>();
    std::deque<Req*> pending;
    std::unordered_map<uint64_t, Req*> inflight;   // submitted, not finished
    std::unordered_map<uint64_t, Req*> finished;   // finished, not yet retrieved
    std::vector<std::unique_ptr<Lane>> lanes;
    std::map<std::vector<int32_t>, size_t> measured;  // key + cap -> bytes a lane allocated
    uint64_t nextTicket = 1;
    bool stopping = false;
    size_t resident = 0, peak = 0;
    // statistics since the last reset
    uint64_t submitted = 0, completed = 0, failed = 0;
    int created = 0, evicted = 0, launches = 0;
    int64_t firstArrival = 0, lastDone = 0;
    double batchSum = 0;
    std::vector<double> latMs;
    std::thread worker;

    explicit Impl(const ServeOptions& o) : opt(o)
    {
        if (opt.max_batch < 1) throw std::invalid_argument("pyrsgm: max_batch must be >= 1");
        if (opt.lanes_per_key < 1) throw std::invalid_argument("pyrsgm: lanes_per_key must be >= 1");
        check(hipGetDevice(&device), "hipGetDevice");
        if (opt.memory_budget == 0) {
            size_t freeB = 0, totalB = 0;
            check(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
            opt.memory_budget = freeB / 10 * 8;
        }
        worker = std::thread([this] { loop(); });
    }

    ~Impl()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stopping = true;
        }
        bell->ring();
        worker.join();
        lanes.clear();
        for (auto& kv : finished) delete kv.second;
        for (auto& kv : inflight) delete kv.second;
    }

    static int outputsOf(const ServeRequest& r)
    {
        return (r.disp16 ? kOut16 : 0) | (r.disp32 ? kOut32 : 0) | (r.conf ? kOutConf : 0);
    }

    size_t predictBytes(const Req& q, int cap) const
    {
        std::vector<int32_t> k = q.key;
        k.push_back(cap);
        auto it = measured.find(k);
        if (it != measured.end()) return it->second;
        return MatcherCore::requiredBytes(q.r.width, q.r.height, cap, q.r.params) +
               stagingBytes(q.r.width, q.r.height, cap, q.outputs);
    }

    // Lane capacity for `demand` waiting requests of q's key: the next power of
    // two, at most max_batch, and halved until the lane takes at most half the budget.
    int capFor(const Req& q, int demand) const
    {
        int n = 1;
        while (n < demand && n < opt.max_batch) n *= 2;
        n = std::min(n, opt.max_batch);
        while (n > 1 && predictBytes(q, n) > opt.memory_budget / 2) n = (n + 1) / 2;
        return n;
    }

    uint64_t submit(const ServeRequest& r, int64_t arrival)
    {
        if (!r.left || !r.right) throw std::invalid_argument("pyrsgm: null input image");
        MatcherCore::checkSize(r.width, r.height, 1);
        if (!r.disp16 && !r.disp32 && !r.conf) throw std::invalid_argument("pyrsgm: request has no outputs");
        std::unique_ptr<Req> q(new Req);
        q->r = r;
        q->r.params = validateParams(r.params);
        q->outputs = outputsOf(r);
        q->key = makeKey(r.width, r.height, q->outputs, q->r.params);
        q->arrival = arrival ? arrival : Server::nowNs();
        std::lock_guard<std::mutex> lk(mu);
        if (predictBytes(*q, 1) > opt.memory_budget)
            throw std::invalid_argument("pyrsgm: request exceeds the server's memory budget");
        q->ticket = nextTicket++;
        if (submitted == 0 || q->arrival < firstArrival) firstArrival = q->arrival;
        ++submitted;
        inflight[q->ticket] = q.get();
        pending.push_back(q.get());
        const uint64_t t = q.release()->ticket;
        bell->ring();
        return t;
    }

    void finish(Req* q, int status, int64_t now)
    {
        q->status = status;
        q->done = now;
        inflight.erase(q->ticket);
        finished[q->ticket] = q;
        if (status == 0) {
            ++completed;
            latMs.push_back((double)(now - q->arrival) * 1e-6);
            lastDone = std::max(lastDone, now);
        } else {
            ++failed;
        }
    }

    // Finished launches (mutex held).
    bool reap()
    {
        bool any = false;
        for (auto& lp : lanes) {
            Lane& l = *lp;
            if (!l.busy) continue;
            const hipError_t e = hipEventQuery(l.done);
            if (e == hipErrorNotReady) continue;
            const int64_t now = Server::nowNs();
            for (Req* q : l.running) finish(q, e == hipSuccess ? 0 : -1, now);
            l.running.clear();
            l.busy = false;
            l.lastUsed = now;
            any = true;
        }
        if (any) cvDone.notify_all();
        return any;
    }

    void evict(size_t i)
    {
        resident -= lanes[i]->bytes;
        lanes.erase(lanes.begin() + (long)i);
        ++evicted;
    }

    // A new lane for q's key if its footprint fits after evicting idle lanes
    // (least recently used first); null if it does not fit now (mutex held).
    Lane* createLane(const Req& q, int demand, bool& allocFailed)
    {
        const int cap = capFor(q, demand);
        const size_t need = predictBytes(q, cap);
        while (resident + need > opt.memory_budget) {
            long victim = -1;
            for (size_t i = 0; i < lanes.size(); ++i)
                if (!lanes[i]->busy && (victim < 0 || lanes[i]->lastUsed < lanes[(size_t)victim]->lastUsed))
                    victim = (long)i;
            if (victim < 0) return nullptr;
            evict((size_t)victim);
        }
        std::unique_ptr<Lane> l;
        try {
            l.reset(new Lane(q.key, q.r.width, q.r.height, q.outputs, cap, q.r.params));
        } catch (const std::exception&) {
            (void)hipGetLastError();
            allocFailed = true;  // out of device memory now: retry once lanes free up
            return nullptr;
        }
        std::vector<int32_t> k = q.key;
        k.push_back(cap);
        measured[k] = l->bytes;
        resident += l->bytes;
        peak = std::max(peak, resident);
        ++created;
        lanes.push_back(std::move(l));
        return lanes.back().get();
    }

    // Pending requests to idle lanes; returns the batches to launch (mutex held).
    std::vector<std::pair<Lane*, std::vector<Req*>>> admit()
    {
        std::vector<std::pair<Lane*, std::vector<Req*>>> out;
        auto batchOf = [&](Lane* l) -> std::vector<Req*>* {
            for (auto& b : out)
                if (b.first == l) return &b.second;
            return nullptr;
        };
        std::map<std::vector<int32_t>, int> demand;  // waiting requests per key
        for (Req* q : pending) ++demand[q->key];
        for (auto it = pending.begin(); it != pending.end();) {
            Req* q = *it;
            Lane* target = nullptr;
            int sameKey = 0;
            for (auto& lp : lanes) {
                if (lp->key != q->key) continue;
                ++sameKey;
                if (std::vector<Req*>* b = batchOf(lp.get())) {
                    if ((int)b->size() < lp->cap) {  // a batch already forming
                        target = lp.get();
                        break;
                    }
                    continue;
                }
                if (!lp->busy && !target) target = lp.get();
            }
            const int want = demand[q->key];
            bool allocFailed = false;
            if (target && !batchOf(target) && 2 * target->cap < std::min(want, opt.max_batch)) {
                // An idle lane much smaller than the demand: replace it by a larger one.
                for (size_t i = 0; i < lanes.size(); ++i)
                    if (lanes[i].get() == target) {
                        evict(i);
                        break;
                    }
                --sameKey;
                target = createLane(*q, want, allocFailed);
                if (!target) target = createLane(*q, 1, allocFailed);
            }
            if (!target && sameKey < opt.lanes_per_key) target = createLane(*q, want, allocFailed);
            if (!target) {
                bool anyBusy = false;
                for (auto& lp : lanes) anyBusy |= lp->busy;
                if (allocFailed && !anyBusy) {
                    // Nothing running will free memory: this request cannot be served.
                    finish(q, -1, Server::nowNs());
                    cvDone.notify_all();
                    it = pending.erase(it);
                    continue;
                }
                ++it;
                continue;
            }
            std::vector<Req*>* b = batchOf(target);
            if (!b) {
                out.push_back({target, {}});
                b = &out.back().second;
            }
            b->push_back(q);
            target->busy = true;  // reserved for this launch
            --demand[q->key];
            it = pending.erase(it);
        }
        return out;
    }

    void loop()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            // Rings after this snapshot (a submit, a finished lane) end the
            // sleep below, so none is lost between the checks and the wait.
            const uint64_t seen = bell->count();
            bool progress = reap();
            auto batches = admit();
            if (!batches.empty()) {
                progress = true;
                const int64_t now = Server::nowNs();
                for (auto& b : batches) {
                    for (Req* q : b.second) {
                        q->launch = now;
                        q->batch = (int)b.second.size();
                    }
                    b.first->running = b.second;
                    ++launches;
                    batchSum += (double)b.second.size();
                }
                lk.unlock();
                std::vector<Lane*> bad;
                for (auto& b : batches) {
                    try {
                        b.first->launch(b.second, opt.use_graphs, bell);
                    } catch (const std::exception&) {
                        bad.push_back(b.first);
                    }
                }
                lk.lock();
                for (Lane* l : bad) {
                    const int64_t t = Server::nowNs();
                    for (Req* q : l->running) finish(q, -1, t);
                    l->running.clear();
                    l->busy = false;
                }
                if (!bad.empty()) cvDone.notify_all();
            }
            bool anyBusy = false;
            for (auto& lp : lanes) anyBusy |= lp->busy;
            if (stopping && pending.empty() && !anyBusy) break;
            if (!progress) {
                // Busy lanes ring on completion; submits and stop ring too.
                lk.unlock();
                bell->waitPast(seen);
                lk.lock();
            }
        }
    }

    static void fill(ServeResult* out, const Req& q)
    {
        out->ticket = q.ticket;
        out->status = q.status;
        out->batch_size = q.batch;
        out->arrival_ns = q.arrival;
        out->launch_ns = q.launch;
        out->done_ns = q.done;
    }
};

Server::Server(const ServeOptions& options) : impl_(new Impl(options)) {}
Server::~Server() = default;

int64_t Server::nowNs()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

uint64_t Server::submit(const ServeRequest& request, int64_t arrival_ns)
{
    return impl_->submit(request, arrival_ns);
}

bool Server::poll(uint64_t ticket, ServeResult* result)
{
    std::lock_guard<std::mutex> lk(impl_->mu);
    auto it = impl_->finished.find(ticket);
    if (it == impl_->finished.end()) return false;
    if (result) Impl::fill(result, *it->second);
    delete it->second;
    impl_->finished.erase(it);
    impl_->cvDone.notify_all();  // a wait() on this ticket ends
    return true;
}

ServeResult Server::wait(uint64_t ticket)
{
    std::unique_lock<std::mutex> lk(impl_->mu);
    if (!impl_->finished.count(ticket) && !impl_->inflight.count(ticket)) {
        // Tickets are issued in increasing order from 1: an issued ticket that
        // is neither in flight nor finished was retrieved already.
        if (ticket != 0 && ticket < impl_->nextTicket)
            throw std::invalid_argument("pyrsgm: ticket was retrieved by another call");
        throw std::invalid_argument("pyrsgm: unknown ticket");
    }
    // A ticket leaves `inflight` only into `finished`, and leaves `finished`
    // when some caller (poll, takeResults, another wait) retrieves it: the
    // wait ends in either case.
    impl_->cvDone.wait(lk, [&] {
        return impl_->finished.count(ticket) != 0 || impl_->inflight.count(ticket) == 0;
    });
    auto it = impl_->finished.find(ticket);
    if (it == impl_->finished.end())
        throw std::invalid_argument("pyrsgm: ticket was retrieved by another call");
    ServeResult r;
    Impl::fill(&r, *it->second);
    delete it->second;
    impl_->finished.erase(it);
    impl_->cvDone.notify_all();
    return r;
}

void Server::drain()
{
    std::unique_lock<std::mutex> lk(impl_->mu);
    impl_->cvDone.wait(lk, [&] { return impl_->inflight.empty(); });
}

size_t Server::takeResults(ServeResult* out, size_t max)
{
    std::lock_guard<std::mutex> lk(impl_->mu);
    size_t n = 0;
    for (auto it = impl_->finished.begin(); it != impl_->finished.end() && n < max;) {
        Impl::fill(&out[n++], *it->second);
        delete it->second;
        it = impl_->finished.erase(it);
    }
    if (n) impl_->cvDone.notify_all();
    return n;
}

ServeStats Server::stats() const
{
    std::lock_guard<std::mutex> lk(impl_->mu);
    const Impl& m = *impl_;
    ServeStats s{};
    s.submitted = m.submitted;
    s.completed = m.completed;
    s.failed = m.failed;
    std::vector<double> v = m.latMs;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
        if (v.empty()) return 0.0;
        const size_t i = (size_t)std::min<double>((double)v.size() - 1, std::ceil(p * (double)v.size()) - 1);
        return v[i];
    };
    s.p50_ms = pct(0.50);
    s.p99_ms = pct(0.99);
    s.max_ms = v.empty() ? 0.0 : v.back();
    double sum = 0;
    for (double x : v) sum += x;
    s.mean_ms = v.empty() ? 0.0 : sum / (double)v.size();
    s.mean_batch = m.launches ? m.batchSum / m.launches : 0.0;
    const double span = (double)(m.lastDone - m.firstArrival) * 1e-9;
    s.pairs_per_s = (m.completed && span > 0) ? (double)m.completed / span : 0.0;
    s.resident_bytes = m.resident;
    s.peak_bytes = m.peak;
    s.budget_bytes = m.opt.memory_budget;
    s.lanes = (int)m.lanes.size();
    s.lanes_created = m.created;
    s.lanes_evicted = m.evicted;
    s.launches = m.launches;
    return s;
}

void Server::resetStats()
{
    std::lock_guard<std::mutex> lk(impl_->mu);
    Impl& m = *impl_;
    m.submitted = m.completed = m.failed = 0;
    m.created = m.evicted = m.launches = 0;
    m.firstArrival = m.lastDone = 0;
    m.batchSum = 0;
    m.latMs.clear();
    m.peak = m.resident;
}

const ServeOptions& Server::options() const { return impl_->opt; }

} // namespace pyrsgm
