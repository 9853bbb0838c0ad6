// mixed.cpp -- pyrsgm::MixedBatch: pairs of different sizes and parameters in
// one fused pass (queued path and winner-take-all launches over all pairs,
// per-pair stages on parallel streams, whole pass capturable as one HIP
// graph), and its tuner.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "matcher_core.h"
#include "pyrsgm/pyrsgm_mixed.h"

namespace pyrsgm {

namespace {

constexpr int kMaxBranches = 4;

// Tunings found by tune(), per pass signature, for the life of the process.
std::mutex g_tuneMutex;
std::map<std::string, MixedTuning> g_tuned;

void checkTuning(const MixedTuning& t)
{
    if (t.branches < 1 || t.branches > kMaxBranches)
        throw std::invalid_argument("pyrsgm: MixedTuning.branches must be in [1, 4]");
    if (t.blocks_per_cu < 0 || t.blocks_per_cu > 64)
        throw std::invalid_argument("pyrsgm: MixedTuning.blocks_per_cu must be in [0, 64]");
    if (t.paired < -1 || t.paired > 1)
        throw std::invalid_argument("pyrsgm: MixedTuning.paired must be -1, 0 or 1");
}

} // namespace

struct MixedBatch::Impl {
    // One queued path launch: the tasks of one step that share a kernel.
    struct Group {
        int step = 0;
        int mode = 0, Dp = 0, adaptive = 0, exact = 0, narrow = 0, paired = 0;
        int taskOff = 0, nTasks = 0, total = 0;
        int wtaTotal = 0;  // blocks of the group's winner-take-all launch
        int grid = 0;
        std::vector<int> patchTask, patchPair;  // single-scale adaptive tasks (per-call image)
    };
    struct Recorded {
        int pair, level;
        pyr::PathTask t;
    };

    std::vector<PairSpec> specs;
    std::vector<std::unique_ptr<MatcherCore>> cores;
    int maxSteps = 0;
    MixedTuning tun;
    std::vector<Recorded> recorded;       // every pair level's path task, as recorded
    std::vector<Group> groups;
    std::vector<pyr::PathArgs> hostTasks;  // device table image
    std::vector<int> hostStart;            // per group nTasks + 1 entries, concatenated
    std::vector<size_t> startOff;          // group -> offset into hostStart
    std::vector<int> hostWtaStart;         // per group nTasks + 1 WTA block offsets, concatenated
    std::vector<size_t> wtaStartOff;       // group -> offset into hostWtaStart
    void* dev = nullptr;
    size_t devBytes = 0;
    pyr::PathArgs* dTasks = nullptr;
    int* dStart = nullptr;
    int* dWtaStart = nullptr;
    uint32_t* dCounters = nullptr;
    int numCUs = 256;
    std::vector<int> branchOf;
    hipStream_t branch[kMaxBranches] = {};
    hipEvent_t evBranch[kMaxBranches] = {};
    hipEvent_t evMain = nullptr;
    hipStream_t captureStream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long perPairLaunches = 0;

    explicit Impl(const std::vector<PairSpec>& pairs, const MixedTuning& t) : specs(pairs), tun(t)
    {
        if (pairs.empty()) throw std::invalid_argument("pyrsgm: MixedBatch needs at least one pair");
        checkTuning(t);
        try {
            for (const PairSpec& p : pairs) {
                cores.emplace_back(new MatcherCore(p.width, p.height, 1, p.params));
                maxSteps = std::max(maxSteps, cores.back()->levels + 1);
            }
            int dev = 0;
            check(hipGetDevice(&dev), "hipGetDevice");
            check(hipDeviceGetAttribute(&numCUs, hipDeviceAttributeMultiprocessorCount, dev),
                  "hipDeviceGetAttribute");
            for (int b = 0; b < kMaxBranches; ++b) {
                check(hipStreamCreateWithFlags(&branch[b], hipStreamNonBlocking), "hipStreamCreateWithFlags");
                check(hipEventCreateWithFlags(&evBranch[b], hipEventDisableTiming), "hipEventCreateWithFlags");
            }
            check(hipEventCreateWithFlags(&evMain, hipEventDisableTiming), "hipEventCreateWithFlags");
            check(hipStreamCreateWithFlags(&captureStream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            record();
            plan();
        } catch (...) {
            destroy();
            throw;
        }
    }

    ~Impl() { destroy(); }

    void destroy()
    {
        for (int b = 0; b < kMaxBranches; ++b)
            if (branch[b]) (void)hipStreamSynchronize(branch[b]);
        dropGraph();
        if (dev) (void)hipFree(dev);
        dev = nullptr;
        for (int b = 0; b < kMaxBranches; ++b) {
            if (evBranch[b]) (void)hipEventDestroy(evBranch[b]);
            if (branch[b]) (void)hipStreamDestroy(branch[b]);
            evBranch[b] = nullptr;
            branch[b] = nullptr;
        }
        if (evMain) (void)hipEventDestroy(evMain);
        if (captureStream) (void)hipStreamDestroy(captureStream);
        evMain = nullptr;
        captureStream = nullptr;
        cores.clear();
    }

    void dropGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    int levelAt(int i, int step) const { return cores[i]->levels - step; }

    // The path launch of every pair level, recorded (not issued) through the
    // same level calls compute() makes.  The input images are placeholders:
    // a recorded task reads only the core's own buffers, except single-scale
    // tasks with adaptive P2, whose left image is patched in per call.
    void record()
    {
        static const uint8_t placeholder[16] = {};
        const BatchBuffer<const uint8_t> img{placeholder, 16, 0};
        for (int i = 0; i < (int)cores.size(); ++i) {
            MatcherCore& c = *cores[i];
            const BatchBuffer<int16_t> d16{reinterpret_cast<int16_t*>(const_cast<uint8_t*>(placeholder)), 16, 0};
            for (int level = c.levels; level >= 0; --level) {
                pyr::PathSink sink;
                if (c.levels == 0)
                    c.singleCall(img, img, d16, {}, {}, nullptr, pyr::kStagePaths, 1, &sink);
                else
                    c.levelCall(level, img, img, d16, {}, {}, nullptr, pyr::kStagePaths, 1, &sink);
                if (sink.tasks.size() != 1) throw std::logic_error("pyrsgm: path task not recorded");
                recorded.push_back({i, level, sink.tasks[0]});
            }
        }
    }

    // Groups, task tables and queue counters for the current tuning.
    void plan()
    {
        groups.clear();
        hostTasks.clear();
        hostStart.clear();
        startOff.clear();
        hostWtaStart.clear();
        wtaStartOff.clear();
        perPairLaunches = 0;
        for (int step = 0; step < maxSteps; ++step) {
            // Tasks of this step, grouped by kernel specialisation.
            std::map<std::vector<int>, std::vector<int>> byKey;  // key -> recorded indices
            for (int r = 0; r < (int)recorded.size(); ++r) {
                const Recorded& rec = recorded[r];
                if (levelAt(rec.pair, step) != rec.level) continue;
                const pyr::PathArgs& a = rec.t.a;
                const int exact = rec.t.mode == pyr::kSingle ? 0 : a.exact;
                const int narrow = exact ? 0 : a.narrow;
                byKey[{rec.t.mode, a.Dp, rec.t.adaptive, exact, narrow}].push_back(r);
            }
            for (auto& kv : byKey) {
                Group g;
                g.step = step;
                g.mode = kv.first[0];
                g.Dp = kv.first[1];
                g.adaptive = kv.first[2];
                g.exact = kv.first[3];
                g.narrow = kv.first[4];
                std::vector<int>& rs = kv.second;
                // Longest lines first.
                std::stable_sort(rs.begin(), rs.end(), [&](int x, int y) {
                    const pyr::PathArgs& a = recorded[x].t.a;
                    const pyr::PathArgs& b = recorded[y].t.a;
                    return std::max(a.W, a.H) > std::max(b.W, b.H);
                });
                const bool pairable = g.narrow && g.mode != pyr::kSingle && g.Dp == 16;
                if (pairable) {
                    if (tun.paired >= 0) {
                        g.paired = tun.paired;
                    } else {
                        size_t blocks4 = 0;
                        for (int r : rs) {
                            const pyr::PathArgs& a = recorded[r].t.a;
                            blocks4 += (size_t)((a.numPaths >= 2 ? 2 : 1) * ((a.H + 3) / 4) +
                                                (a.numPaths >= 3 ? (a.W + 3) / 4 : 0));
                        }
                        g.paired = pyr::pairedLinesAuto(blocks4);
                    }
                }
                const int lpb = pyr::queuedLinesPerBlock(g.Dp, g.paired);
                g.taskOff = (int)hostTasks.size();
                startOff.push_back(hostStart.size());
                wtaStartOff.push_back(hostWtaStart.size());
                int total = 0, wtaTotal = 0;
                for (int r : rs) {
                    pyr::PathArgs a = recorded[r].t.a;
                    a.nLRBlocks = a.numPaths >= 2 ? (a.H + lpb - 1) / lpb : 0;
                    a.nRLBlocks = (a.H + lpb - 1) / lpb;
                    const int nTB = a.numPaths >= 3 ? (a.W + lpb - 1) / lpb : 0;
                    if (g.mode == pyr::kSingle && g.adaptive) {
                        g.patchTask.push_back((int)hostTasks.size());
                        g.patchPair.push_back(recorded[r].pair);
                    }
                    hostStart.push_back(total);
                    total += a.nLRBlocks + a.nRLBlocks + nTB;
                    hostWtaStart.push_back(wtaTotal);
                    wtaTotal += pyr::wtaBlocks(a);
                    hostTasks.push_back(a);
                }
                hostStart.push_back(total);
                hostWtaStart.push_back(wtaTotal);
                g.nTasks = (int)rs.size();
                g.total = total;
                g.wtaTotal = wtaTotal;
                const int occ = pyr::queuedBlocksPerCU(g.mode, g.Dp, g.adaptive, g.exact, g.narrow, g.paired);
                const int bpc = tun.blocks_per_cu > 0 ? tun.blocks_per_cu : occ;
                g.grid = (int)std::min<long long>((long long)total, (long long)numCUs * bpc);
                groups.push_back(std::move(g));
            }
            for (int i = 0; i < (int)cores.size(); ++i)
                if (levelAt(i, step) >= 0) perPairLaunches += 3;  // prep, cost, post
        }
        for (auto& c : cores) perPairLaunches += c->levels;  // pyrDown
        // Device tables: tasks, starts, two counters per group.
        const size_t tasksBytes = (hostTasks.size() * sizeof(pyr::PathArgs) + 255) / 256 * 256;
        const size_t startBytes = (hostStart.size() * sizeof(int) + 255) / 256 * 256;
        const size_t wtaBytes = (hostWtaStart.size() * sizeof(int) + 255) / 256 * 256;
        const size_t counterBytes = groups.size() * 2 * sizeof(uint32_t);
        const size_t need = tasksBytes + startBytes + wtaBytes + counterBytes;
        if (need > devBytes) {
            if (dev) check(hipFree(dev), "hipFree");
            dev = nullptr;
            devBytes = 0;
            check(hipMalloc(&dev, need), "hipMalloc");
            devBytes = need;
        }
        char* base = static_cast<char*>(dev);
        dTasks = reinterpret_cast<pyr::PathArgs*>(base);
        dStart = reinterpret_cast<int*>(base + tasksBytes);
        dWtaStart = reinterpret_cast<int*>(base + tasksBytes + startBytes);
        dCounters = reinterpret_cast<uint32_t*>(base + tasksBytes + startBytes + wtaBytes);
        check(hipMemcpy(dTasks, hostTasks.data(), hostTasks.size() * sizeof(pyr::PathArgs),
                        hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dStart, hostStart.data(), hostStart.size() * sizeof(int), hipMemcpyHostToDevice),
              "hipMemcpy");
        check(hipMemcpy(dWtaStart, hostWtaStart.data(), hostWtaStart.size() * sizeof(int),
                        hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemset(dCounters, 0, counterBytes), "hipMemset");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        balanceBranches();
    }

    // Pairs to branches: largest estimated work first, each to the least loaded.
    void balanceBranches()
    {
        const int nb = tun.branches;
        std::vector<int> order(cores.size());
        std::vector<double> work(cores.size());
        for (size_t i = 0; i < cores.size(); ++i) {
            order[i] = (int)i;
            double w = 0;
            for (int level = 0; level <= cores[i]->levels; ++level) {
                const LevelInfo& li = cores[i]->info[level];
                w += (double)li.width * li.height * (16.0 + li.num_labels);
            }
            work[i] = w;
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
        std::vector<double> load(nb, 0.0);
        branchOf.assign(cores.size(), 0);
        for (int i : order) {
            const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            branchOf[i] = b;
            load[b] += work[i];
        }
    }

    void checkIO(const pyrsgm_pair_io_c* io) const
    {
        if (!io) throw std::invalid_argument("pyrsgm: null pair list");
        for (size_t i = 0; i < cores.size(); ++i) {
            if (!io[i].left || !io[i].right) throw std::invalid_argument("pyrsgm: null input image");
            if (!io[i].disp16 && !io[i].disp32 && !io[i].conf)
                throw std::invalid_argument("pyrsgm: a pair of a MixedBatch has no output");
        }
    }

    // The per-pair stages of pair i at `level` on stream s.
    void pairStages(int i, int level, const pyrsgm_pair_io_c& io, hipStream_t s, int stages)
    {
        MatcherCore& c = *cores[i];
        const BatchBuffer<const uint8_t> L{io.left, io.left_pitch, 0}, R{io.right, io.right_pitch, 0};
        const BatchBuffer<int16_t> d16{io.disp16, io.disp16_pitch, 0};
        const BatchBuffer<float> d32{io.disp32, io.disp32_pitch, 0};
        const BatchBuffer<float> oc = c.confRoute(d32, {io.conf, io.conf_pitch, 0});
        if (c.levels == 0)
            c.singleCall(L, R, d16, d32, oc, s, stages, 1);
        else
            c.levelCall(level, L, R, d16, d32, oc, s, stages, 1);
    }

    void fork(hipStream_t main)
    {
        if (tun.branches < 2) return;
        check(hipEventRecord(evMain, main), "hipEventRecord");
        for (int b = 0; b < tun.branches; ++b) check(hipStreamWaitEvent(branch[b], evMain, 0), "hipStreamWaitEvent");
    }

    void join(hipStream_t main)
    {
        if (tun.branches < 2) return;
        for (int b = 0; b < tun.branches; ++b) {
            check(hipEventRecord(evBranch[b], branch[b]), "hipEventRecord");
            check(hipStreamWaitEvent(main, evBranch[b], 0), "hipStreamWaitEvent");
        }
    }

    void compute(const pyrsgm_pair_io_c* io, hipStream_t main)
    {
        checkIO(io);
        const int n = (int)cores.size();
        auto br = [&](int i) { return tun.branches > 1 ? branch[branchOf[i]] : main; };
        fork(main);
        for (int i = 0; i < n; ++i)
            if (cores[i]->levels > 0)
                cores[i]->pyrDownChain({io[i].left, io[i].left_pitch, 0}, {io[i].right, io[i].right_pitch, 0},
                                       br(i), 1);
        size_t gi = 0;
        for (int step = 0; step < maxSteps; ++step) {
            for (int i = 0; i < n; ++i)
                if (levelAt(i, step) >= 0)
                    pairStages(i, levelAt(i, step), io[i], br(i), pyr::kStagePrep | pyr::kStageCost);
            join(main);
            for (; gi < groups.size() && groups[gi].step == step; ++gi) {
                const Group& g = groups[gi];
                for (size_t p0 = 0; p0 < g.patchTask.size(); p0 += pyr::kMaxImgPatches) {
                    pyr::ImgPatch patch;
                    memset(&patch, 0, sizeof(patch));
                    patch.tasks = dTasks;
                    patch.n = (int)std::min<size_t>(pyr::kMaxImgPatches, g.patchTask.size() - p0);
                    for (int k = 0; k < patch.n; ++k) {
                        const int pr = g.patchPair[p0 + k];
                        patch.idx[k] = g.patchTask[p0 + k];
                        patch.img[k] = io[pr].left;
                        patch.istep[k] = (int)io[pr].left_pitch;
                    }
                    check(pyr::launchImgPatch(patch, main), "image patch");
                }
                pyr::QueueArgs q;
                q.tasks = dTasks + g.taskOff;
                q.start = dStart + startOff[gi];
                q.counter = dCounters + 2 * gi;
                q.nTasks = g.nTasks;
                q.total = g.total;
                pyr::profBegin(PYRSGM_STAGE_PATHS, 0, 0, g.Dp, main);
                const hipError_t e = pyr::launchPathsQueued(q, g.mode, g.Dp, g.adaptive, g.exact,
                                                            g.narrow, g.paired, g.grid, main);
                pyr::profEnd(main);
                if (e != hipSuccess) {
                    std::ostringstream oss;
                    oss << "queued paths (step " << step << ", labels " << g.Dp << ", " << g.nTasks << " tasks)";
                    throwHip(oss.str(), e);
                }
                pyr::WtaQueueArgs wq;
                wq.tasks = dTasks + g.taskOff;
                wq.start = dWtaStart + wtaStartOff[gi];
                wq.nTasks = g.nTasks;
                wq.total = g.wtaTotal;
                pyr::profBegin(PYRSGM_STAGE_WTA, 0, 0, g.Dp, main);
                const hipError_t ew = pyr::launchWTAQueued(wq, g.mode, g.Dp, g.narrow, main);
                pyr::profEnd(main);
                if (ew != hipSuccess) throwHip("queued winner-take-all", ew);
            }
            fork(main);
            for (int i = 0; i < n; ++i)
                if (levelAt(i, step) >= 0) pairStages(i, levelAt(i, step), io[i], br(i), pyr::kStagePost);
        }
        join(main);
    }

    void capture(const pyrsgm_pair_io_c* io)
    {
        checkIO(io);
        dropGraph();
        check(hipStreamBeginCapture(captureStream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            compute(io, captureStream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(captureStream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(captureStream, &graph), "hipStreamEndCapture");
        check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }

    std::string signature() const
    {
        std::ostringstream oss;
        for (size_t i = 0; i < cores.size(); ++i) {
            const MatcherCore& c = *cores[i];
            const Params& p = c.params;
            oss << c.width << 'x' << c.height << ':' << p.num_disparities << ':' << p.pyramid_levels << ':'
                << (int)p.cost_mode << ':' << p.adaptive_p2 << ':' << p.p2 << ':' << p.num_paths << ';';
        }
        return oss.str();
    }

    void setTuning(const MixedTuning& t)
    {
        checkTuning(t);
        for (int b = 0; b < kMaxBranches; ++b) check(hipStreamSynchronize(branch[b]), "hipStreamSynchronize");
        dropGraph();
        tun = t;
        plan();
    }

    double timePass(const pyrsgm_pair_io_c* io, hipStream_t s, int reps)
    {
        hipEvent_t a = nullptr, b = nullptr;
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventCreate(&b), "hipEventCreate");
        compute(io, s);  // warm-up
        check(hipEventRecord(a, s), "hipEventRecord");
        for (int r = 0; r < reps; ++r) compute(io, s);
        check(hipEventRecord(b, s), "hipEventRecord");
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return (double)ms / reps;
    }

    // Coordinate descent over the three knobs, from the current tuning.
    MixedTuning tune(const pyrsgm_pair_io_c* io, hipStream_t s, int reps, double* bestMs)
    {
        checkIO(io);
        if (reps < 1) reps = 1;
        MixedTuning best = tun;
        double bestT = timePass(io, s, reps);
        auto consider = [&](MixedTuning t) {
            setTuning(t);
            const double ms = timePass(io, s, reps);
            if (ms < bestT) {
                bestT = ms;
                best = t;
            }
        };
        for (int b : {1, 2, 4}) {
            MixedTuning t = best;
            if (t.branches == b) continue;
            t.branches = b;
            consider(t);
        }
        for (int bpc : {0, 1, 2, 4, 8}) {
            MixedTuning t = best;
            if (t.blocks_per_cu == bpc) continue;
            t.blocks_per_cu = bpc;
            consider(t);
        }
        for (int pk : {-1, 0, 1}) {
            MixedTuning t = best;
            if (t.paired == pk) continue;
            t.paired = pk;
            consider(t);
        }
        setTuning(best);
        {
            std::lock_guard<std::mutex> lock(g_tuneMutex);
            g_tuned[signature()] = best;
        }
        if (bestMs) *bestMs = bestT;
        return best;
    }
};

MixedBatch::MixedBatch(const std::vector<PairSpec>& pairs, const MixedTuning& tuning)
{
    MixedTuning t = tuning;
    impl_.reset(new Impl(pairs, t));
    // A tuning found earlier in this process for the same pass replaces the default.
    std::lock_guard<std::mutex> lock(g_tuneMutex);
    auto it = g_tuned.find(impl_->signature());
    if (it != g_tuned.end() && tuning.branches == MixedTuning{}.branches &&
        tuning.blocks_per_cu == MixedTuning{}.blocks_per_cu && tuning.paired == MixedTuning{}.paired) {
        impl_->tun = it->second;
        impl_->plan();
    }
}

MixedBatch::~MixedBatch() = default;

void MixedBatch::compute(const pyrsgm_pair_io_c* io, hipStream_t stream) { impl_->compute(io, stream); }
void MixedBatch::capture(const pyrsgm_pair_io_c* io) { impl_->capture(io); }
void MixedBatch::launch(hipStream_t stream)
{
    if (!impl_->exec) throw std::logic_error("pyrsgm: MixedBatch::launch before capture");
    check(hipGraphLaunch(impl_->exec, stream), "hipGraphLaunch");
}
void MixedBatch::setTuning(const MixedTuning& t) { impl_->setTuning(t); }
MixedTuning MixedBatch::tuning() const { return impl_->tun; }
MixedTuning MixedBatch::tune(const pyrsgm_pair_io_c* io, hipStream_t stream, int reps, double* best_ms)
{
    return impl_->tune(io, stream, reps, best_ms);
}
int MixedBatch::pairs() const { return (int)impl_->cores.size(); }
const Params& MixedBatch::params(int i) const { return impl_->cores.at((size_t)i)->params; }
size_t MixedBatch::deviceMemoryBytes() const
{
    size_t total = impl_->devBytes;
    for (const auto& c : impl_->cores) total += c->deviceBytes();
    return total;
}
std::vector<long long> MixedBatch::plan() const
{
    long long tasks = 0, items = 0;
    for (const auto& g : impl_->groups) {
        tasks += g.nTasks;
        items += g.total;
    }
    return {(long long)impl_->groups.size(), tasks, items, impl_->perPairLaunches};
}

} // namespace pyrsgm
