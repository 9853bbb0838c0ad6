// mixed.cpp -- pyrsgm::MixedBatch: pairs of different sizes and parameters in
// one fused pass (queued path and winner-take-all launches over all pairs,
// per-pair stages on parallel streams, whole pass capturable as one HIP
// graph), and its tuner.
#include <stdint.h>
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
    if (t.queued_stages != 0 && t.queued_stages != 1)
        throw std::invalid_argument("pyrsgm: MixedTuning.queued_stages must be 0 or 1");
}

// Phases of a pyramid step's stage launches, in issue order: the cost pass
// that follows another cost pass of the same pair (the BT vertical window sums
// after the row sums) runs in a phase of its own.
enum StagePhase : int { kPhPyrDown = 0, kPhPrep, kPhCost, kPhCost2, kPhPost };
int phaseOf(int kind)
{
    switch (kind) {
    case pyr::kStPyrDown: return kPhPyrDown;
    case pyr::kStPrepCensus:
    case pyr::kStPrepBT: return kPhPrep;
    case pyr::kStBoxV: return kPhCost2;
    case pyr::kStPost: return kPhPost;
    default: return kPhCost;
    }
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
    // A recorded stage launch of one pair: pyramid level (or pyrDown depth,
    // with level -1), its launch and its I/O flags.
    struct RecordedStage {
        int pair, level, depth;
        pyr::StageLaunch l;
    };
    // One queued stage launch: the tasks of one step and phase that share a
    // stage kernel (pyrDown: one per depth, step = depth - 1).
    struct StageGroup {
        int step = 0, phase = 0, kind = 0, variant = 0;
        int taskOff = 0, nTasks = 0, blocks = 0;
        size_t lds = 0, startOff = 0;
    };

    std::vector<PairSpec> specs;
    std::vector<std::unique_ptr<MatcherCore>> cores;
    int maxSteps = 0;
    MixedTuning tun;
    std::vector<Recorded> recorded;       // every pair level's path task, as recorded
    std::vector<RecordedStage> recordedStages;  // every other launch of every pair, as recorded
    std::vector<StageGroup> stageGroups;        // in issue order
    std::vector<pyr::StageTask> hostStageTasks;
    std::vector<int> hostStageStart;
    pyr::StageTask* dStageTasks = nullptr;
    int* dStageStart = nullptr;
    pyr::StageIo* dIo = nullptr;
    long long stageLaunches = 0;
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
                // The pair's other launches of this level; level 0 takes the
                // caller's images (and writes the caller's outputs) per call.
                pyr::StageSink ss;
                const int st = pyr::kStagePrep | pyr::kStageCost | pyr::kStagePost;
                if (c.levels == 0)
                    c.singleCall(img, img, d16, {}, {}, nullptr, st, 1, nullptr, &ss);
                else
                    c.levelCall(level, img, img, d16, {}, {}, nullptr, st, 1, nullptr, &ss);
                for (pyr::StageLaunch& l : ss.launches) {
                    if (level == 0) l.io = l.kind == pyr::kStPost ? pyr::kIoOut : pyr::kIoIn;
                    recordedStages.push_back({i, level, 0, l});
                }
            }
            if (c.levels > 0) {
                pyr::StageSink ss;
                c.pyrDownChain(img, img, nullptr, 1, &ss);
                if ((int)ss.launches.size() != c.levels) throw std::logic_error("pyrsgm: pyrDown launches not recorded");
                for (int k = 0; k < c.levels; ++k) {
                    pyr::StageLaunch& l = ss.launches[(size_t)k];
                    if (k == 0) l.io = pyr::kIoIn;  // the first reduction reads the caller's images
                    recordedStages.push_back({i, -1, k + 1, l});
                }
            }
        }
    }

    // The queued stage launches: every recorded stage launch of one step (or
    // pyrDown depth) and phase, grouped by kernel, largest tasks first.
    void planStages()
    {
        stageGroups.clear();
        hostStageTasks.clear();
        hostStageStart.clear();
        stageLaunches = 0;
        if (!tun.queued_stages) return;
        auto addGroups = [&](int step, int phase, const std::vector<int>& members) {
            std::map<std::pair<int, int>, std::vector<int>> byKernel;
            for (int r : members) byKernel[{recordedStages[r].l.kind, recordedStages[r].l.variant}].push_back(r);
            for (auto& kv : byKernel) {
                std::vector<int>& rs = kv.second;
                auto blocksOf = [&](int r) {
                    const dim3& g = recordedStages[r].l.grid;
                    return (long long)g.x * g.y * g.z;
                };
                std::stable_sort(rs.begin(), rs.end(), [&](int x, int y) { return blocksOf(x) > blocksOf(y); });
                StageGroup g;
                g.step = step;
                g.phase = phase;
                g.kind = kv.first.first;
                g.variant = kv.first.second;
                g.taskOff = (int)hostStageTasks.size();
                g.startOff = hostStageStart.size();
                long long total = 0;
                for (int r : rs) {
                    const RecordedStage& rec = recordedStages[r];
                    pyr::StageTask t;
                    memset(&t, 0, sizeof(t));
                    t.a = rec.l.a;
                    t.gx = (int)rec.l.grid.x;
                    t.gy = (int)rec.l.grid.y;
                    t.io = rec.l.io;
                    t.pair = rec.pair;
                    hostStageStart.push_back((int)total);
                    total += blocksOf(r);
                    g.lds = std::max(g.lds, rec.l.lds);
                    hostStageTasks.push_back(t);
                }
                if (total > INT32_MAX) throw std::length_error("pyrsgm: queued stage launch too large");
                hostStageStart.push_back((int)total);
                g.nTasks = (int)rs.size();
                g.blocks = (int)total;
                stageGroups.push_back(g);
                ++stageLaunches;
            }
        };
        int maxDepth = 0;
        for (auto& c : cores) maxDepth = std::max(maxDepth, c->levels);
        for (int depth = 1; depth <= maxDepth; ++depth) {
            std::vector<int> m;
            for (int r = 0; r < (int)recordedStages.size(); ++r)
                if (recordedStages[r].depth == depth) m.push_back(r);
            addGroups(depth - 1, kPhPyrDown, m);
        }
        for (int step = 0; step < maxSteps; ++step)
            for (int phase : {kPhPrep, kPhCost, kPhCost2, kPhPost}) {
                std::vector<int> m;
                for (int r = 0; r < (int)recordedStages.size(); ++r) {
                    const RecordedStage& rec = recordedStages[r];
                    if (rec.depth == 0 && levelAt(rec.pair, step) == rec.level && phaseOf(rec.l.kind) == phase)
                        m.push_back(r);
                }
                addGroups(step, phase, m);
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
            if (!tun.queued_stages)
                for (int i = 0; i < (int)cores.size(); ++i)
                    if (levelAt(i, step) >= 0) perPairLaunches += 3;  // prep, cost, post
        }
        if (!tun.queued_stages)
            for (auto& c : cores) perPairLaunches += c->levels;  // pyrDown
        planStages();
        // Device tables: path tasks, starts, two counters per group; stage
        // tasks, their starts and the per-pair I/O table.
        auto round = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t tasksBytes = round(hostTasks.size() * sizeof(pyr::PathArgs));
        const size_t startBytes = round(hostStart.size() * sizeof(int));
        const size_t wtaBytes = round(hostWtaStart.size() * sizeof(int));
        const size_t counterBytes = round(groups.size() * 2 * sizeof(uint32_t));
        const size_t stageBytes = round(hostStageTasks.size() * sizeof(pyr::StageTask));
        const size_t stageStartBytes = round(hostStageStart.size() * sizeof(int));
        const size_t ioBytes = round(cores.size() * sizeof(pyr::StageIo));
        const size_t need = tasksBytes + startBytes + wtaBytes + counterBytes + stageBytes +
                            stageStartBytes + ioBytes;
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
        char* const stageBase = base + tasksBytes + startBytes + wtaBytes + counterBytes;
        dStageTasks = reinterpret_cast<pyr::StageTask*>(stageBase);
        dStageStart = reinterpret_cast<int*>(stageBase + stageBytes);
        dIo = reinterpret_cast<pyr::StageIo*>(stageBase + stageBytes + stageStartBytes);
        if (!hostStageTasks.empty()) {
            check(hipMemcpy(dStageTasks, hostStageTasks.data(), hostStageTasks.size() * sizeof(pyr::StageTask),
                            hipMemcpyHostToDevice), "hipMemcpy");
            check(hipMemcpy(dStageStart, hostStageStart.data(), hostStageStart.size() * sizeof(int),
                            hipMemcpyHostToDevice), "hipMemcpy");
        }
        check(hipMemset(dIo, 0, ioBytes), "hipMemset");
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

    // The queued path and winner-take-all launches of one step (in compute()
    // and computeQueued()).
    void pathGroups(const pyrsgm_pair_io_c* io, hipStream_t main, int step, size_t& gi)
    {
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
    }

    // The whole pass as queued launches on one stream: the caller's buffers
    // into the I/O table, the pyrDown depths, then per step the descriptor,
    // cost, path, winner-take-all and post launches, each over every pair
    // that has a level at that step.
    void computeQueued(const pyrsgm_pair_io_c* io, hipStream_t main)
    {
        const int n = (int)cores.size();
        for (int p0 = 0; p0 < n; p0 += pyr::kMaxIoPatch) {
            pyr::IoPatch patch;
            memset(&patch, 0, sizeof(patch));
            patch.table = dIo;
            patch.n = std::min(pyr::kMaxIoPatch, n - p0);
            for (int k = 0; k < patch.n; ++k) {
                const int i = p0 + k;
                const MatcherCore& c = *cores[i];
                const BatchBuffer<float> oc = c.confRoute({io[i].disp32, io[i].disp32_pitch, 0},
                                                          {io[i].conf, io[i].conf_pitch, 0});
                pyr::StageIo& x = patch.io[k];
                x.left = io[i].left;
                x.right = io[i].right;
                x.lstep = io[i].left_pitch;
                x.rstep = io[i].right_pitch;
                x.out16 = io[i].disp16;
                x.s16 = io[i].disp16_pitch;
                x.out32 = io[i].disp32;
                x.s32 = io[i].disp32_pitch;
                x.outc = oc.data;
                x.sc = oc.pitch;
                patch.idx[k] = i;
            }
            check(pyr::launchIoPatch(patch, main), "I/O table");
        }
        size_t sg = 0, gi = 0;
        auto stages = [&](int step, int lastPhase) {
            for (; sg < stageGroups.size() && stageGroups[sg].step == step && stageGroups[sg].phase <= lastPhase; ++sg) {
                const StageGroup& g = stageGroups[sg];
                pyr::StageQueue q;
                q.tasks = dStageTasks + g.taskOff;
                q.start = dStageStart + g.startOff;
                q.io = dIo;
                q.nTasks = g.nTasks;
                const hipError_t e = pyr::launchStageQueued(g.kind, g.variant, q, g.blocks, g.lds, main);
                if (e != hipSuccess) {
                    std::ostringstream oss;
                    oss << "queued stage (kind " << g.kind << ", step " << step << ", " << g.nTasks << " tasks)";
                    throwHip(oss.str(), e);
                }
            }
        };
        // pyrDown groups carry step = depth - 1 and come first.
        while (sg < stageGroups.size() && stageGroups[sg].phase == kPhPyrDown) {
            const int d = stageGroups[sg].step;
            stages(d, kPhPyrDown);
        }
        for (int step = 0; step < maxSteps; ++step) {
            stages(step, kPhCost2);
            pathGroups(io, main, step, gi);
            stages(step, kPhPost);
        }
    }

    void compute(const pyrsgm_pair_io_c* io, hipStream_t main)
    {
        checkIO(io);
        if (tun.queued_stages) {
            computeQueued(io, main);
            return;
        }
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
            pathGroups(io, main, step, gi);
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
        for (int qs : {1, 0}) {
            MixedTuning t = best;
            if (t.queued_stages == qs) continue;
            t.queued_stages = qs;
            consider(t);
        }
        // Streams matter only to the per-pair stage launches.
        if (!best.queued_stages)
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
        tuning.blocks_per_cu == MixedTuning{}.blocks_per_cu && tuning.paired == MixedTuning{}.paired &&
        tuning.queued_stages == MixedTuning{}.queued_stages) {
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
    return {(long long)impl_->groups.size(), tasks, items, impl_->perPairLaunches,
            impl_->stageLaunches, (long long)impl_->hostStageTasks.size()};
}

} // namespace pyrsgm
