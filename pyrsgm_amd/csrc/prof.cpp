// prof.cpp -- stage timing hook (pyrsgm_prof.h): HIP event pairs around the
// launches of each level, recorded on the launch stream while enabled.
#include <vector>

#include "internal.h"
#include "pyrsgm/pyrsgm_prof.h"

namespace pyr {
namespace {

struct Record {
    int stage, W, H, D;
    hipEvent_t a, b;
};

struct Prof {
    bool on = false;
    int fStage = -1, fW = -1, fH = -1;  // record filter; -1 matches anything
    int used = 0;
    std::vector<Record> recs;
    int open = -1;
};

Prof& prof()
{
    static Prof p;
    return p;
}

} // namespace

void profBegin(int stage, int W, int H, int D, hipStream_t s)
{
    Prof& p = prof();
    if (!p.on || p.used >= (int)p.recs.size()) return;
    if ((p.fStage >= 0 && stage != p.fStage) || (p.fW >= 0 && W != p.fW) ||
        (p.fH >= 0 && H != p.fH))
        return;
    Record& r = p.recs[p.used];
    r.stage = stage;
    r.W = W;
    r.H = H;
    r.D = D;
    if (hipEventRecord(r.a, s) != hipSuccess) return;
    p.open = p.used;
}

void profEnd(hipStream_t s)
{
    Prof& p = prof();
    if (!p.on || p.open < 0) return;
    if (hipEventRecord(p.recs[p.open].b, s) == hipSuccess) ++p.used;
    p.open = -1;
}

} // namespace pyr

using namespace pyr;

extern "C" {

int pyrsgm_prof_start(int max_records)
{
    Prof& p = prof();
    if ((int)p.recs.size() < max_records) {
        const size_t old = p.recs.size();
        p.recs.resize(max_records);
        for (size_t i = old; i < p.recs.size(); ++i) {
            if (hipEventCreate(&p.recs[i].a) != hipSuccess ||
                hipEventCreate(&p.recs[i].b) != hipSuccess) {
                p.recs.resize(i);
                return -1;
            }
        }
    }
    p.used = 0;
    p.open = -1;
    p.on = true;
    return 0;
}

void pyrsgm_prof_filter(int stage, int width, int height)
{
    Prof& p = prof();
    p.fStage = stage;
    p.fW = width;
    p.fH = height;
}

void pyrsgm_prof_stop(void) { prof().on = false; }

int pyrsgm_prof_count(void) { return prof().used; }

int pyrsgm_prof_get(int i, int* stage, int* width, int* height, int* labels, float* ms)
{
    Prof& p = prof();
    if (i < 0 || i >= p.used) return -1;
    Record& r = p.recs[i];
    if (hipEventSynchronize(r.b) != hipSuccess) return -1;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) return -1;
    if (stage) *stage = r.stage;
    if (width) *width = r.W;
    if (height) *height = r.H;
    if (labels) *labels = r.D;
    if (ms) *ms = t;
    return 0;
}

} // extern "C"
