// Per-kernel timing of the library's own launches with HIP events.
//
// When enabled, every launch site brackets its kernel with a pair of events
// recorded on the launch stream and notes the launch's algorithmic bytes
// (compulsory reads + writes of that launch) and flops.  Summaries aggregate
// by kernel label after synchronising on the recorded events.  Disabled by
// default; no effect on results.
#include <atomic>
#include <map>
#include <vector>

#include "fm_common.hpp"

namespace fm {

namespace {
struct Rec {
    std::string name;
    hipEvent_t a, b;
    double bytes, flops, pass_bytes;
};
std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<Rec> g_recs;
struct Agg {
    int64_t launches = 0;
    double ms = 0, bytes = 0, flops = 0, pass_bytes = 0;
};
std::vector<std::pair<std::string, Agg>> g_summary;
}  // namespace

bool prof_on() { return g_on.load(std::memory_order_relaxed); }

void prof_begin(hipStream_t s, ProfToken *t) {
    t->active = false;
    if (!prof_on()) return;
    // launches recorded into a HIP graph are not timed (an event recorded
    // during capture becomes a graph node, not a timestamp)
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return;
    if (hipEventCreate(&t->a) != hipSuccess) return;
    if (hipEventCreate(&t->b) != hipSuccess) {
        (void)hipEventDestroy(t->a);
        return;
    }
    (void)hipEventRecord(t->a, s);
    t->active = true;
}

void prof_end(hipStream_t s, ProfToken *t, const char *name, double bytes, double flops,
              double pass_bytes) {
    if (!t->active) return;
    (void)hipEventRecord(t->b, s);
    std::lock_guard<std::mutex> lock(g_mu);
    g_recs.push_back(Rec{name, t->a, t->b, bytes, flops, pass_bytes < 0 ? bytes : pass_bytes});
}

static void clear_locked() {
    for (auto &r : g_recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_recs.clear();
    g_summary.clear();
}

}  // namespace fm

using namespace fm;

extern "C" fm_status fm_profiler_enable(int on) {
    g_on.store(on != 0);
    return FM_SUCCESS;
}

extern "C" fm_status fm_profiler_reset(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    clear_locked();
    return FM_SUCCESS;
}

extern "C" fm_status fm_profiler_count(int *count) {
    FM_REQUIRE(count != nullptr, FM_ERR_VALUE, "NULL argument");
    std::lock_guard<std::mutex> lock(g_mu);
    std::map<std::string, Agg> agg;
    for (auto &r : g_recs) {
        FM_CHECK_HIP(hipEventSynchronize(r.b));
        float ms = 0.f;
        FM_CHECK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        Agg &g = agg[r.name];
        g.launches += 1;
        g.ms += ms;
        g.bytes += r.bytes;
        g.flops += r.flops;
        g.pass_bytes += r.pass_bytes;
    }
    g_summary.assign(agg.begin(), agg.end());
    *count = int(g_summary.size());
    return FM_SUCCESS;
}

extern "C" fm_status fm_profiler_entry(int index, char *name, int name_len, int64_t *launches,
                                       double *total_ms, double *total_bytes,
                                       double *total_flops, double *total_pass_bytes) {
    std::lock_guard<std::mutex> lock(g_mu);
    FM_REQUIRE(index >= 0 && index < int(g_summary.size()), FM_ERR_VALUE,
               "profiler entry %d out of range (call fm_profiler_count first)", index);
    const auto &e = g_summary[index];
    if (name && name_len > 0) {
        strncpy(name, e.first.c_str(), size_t(name_len) - 1);
        name[name_len - 1] = 0;
    }
    if (launches) *launches = e.second.launches;
    if (total_ms) *total_ms = e.second.ms;
    if (total_bytes) *total_bytes = e.second.bytes;
    if (total_flops) *total_flops = e.second.flops;
    if (total_pass_bytes) *total_pass_bytes = e.second.pass_bytes;
    return FM_SUCCESS;
}
