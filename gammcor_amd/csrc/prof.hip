// Kernel timing registry behind gmc_prof_* (see prof.hpp).
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gammcor_hip.h"
#include "prof.hpp"

namespace gmc {

namespace {
struct Rec {
    std::string name;
    hipEvent_t a, b;
    double work;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool prof_enabled() { return g_on; }

void prof_release()
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = false;
    for (auto& r : g_recs) {
        (void)hipEventSynchronize(r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : g_pool) (void)hipEventDestroy(e);
    g_recs.clear();
    g_pool.clear();
}

ProfScope::ProfScope(hipStream_t st, const char* name, double work) : s(st)
{
    if (!g_on) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive) return;
    std::lock_guard<std::mutex> lk(g_mu);
    Rec r{name, take_event(), take_event(), work};
    if (!r.a || !r.b) return;
    if (hipEventRecord(r.a, st) != hipSuccess) return;
    g_recs.push_back(r);
    slot = (int)g_recs.size() - 1;
}

void ProfScope::close()
{
    if (slot < 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    (void)hipEventRecord(g_recs[slot].b, s);
    slot = -1;
}

}  // namespace gmc

using namespace gmc;

extern "C" {

int gmc_prof_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = on != 0;
    return 0;
}

int gmc_prof_reset(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& r : g_recs) {
        (void)hipEventSynchronize(r.b);
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_recs.clear();
    return 0;
}

int gmc_prof_get(const char* name, double* total_ms, int* count, double* total_work)
{
    std::lock_guard<std::mutex> lk(g_mu);
    double t = 0.0, w = 0.0;
    int c = 0;
    for (auto& r : g_recs) {
        if (r.name != name) continue;
        GMC_HIP(hipEventSynchronize(r.b));
        float ms = 0.f;
        GMC_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        t += ms;
        w += r.work;
        c++;
    }
    if (total_ms) *total_ms = t;
    if (count) *count = c;
    if (total_work) *total_work = w;
    return 0;
}

}  // extern "C"
