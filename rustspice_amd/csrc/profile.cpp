// profile.cpp - see profile.hpp.
#include "profile.hpp"

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/rsgpu.h"

namespace rsgpu {
namespace {
struct Rec {
    hipEvent_t a, b;
    int kernel;
    bool closed;  // b recorded
};
bool g_on = false;
std::vector<Rec> g_recs;
int g_open = 0;  // records begun and not yet ended (nested intervals)
double g_ms[PK_COUNT];
int64_t g_cnt[PK_COUNT];

// Folds every record into the totals. Only called with no interval open:
// an open record's end event is not recorded yet, and the elapsed time of an
// unrecorded event is an error that HIP would also leave as the thread's
// last error, to be reported by the next launch check.
void drain() {
    for (auto& r : g_recs) {
        float ms = 0.f;
        if (r.closed && hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            g_ms[r.kernel] += ms;
            g_cnt[r.kernel]++;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_recs.clear();
    g_open = 0;
    (void)hipGetLastError();
}
}  // namespace

bool prof_on() { return g_on; }

int prof_begin(int kernel, hipStream_t s) {
    if (!g_on) return -1;
    if (g_recs.size() >= 4096 && g_open == 0) drain();
    Rec r{};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
    r.kernel = kernel;
    r.closed = false;
    (void)hipEventRecord(r.a, s);
    g_recs.push_back(r);
    g_open++;
    return (int)g_recs.size() - 1;
}

void prof_end(int slot, hipStream_t s) {
    if (slot < 0 || slot >= (int)g_recs.size() || g_recs[slot].closed) return;
    (void)hipEventRecord(g_recs[slot].b, s);
    g_recs[slot].closed = true;
    g_open--;
}
}  // namespace rsgpu

extern "C" {
void rs_profile_enable(int32_t on) {
    rsgpu::drain();
    std::memset(rsgpu::g_ms, 0, sizeof rsgpu::g_ms);
    std::memset(rsgpu::g_cnt, 0, sizeof rsgpu::g_cnt);
    rsgpu::g_on = on != 0;
}

void rs_profile_get(double* ms_total, int64_t* counts) {
    rsgpu::drain();
    for (int k = 0; k < rsgpu::PK_COUNT; k++) {
        ms_total[k] = rsgpu::g_ms[k];
        counts[k] = rsgpu::g_cnt[k];
    }
}
}
