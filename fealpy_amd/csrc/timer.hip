// In-step kernel timer (fa_timer_arm / fa_timer_read): the launches of the
// armed kernel families are bracketed by HIP events recorded on the stream
// each launch goes to, so a benchmark can time the kernels of the real step
// (not a stand-alone replay) and set them against the roofline.
#include <cstring>

#include "common.hpp"

namespace fa {

namespace {

constexpr int TIMER_KINDS = 16;
constexpr int TIMER_MAX = 128;  // timed launches per kernel family and arming

struct TimerSlot {
    hipEvent_t ev[2 * TIMER_MAX];
    int count = 0;
    bool open = false;
    bool created = false;
};

unsigned g_armed = 0;
TimerSlot g_slot[TIMER_KINDS];

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

}  // namespace

void timer_begin(int kind, hipStream_t s) {
    if (kind <= 0 || kind >= TIMER_KINDS || !(g_armed & (1u << kind))) return;
    TimerSlot& t = g_slot[kind];
    if (t.count >= TIMER_MAX || capturing(s)) return;
    if (hipEventRecord(t.ev[2 * t.count], s) == hipSuccess) t.open = true;
}

void timer_end(int kind, hipStream_t s) {
    if (kind <= 0 || kind >= TIMER_KINDS || !(g_armed & (1u << kind))) return;
    TimerSlot& t = g_slot[kind];
    if (!t.open) return;
    t.open = false;
    if (hipEventRecord(t.ev[2 * t.count + 1], s) == hipSuccess) ++t.count;
}

}  // namespace fa

using namespace fa;

extern "C" int fa_timer_arm(unsigned mask) {
    FA_REQUIRE((mask & 1u) == 0 && mask < (1u << TIMER_KINDS), "fa_timer_arm: bad kernel mask 0x%x", mask);
    for (int k = 1; k < TIMER_KINDS; ++k) {
        if (!(mask & (1u << k)) || g_slot[k].created) continue;
        for (int i = 0; i < 2 * TIMER_MAX; ++i) FA_HIP(hipEventCreate(&g_slot[k].ev[i]));
        g_slot[k].created = true;
    }
    for (int k = 0; k < TIMER_KINDS; ++k) {
        g_slot[k].count = 0;
        g_slot[k].open = false;
    }
    g_armed = mask;
    return FA_OK;
}

extern "C" int fa_timer_read(int kind, int64_t* launches, double* total_ms) {
    FA_REQUIRE(kind > 0 && kind < TIMER_KINDS && launches && total_ms, "fa_timer_read: bad arguments");
    TimerSlot& t = g_slot[kind];
    double sum = 0.0;
    for (int i = 0; i < t.count; ++i) {
        FA_HIP(hipEventSynchronize(t.ev[2 * i + 1]));
        float ms = 0.0f;
        FA_HIP(hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]));
        sum += ms;
    }
    *launches = t.count;
    *total_ms = sum;
    return FA_OK;
}
