// Kernel naming and per-stage timing of the drivers (no reference
// counterpart: the reference times whole solves with host timers,
// WaveletPotentials/TestLocalProjection.cpp:757-775).
#include <cstdio>
#include <string>

#include "whhd_internal.h"

namespace whhd {

namespace {
thread_local char g_last_kernel[128] = "";
}

const char* wavelet_name(int wt) {
    // lower-case names of WaveletPotentials/WaveletTypes.cpp:4-67
    static const char* const names[] = {
        "cdf3.9", "cdf3.11", "cdf3.5", "cdf2.10", "cdf2.12", "cdf2.6", "cdf4.4",
        "cdf4.6", "cdf4.8", "cdf4.10", "cdf2.8", "cdf3.7", "cdf26.44", "cdf28.46",
        "cdf44.26", "cdf46.28", "cdf2.0", "cdf1.1", "cdf1.7"};
    return (wt >= 0 && wt < (int)(sizeof(names) / sizeof(names[0]))) ? names[wt] : "?";
}

const char* last_kernel() { return g_last_kernel; }

void set_last_kernel(const char* name) {
    std::snprintf(g_last_kernel, sizeof(g_last_kernel), "%s", name);
}

void StageTrace::release() {
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    ev.clear();
    labels.clear();
    cap = solves = stages = 0;
    cur = -1;
}

int StageTrace::begin(int max_solves) {
    if (max_solves < 1 || max_solves > 4096)
        return fail(WHHD_INVALID_ARG, "trace_begin: max_solves must be in [1, 4096]");
    release();
    ev.assign((size_t)max_solves * (kMaxStages + 1), nullptr);
    for (hipEvent_t& e : ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) {
            e = nullptr;
            release();
            return hip_fail(r, "trace_begin: hipEventCreate");
        }
    }
    cap = max_solves;
    return WHHD_OK;
}

void StageTrace::solve_start(hipStream_t s) {
    cur = -1;
    if (solves >= cap) return;
    if (hipEventRecord(ev[(size_t)solves * (kMaxStages + 1)], s) == hipSuccess) cur = 0;
}

void StageTrace::mark(hipStream_t s, const char* stage) {
    if (cur < 0 || cur >= kMaxStages) return;
    if (hipEventRecord(ev[(size_t)solves * (kMaxStages + 1) + cur + 1], s) != hipSuccess) {
        cur = -1;
        return;
    }
    if (solves == 0) labels.push_back(std::string(stage) + " | " + last_kernel());
    ++cur;
}

void StageTrace::solve_end() {
    if (cur < 0) return;
    if (solves == 0) stages = cur;
    if (cur == stages) ++solves;  // a solve with a different stage list is not counted
    cur = -1;
}

int StageTrace::end(float* stage_us, int max_stages, int* n_stages, int* n_solves) {
    if (cap == 0) return fail(WHHD_INVALID_ARG, "trace_end: no trace_begin");
    if (n_stages) *n_stages = stages;
    if (n_solves) *n_solves = solves;
    if (solves == 0) return WHHD_OK;
    hipError_t r = hipEventSynchronize(ev[(size_t)(solves - 1) * (kMaxStages + 1) + stages]);
    if (r != hipSuccess) return hip_fail(r, "trace_end: hipEventSynchronize");
    for (int k = 0; k < stages && k < max_stages; ++k) {
        double sum = 0.0;
        for (int i = 0; i < solves; ++i) {
            float ms = 0.f;
            const size_t b = (size_t)i * (kMaxStages + 1);
            r = hipEventElapsedTime(&ms, ev[b + k], ev[b + k + 1]);
            if (r != hipSuccess) return hip_fail(r, "trace_end: hipEventElapsedTime");
            sum += ms;
        }
        if (stage_us) stage_us[k] = (float)(1000.0 * sum / solves);
    }
    return WHHD_OK;
}

}  // namespace whhd
