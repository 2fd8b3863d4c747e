// Batched recovery of many independent grids of uneven sizes.  No reference
// counterpart: the reference recovers one grid per CuWaveletPotentialRecover3D
// handle (WaveletPotentials/cuWaveletPotentialRecover3D.cu:6-61); each problem
// here gets exactly that solve.
//
// Problems with equal extents and spacing form a size class.  A class is one
// recovery driver over its problems stored back to back, so each of its 19
// stages is one launch whose grid rows are the problems.  Classes are
// independent and run as parallel branches: the first on the given stream, the
// others on internal streams forked from it and joined back with events.
// Captured once, the whole batch is one hipGraph replayed by a single call.
// The caller owns the arena holding every problem's six arrays; `layout` says
// where each lies.  The autotuner times the launch shapes of each class (line
// passes: 1, 2 or 4 wavefronts per workgroup; projection: 1, 2, 4 or 8 z cells
// per thread) on the class's own problems and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "whhd_internal.h"

struct whhd_batch3d_s {
    struct Class {
        int nx, ny, nz;
        double dx, dy, dz;
        std::vector<int> members;  // problem indices in slot order
        whhd_recover3d_s* rec = nullptr;
        int64_t base[6] = {};  // arena offsets (floats) of the xv, yv, zv, qx, qy, qz regions
        int64_t size[6] = {};  // floats per problem in each region
        int waves = 4, zb = 4;
        hipStream_t stream = nullptr;  // classes after the first
        hipEvent_t joined = nullptr;
    };
    std::vector<Class> cls;
    std::vector<int> cls_of, slot_of;
    int64_t arena_floats = 0;
    hipStream_t capture = nullptr;
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float* graph_arena = nullptr;
};

namespace whhd {
namespace {

void drop_graph(whhd_batch3d_s* h) {
    if (h->exec) (void)hipGraphExecDestroy(h->exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->exec = nullptr;
    h->graph = nullptr;
}

void release(whhd_batch3d_s* h) {
    if (!h) return;
    drop_graph(h);
    for (auto& c : h->cls) {
        if (c.rec) whhd_recover3d_destroy(c.rec);
        if (c.stream) (void)hipStreamDestroy(c.stream);
        if (c.joined) (void)hipEventDestroy(c.joined);
    }
    if (h->capture) (void)hipStreamDestroy(h->capture);
    if (h->fork) (void)hipEventDestroy(h->fork);
    delete h;
}

int solve_class(whhd_batch3d_s::Class& c, float* arena, hipStream_t s) {
    float* p[6];
    for (int k = 0; k < 6; ++k) p[k] = arena + c.base[k];
    return recover3d_solve(c.rec, p[3], p[4], p[5], nullptr, nullptr, nullptr, p[0], p[1], p[2], s);
}

// All classes: the first on `s`, the rest forked from `s` and joined back.
int enqueue(whhd_batch3d_s* h, float* arena, hipStream_t s) {
    const size_t nc = h->cls.size();
    if (nc > 1 && hipEventRecord(h->fork, s) != hipSuccess)
        return fail(WHHD_HIP_ERROR, "batch3d: fork event");
    for (size_t i = 1; i < nc; ++i) {
        auto& c = h->cls[i];
        if (hipStreamWaitEvent(c.stream, h->fork, 0) != hipSuccess)
            return fail(WHHD_HIP_ERROR, "batch3d: fork wait");
        int st = solve_class(c, arena, c.stream);
        if (st != WHHD_OK) return st;
        if (hipEventRecord(c.joined, c.stream) != hipSuccess)
            return fail(WHHD_HIP_ERROR, "batch3d: join event");
    }
    int st = solve_class(h->cls[0], arena, s);
    if (st != WHHD_OK) return st;
    for (size_t i = 1; i < nc; ++i)
        if (hipStreamWaitEvent(s, h->cls[i].joined, 0) != hipSuccess)
            return fail(WHHD_HIP_ERROR, "batch3d: join wait");
    return WHHD_OK;
}

int capture(whhd_batch3d_s* h, float* arena) {
    drop_graph(h);
    hipError_t e = hipStreamBeginCapture(h->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "batch3d: begin capture");
    const int st = enqueue(h, arena, h->capture);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(h->capture, &g);
    if (st != WHHD_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "batch3d: end capture");
    h->graph = g;
    e = hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(h);
        return hip_fail(e, "batch3d: graph instantiate");
    }
    h->graph_arena = arena;
    return WHHD_OK;
}

bool proj_fits(const whhd_batch3d_s::Class& c, int zb) {
    return (int64_t)((c.nz + zb - 1) / zb) * (int64_t)c.members.size() <= 65535;
}

// Average time of `reps` solves of class c with its current launch shapes.
int time_class(whhd_batch3d_s::Class& c, float* arena, int reps, hipStream_t s, float* ms) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        if (a) (void)hipEventDestroy(a);
        return fail(WHHD_HIP_ERROR, "batch3d autotune: events");
    }
    int st = solve_class(c, arena, s);  // warm-up
    if (st == WHHD_OK && hipEventRecord(a, s) != hipSuccess) st = fail(WHHD_HIP_ERROR, "autotune event");
    for (int r = 0; r < reps && st == WHHD_OK; ++r) st = solve_class(c, arena, s);
    if (st == WHHD_OK && hipEventRecord(b, s) != hipSuccess) st = fail(WHHD_HIP_ERROR, "autotune event");
    if (st == WHHD_OK) {
        hipError_t e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
        if (e != hipSuccess) st = hip_fail(e, "batch3d autotune: timing");
        *ms /= reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return st;
}

}  // namespace
}  // namespace whhd

extern "C" {

WHHD_API int whhd_batch3d_create(int n, const int* dims, const double* spacing, char bc,
                                 int wavelet_curl, int wavelet_div, whhd_batch3d_t* out) {
    using namespace whhd;
    if (!out) return fail(WHHD_INVALID_ARG, "batch3d_create: null handle pointer");
    *out = nullptr;
    if (n < 1 || !dims || !spacing) return fail(WHHD_INVALID_ARG, "batch3d_create: empty batch");
    auto* h = new whhd_batch3d_s;
    h->cls_of.assign(n, -1);
    h->slot_of.assign(n, -1);
    for (int i = 0; i < n; ++i) {
        const int* d = dims + 3 * i;
        const double* x = spacing + 3 * i;
        int c = 0;
        for (; c < (int)h->cls.size(); ++c) {
            const auto& k = h->cls[c];
            if (k.nx == d[0] && k.ny == d[1] && k.nz == d[2] && k.dx == x[0] && k.dy == x[1] &&
                k.dz == x[2])
                break;
        }
        if (c == (int)h->cls.size()) {
            whhd_batch3d_s::Class k;
            k.nx = d[0]; k.ny = d[1]; k.nz = d[2];
            k.dx = x[0]; k.dy = x[1]; k.dz = x[2];
            h->cls.push_back(k);
        }
        h->slot_of[i] = (int)h->cls[c].members.size();
        h->cls[c].members.push_back(i);
        h->cls_of[i] = c;
    }
    // biggest class first: it runs on the caller's stream
    std::vector<int> order(h->cls.size());
    for (size_t c = 0; c < order.size(); ++c) order[c] = (int)c;
    auto work = [&](int c) {
        const auto& k = h->cls[c];
        return (double)k.nx * k.ny * k.nz * (double)k.members.size();
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    std::vector<whhd_batch3d_s::Class> sorted;
    for (int c : order) sorted.push_back(h->cls[c]);
    h->cls.swap(sorted);
    int64_t off = 0;
    for (int c = 0; c < (int)h->cls.size(); ++c) {
        auto& k = h->cls[c];
        for (int p : k.members) h->cls_of[p] = c;
        const int64_t X = k.nx, Y = k.ny, Z = k.nz;
        const int64_t sz[6] = {(X + 1) * Y * Z,       X * (Y + 1) * Z,       X * Y * (Z + 1),
                               X * (Y + 1) * (Z + 1), (X + 1) * Y * (Z + 1), (X + 1) * (Y + 1) * Z};
        for (int r = 0; r < 6; ++r) {
            k.size[r] = sz[r];
            k.base[r] = off;
            off += (sz[r] * (int64_t)k.members.size() + 63) / 64 * 64;  // 256-byte aligned regions
        }
        int st = recover3d_create(k.nx, k.ny, k.nz, k.dx, k.dy, k.dz, bc, wavelet_curl, wavelet_div,
                                  &k.rec, (int)k.members.size());
        if (st == WHHD_OK && !proj_fits(k, 4))
            st = fail(WHHD_INVALID_ARG, "batch3d_create: too many problems in one size class");
        if (st == WHHD_OK) st = recover3d_set_tiles(k.rec, k.waves, k.zb);
        if (st == WHHD_OK && c > 0 &&
            (hipStreamCreateWithFlags(&k.stream, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreateWithFlags(&k.joined, hipEventDisableTiming) != hipSuccess))
            st = fail(WHHD_HIP_ERROR, "batch3d_create: stream or event");
        if (st != WHHD_OK) {
            release(h);
            return st;
        }
    }
    h->arena_floats = off;
    if (hipStreamCreateWithFlags(&h->capture, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) != hipSuccess) {
        release(h);
        return fail(WHHD_HIP_ERROR, "batch3d_create: stream or event");
    }
    *out = h;
    return WHHD_OK;
}

WHHD_API int whhd_batch3d_layout(whhd_batch3d_t h, long long* arena_floats, long long* offsets) {
    using namespace whhd;
    if (!h) return fail(WHHD_INVALID_ARG, "batch3d_layout: null handle");
    if (arena_floats) *arena_floats = h->arena_floats;
    if (offsets)
        for (size_t i = 0; i < h->cls_of.size(); ++i) {
            const auto& k = h->cls[h->cls_of[i]];
            for (int r = 0; r < 6; ++r) offsets[6 * i + r] = k.base[r] + h->slot_of[i] * k.size[r];
        }
    return WHHD_OK;
}

WHHD_API int whhd_batch3d_solve(whhd_batch3d_t h, float* arena, void* stream) {
    using namespace whhd;
    if (!h || !arena) return fail(WHHD_INVALID_ARG, "batch3d_solve: null argument");
    return enqueue(h, arena, (hipStream_t)stream);
}

WHHD_API int whhd_batch3d_capture(whhd_batch3d_t h, float* arena) {
    using namespace whhd;
    if (!h || !arena) return fail(WHHD_INVALID_ARG, "batch3d_capture: null argument");
    return capture(h, arena);
}

WHHD_API int whhd_batch3d_replay(whhd_batch3d_t h, void* stream) {
    using namespace whhd;
    if (!h) return fail(WHHD_INVALID_ARG, "batch3d_replay: null handle");
    if (!h->exec) return fail(WHHD_INVALID_ARG, "batch3d_replay: no captured graph");
    hipError_t e = hipGraphLaunch(h->exec, (hipStream_t)stream);
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "batch3d: graph launch");
}

WHHD_API int whhd_batch3d_autotune(whhd_batch3d_t h, float* arena, int reps, void* stream) {
    using namespace whhd;
    if (!h || !arena || reps < 1) return fail(WHHD_INVALID_ARG, "batch3d_autotune: bad argument");
    hipStream_t s = (hipStream_t)stream;
    for (auto& c : h->cls) {
        // line-pass shape first (with the current projection shape), then the projection's
        float best = INFINITY;
        int bw = c.waves;
        for (int w : {1, 2, 4}) {
            int st = recover3d_set_tiles(c.rec, w, c.zb);
            float ms = 0.f;
            if (st == WHHD_OK) st = time_class(c, arena, reps, s, &ms);
            if (st != WHHD_OK) return st;
            if (ms < best) { best = ms; bw = w; }
        }
        c.waves = bw;
        best = INFINITY;
        int bz = c.zb;
        for (int z : {1, 2, 4, 8}) {
            if (!proj_fits(c, z)) continue;
            int st = recover3d_set_tiles(c.rec, c.waves, z);
            float ms = 0.f;
            if (st == WHHD_OK) st = time_class(c, arena, reps, s, &ms);
            if (st != WHHD_OK) return st;
            if (ms < best) { best = ms; bz = z; }
        }
        c.zb = bz;
        int st = recover3d_set_tiles(c.rec, c.waves, c.zb);
        if (st != WHHD_OK) return st;
    }
    if (h->exec) {
        // the graph holds the old launch shapes: capture it again on the same arena
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "batch3d autotune: sync");
        return capture(h, h->graph_arena);
    }
    return WHHD_OK;
}

WHHD_API int whhd_batch3d_num_classes(whhd_batch3d_t h) {
    return h ? (int)h->cls.size() : whhd::fail(WHHD_INVALID_ARG, "batch3d: null handle");
}

WHHD_API int whhd_batch3d_class_info(whhd_batch3d_t h, int c, int* info) {
    using namespace whhd;
    if (!h || !info || c < 0 || c >= (int)h->cls.size())
        return fail(WHHD_INVALID_ARG, "batch3d_class_info: bad argument");
    const auto& k = h->cls[c];
    info[0] = k.nx; info[1] = k.ny; info[2] = k.nz;
    info[3] = (int)k.members.size();
    info[4] = k.waves;
    info[5] = k.zb;
    return WHHD_OK;
}

WHHD_API int whhd_batch3d_set_tiles(whhd_batch3d_t h, int c, int lp_waves, int proj_zb) {
    using namespace whhd;
    if (!h || c < 0 || c >= (int)h->cls.size())
        return fail(WHHD_INVALID_ARG, "batch3d_set_tiles: bad argument");
    auto& k = h->cls[c];
    if (!proj_fits(k, proj_zb == 1 || proj_zb == 2 || proj_zb == 8 ? proj_zb : 4))
        return fail(WHHD_INVALID_ARG, "batch3d_set_tiles: projection z block too small for the class");
    int st = recover3d_set_tiles(k.rec, lp_waves, proj_zb);
    if (st != WHHD_OK) return st;
    k.waves = lp_waves;
    k.zb = proj_zb;
    drop_graph(h);  // a captured graph would keep the old shapes
    return WHHD_OK;
}

WHHD_API void whhd_batch3d_destroy(whhd_batch3d_t h) { whhd::release(h); }

}  // extern "C"
