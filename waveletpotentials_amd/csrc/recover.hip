// Potential-recovery drivers: CuWaveletPotentialRecover3D / 2D
// (WaveletPotentials/cuWaveletPotentialRecover3D.cu:6-239,
//  WaveletPotentials/cuWaveletPotentialRecover2D.cu:6-37).
//
// Per component, the velocity's own axis uses the divergence wavelet with
// bc_div = proj_bc(bc) and the other axes the curl wavelet with bc; the
// potential's own axis uses the curl wavelet and the other axes the
// divergence wavelet.  The handle owns its scratch (per component one
// velocity-sized array and one array of the larger of the velocity and the
// potential sizes, allocated at create: passes run out of place) and enqueues
// everything on the caller's stream without synchronizing, so a solve can be
// captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "whhd_internal.h"

struct whhd_project3d_s;
struct whhd_project2d_s;

namespace whhd {
int project3d_create(int, int, int, double, double, double, int, char, whhd_project3d_s**);
int project3d_run(const whhd_project3d_s*, float*, float*, float*, float*, float*, float*,
                  const float*, const float*, const float*, hipStream_t, int nb, int zb);
int project2d_create(int, int, double, double, int, char, whhd_project2d_s**);
int project2d_run(const whhd_project2d_s*, float*, const float*, const float*, hipStream_t);
}  // namespace whhd

extern "C" void whhd_project3d_destroy(whhd_project3d_s* h);
extern "C" void whhd_project2d_destroy(whhd_project2d_s* h);

struct whhd_recover3d_s {
    int nx, ny, nz;
    int lx, ly, lz;
    char bc, bd;
    int wc, wd;
    int nb = 1;        // problems per solve, arrays back to back
    int lp_waves = 1;  // wavefronts per workgroup of the line passes and z cells
    int proj_zb = 8;   // per projection thread: fastest at 256^3 (0.713 ms against
                       // 0.727 with 2 and 4, 0.753 with 4 and 8)
    int lp_waves_y = 1;  // the y (contiguous) passes: 0.691 ms against 0.695 with 2
    whhd_project3d_s* proj = nullptr;
    float* tmp[3] = {nullptr, nullptr, nullptr};  // transformed velocity (velocity-sized)
    // ping-pong scratch of component c: the potential's coefficients and the
    // middle pass of both of component c's three-pass chains, so it holds
    // max(velocity size, potential size) elements per problem
    float* sq[3] = {nullptr, nullptr, nullptr};
    size_t sq_elems[3] = {0, 0, 0};  // elements per problem of sq[c]
    whhd::StageTrace trace;
};

struct whhd_recover2d_s {
    int nx, ny, lx, ly;
    char bc, bd;
    int wc, wd;
    whhd_project2d_s* proj = nullptr;
    float* tmp[2] = {nullptr, nullptr};
};

namespace whhd {

#define TRY(e) do { int s_ = (e); if (s_ != WHHD_OK) return s_; } while (0)

// cuCutCoefs (WaveletPotentials/cuWaveletPotentialRecover3D.cu:123-197) over the
// (nx+1, ny+1, nz+1) node grid; writes that would fall outside an array are
// skipped.
__global__ void cut_kernel(float* qx, float* qy, float* qz, int nx, int ny, int nz, int lx, int ly,
                           int lz, int cut) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y > ny || x > nx) return;
    const int64_t NX = nx, NY = ny, NZ = nz;
    const int64_t iqx = y + (x - 1) * (NY + 1) + z * NX * (NY + 1);
    const int64_t iqy = (y - 1) + x * NY + z * (NX + 1) * NY;
    const int64_t iqz = y + x * (NY + 1) + (z - 1) * (NX + 1) * (NY + 1);
    const int64_t nqx = NX * (NY + 1) * (NZ + 1), nqy = (NX + 1) * NY * (NZ + 1),
                  nqz = (NX + 1) * (NY + 1) * NZ;
    auto bl = [](int v) { return v > 0 ? 32 - __clz(v) : 0; };
    const int jy = ly - bl(y - 1 > 0 ? y - 1 : 0);
    const int jx = lx - bl(x - 1 > 0 ? x - 1 : 0);
    const int jz = lz - bl(z - 1 > 0 ? z - 1 : 0);
    bool zx = false, zy = false, zz = false;
    if (jx < cut) { zx = true; if (y > 0) zy = true; if (z > 0) zz = true; }
    if (jy < cut) { zy = true; if (x > 0) zx = true; if (z > 0) zz = true; }
    if (jz < cut) { zz = true; if (x > 0) zx = true; if (y > 0) zy = true; }
    if (zx && iqx >= 0 && iqx < nqx) qx[iqx] = 0.f;
    if (zy && iqy >= 0 && iqy < nqy) qy[iqy] = 0.f;
    if (zz && iqz >= 0 && iqz < nqz) qz[iqz] = 0.f;
}

// Three line passes of one array: forward y, x, z; inverse z, y, x.  The
// separable transforms commute, so the order changes only FP32 rounding; the
// orders are the fastest of those measured at 256^3 (forward y, x, z against
// x, y, z: 0.697 vs 0.705 ms; inverse z, y, x against x, y, z: 0.695 vs
// 0.698 ms).  The line-parallel kernels read and write different arrays, so
// the passes ping-pong: in -> out -> scratch -> out.  `scratch` may be `in`
// (consumed by the first pass) but not `out`.
struct Pass {
    char bc;
    int wt;
};
// Stage labels for the trace: "<fwd|inv> <array> <axis>".
static void mark(StageTrace* tr, hipStream_t s, bool fwd, const char* arr, char axis) {
    if (!tr || tr->cur < 0) return;
    char lab[32];
    std::snprintf(lab, sizeof(lab), "%s %s %c", fwd ? "fwd" : "inv", arr, axis);
    tr->mark(s, lab);
}
static int three_pass(const float* in, float* scratch, size_t scratch_elems, float* out, int nx,
                      int ny, int nz, const int lv[3], Pass px, Pass py, Pass pz, bool fwd,
                      hipStream_t s, const whhd_recover3d_s* h, const char* arr) {
    if ((size_t)nx * ny * nz > scratch_elems)
        return fail(WHHD_INVALID_ARG, "recover3d: scratch smaller than the array a pass writes");
    StageTrace* tr = const_cast<StageTrace*>(&h->trace);
    const LineBatch b{h->nb, (int64_t)nx * ny * nz, h->lp_waves};
    const LineBatch by{h->nb, (int64_t)nx * ny * nz, h->lp_waves_y};
    if (fwd) {
        TRY(transform_lines<float>(out, in, nx, ny, nz, lv[1], 'y', py.bc, py.wt, fwd, s, nullptr, &by));
        mark(tr, s, fwd, arr, 'y');
        TRY(transform_lines<float>(scratch, out, nx, ny, nz, lv[0], 'x', px.bc, px.wt, fwd, s, nullptr, &b));
        mark(tr, s, fwd, arr, 'x');
        TRY(transform_lines<float>(out, scratch, nx, ny, nz, lv[2], 'z', pz.bc, pz.wt, fwd, s, nullptr, &b));
        mark(tr, s, fwd, arr, 'z');
        return WHHD_OK;
    }
    TRY(transform_lines<float>(out, in, nx, ny, nz, lv[2], 'z', pz.bc, pz.wt, fwd, s, nullptr, &b));
    mark(tr, s, fwd, arr, 'z');
    TRY(transform_lines<float>(scratch, out, nx, ny, nz, lv[1], 'y', py.bc, py.wt, fwd, s, nullptr, &by));
    mark(tr, s, fwd, arr, 'y');
    TRY(transform_lines<float>(out, scratch, nx, ny, nz, lv[0], 'x', px.bc, px.wt, fwd, s, nullptr, &b));
    mark(tr, s, fwd, arr, 'x');
    return WHHD_OK;
}

// Velocity component c (0: xv, 1: yv, 2: zv): its own axis uses the divergence
// wavelet, the others the curl wavelet.
static int vel_pass(whhd_recover3d_s* h, int c, const float* in, float* scratch, size_t cap,
                    float* out, bool fwd, hipStream_t s) {
    const Pass C{h->bc, h->wc}, D{h->bd, h->wd};
    const int lv[3] = {h->lx, h->ly, h->lz};
    const int X = h->nx, Y = h->ny, Z = h->nz;
    switch (c) {
    case 0: return three_pass(in, scratch, cap, out, X + 1, Y, Z, lv, D, C, C, fwd, s, h, "xv");
    case 1: return three_pass(in, scratch, cap, out, X, Y + 1, Z, lv, C, D, C, fwd, s, h, "yv");
    default: return three_pass(in, scratch, cap, out, X, Y, Z + 1, lv, C, C, D, fwd, s, h, "zv");
    }
}

// Potential component c (0: qx, 1: qy, 2: qz): its own axis uses the curl
// wavelet, the others the divergence wavelet.
static int pot_pass(whhd_recover3d_s* h, int c, const float* in, float* scratch, size_t cap,
                    float* out, bool fwd, hipStream_t s) {
    const Pass C{h->bc, h->wc}, D{h->bd, h->wd};
    const int lv[3] = {h->lx, h->ly, h->lz};
    const int X = h->nx, Y = h->ny, Z = h->nz;
    switch (c) {
    case 0: return three_pass(in, scratch, cap, out, X, Y + 1, Z + 1, lv, C, D, D, fwd, s, h, "qx");
    case 1: return three_pass(in, scratch, cap, out, X + 1, Y, Z + 1, lv, D, C, D, fwd, s, h, "qy");
    default: return three_pass(in, scratch, cap, out, X + 1, Y + 1, Z, lv, D, D, C, fwd, s, h, "qz");
    }
}

// Scratch of a recovery handle, elements per problem: tmp[c] holds velocity
// component c; sq[c] holds potential component c's coefficients and the
// middle pass of velocity c's and potential c's chains.
void recover3d_scratch_elems(int nx, int ny, int nz, size_t tmp[3], size_t sq[3]) {
    const size_t X = nx, Y = ny, Z = nz;
    const size_t vel[3] = {(X + 1) * Y * Z, X * (Y + 1) * Z, X * Y * (Z + 1)};
    const size_t pot[3] = {X * (Y + 1) * (Z + 1), (X + 1) * Y * (Z + 1), (X + 1) * (Y + 1) * Z};
    for (int c = 0; c < 3; ++c) {
        tmp[c] = vel[c];
        sq[c] = vel[c] > pot[c] ? vel[c] : pot[c];
    }
}

static bool pow2(int n) { return n >= 4 && (n & (n - 1)) == 0; }

int recover3d_create(int nx, int ny, int nz, double dx, double dy, double dz, char bc, int wc,
                     int wd, whhd_recover3d_s** out, int nb) {
    if (!out) return fail(WHHD_INVALID_ARG, "recover3d_create: null handle pointer");
    *out = nullptr;
    if (nb < 1 || nb > 4096) return fail(WHHD_INVALID_ARG, "recover3d_create: batch count out of range");
    if (!pow2(nx) || !pow2(ny) || !pow2(nz))
        return fail(WHHD_NOT_POW2, "recover3d: cell counts must be powers of two (>= 4)");
    const char bd = proj_bc(bc);
    TRY(check_wavelet(wc, bc));
    TRY(check_wavelet(wd, bd));
    TRY(check_line(nx, ilog2(nx), wc));
    TRY(check_line(nx + 1, ilog2(nx + 1), wd));
    TRY(check_line(ny, ilog2(ny), wc));
    TRY(check_line(ny + 1, ilog2(ny + 1), wd));
    TRY(check_line(nz, ilog2(nz), wc));
    TRY(check_line(nz + 1, ilog2(nz + 1), wd));
    auto* h = new whhd_recover3d_s;
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->lx = ilog2(nx); h->ly = ilog2(ny); h->lz = ilog2(nz);
    h->bc = bc; h->bd = bd; h->wc = wc; h->wd = wd;
    h->nb = nb;
    int st = project3d_create(nx, ny, nz, dx, dy, dz, wd, bd, &h->proj);
    if (st != WHHD_OK) {
        delete h;
        return st;
    }
    size_t sz[6];
    recover3d_scratch_elems(nx, ny, nz, sz, sz + 3);
    for (int c = 0; c < 3; ++c) h->sq_elems[c] = sz[3 + c];
    float** slot[6] = {&h->tmp[0], &h->tmp[1], &h->tmp[2], &h->sq[0], &h->sq[1], &h->sq[2]};
    for (int i = 0; i < 6; ++i)
        if (hipMalloc(slot[i], sz[i] * nb * sizeof(float)) != hipSuccess) {
            for (int j = 0; j < i; ++j) hipFree(*slot[j]);
            whhd_project3d_destroy(h->proj);
            delete h;
            return fail(WHHD_OOM, "recover3d: scratch allocation failed");
        }
    *out = h;
    return WHHD_OK;
}

int recover3d_solve(whhd_recover3d_s* h, float* qx, float* qy, float* qz, float* xo, float* yo,
                    float* zo, const float* xv, const float* yv, const float* zv, hipStream_t s) {
    if (!h || !qx || !qy || !qz || !xv || !yv || !zv)
        return fail(WHHD_INVALID_ARG, "recover3d_solve: null argument");
    const bool vout = xo || yo || zo;
    if (vout && !(xo && yo && zo))
        return fail(WHHD_INVALID_ARG, "recover3d_solve_proj: all three v_out arrays are needed");
    float* const* t = h->tmp;
    float* const* S = h->sq;
    const float* v[3] = {xv, yv, zv};
    float* q[3] = {qx, qy, qz};
    float* vo[3] = {xo, yo, zo};
    h->trace.solve_start(s);
    const size_t* cap = h->sq_elems;
    for (int c = 0; c < 3; ++c) TRY(vel_pass(h, c, v[c], S[c], cap[c], t[c], true, s));
    // q coefficients land in the scratch; with v_out, s x q overwrites the
    // transformed velocity in place at the projected coefficients
    if (vout)
        TRY(project3d_run(h->proj, S[0], S[1], S[2], t[0], t[1], t[2], t[0], t[1], t[2], s, h->nb,
                          h->proj_zb));
    else
        TRY(project3d_run(h->proj, S[0], S[1], S[2], nullptr, nullptr, nullptr, t[0], t[1], t[2], s,
                          h->nb, h->proj_zb));
    h->trace.mark(s, "project");
    for (int c = 0; c < 3; ++c) TRY(pot_pass(h, c, S[c], S[c], cap[c], q[c], false, s));
    if (vout)
        for (int c = 0; c < 3; ++c) TRY(vel_pass(h, c, t[c], S[c], cap[c], vo[c], false, s));
    h->trace.solve_end();
    return WHHD_OK;
}

int recover3d_set_tiles(whhd_recover3d_s* h, int lp_waves, int proj_zb) {
    if (!h) return fail(WHHD_INVALID_ARG, "recover3d_set_tiles: null handle");
    if (lp_waves != 1 && lp_waves != 2 && lp_waves != 4)
        return fail(WHHD_INVALID_ARG, "line-pass wavefronts per workgroup must be 1, 2 or 4");
    if (proj_zb != 1 && proj_zb != 2 && proj_zb != 4 && proj_zb != 8)
        return fail(WHHD_INVALID_ARG, "projection z block must be 1, 2, 4 or 8");
    h->lp_waves = lp_waves;
    h->lp_waves_y = lp_waves;
    h->proj_zb = proj_zb;
    return WHHD_OK;
}

int recover3d_cut(whhd_recover3d_s* h, float* ox, float* oy, float* oz, const float* qx,
                  const float* qy, const float* qz, int level, hipStream_t s) {
    if (!h || !ox || !oy || !oz || !qx || !qy || !qz)
        return fail(WHHD_INVALID_ARG, "recover3d_cut: null argument");
    if (h->nb != 1) return fail(WHHD_INVALID_ARG, "recover3d_cut: one problem per handle");
    float* const* S = h->sq;
    const float* q[3] = {qx, qy, qz};
    float* o[3] = {ox, oy, oz};
    h->trace.solve_start(s);
    // the forward passes ping-pong through the caller's output (potential-sized)
    const size_t pot_e[3] = {(size_t)h->nx * (h->ny + 1) * (h->nz + 1),
                             (size_t)(h->nx + 1) * h->ny * (h->nz + 1),
                             (size_t)(h->nx + 1) * (h->ny + 1) * h->nz};
    for (int c = 0; c < 3; ++c) TRY(pot_pass(h, c, q[c], o[c], pot_e[c], S[c], true, s));
    dim3 block(64, 4, 1);
    dim3 grid((h->ny + 1 + 63) / 64, (h->nx + 1 + 3) / 4, h->nz + 1);
    hipLaunchKernelGGL(cut_kernel, grid, block, 0, s, S[0], S[1], S[2], h->nx, h->ny, h->nz, h->lx,
                       h->ly, h->lz, level);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "cut kernel");
    set_last_kernel("cut_kernel");
    h->trace.mark(s, "cut");
    for (int c = 0; c < 3; ++c) TRY(pot_pass(h, c, S[c], S[c], h->sq_elems[c], o[c], false, s));
    h->trace.solve_end();
    return WHHD_OK;
}

int recover2d_create(int nx, int ny, double dx, double dy, char bc, int wc, int wd,
                     whhd_recover2d_s** out) {
    if (!out) return fail(WHHD_INVALID_ARG, "recover2d_create: null handle pointer");
    *out = nullptr;
    if (!pow2(nx) || !pow2(ny))
        return fail(WHHD_NOT_POW2, "recover2d: cell counts must be powers of two (>= 4)");
    const char bd = proj_bc(bc);
    TRY(check_wavelet(wc, bc));
    TRY(check_wavelet(wd, bd));
    TRY(check_line(nx, ilog2(nx), wc));
    TRY(check_line(nx + 1, ilog2(nx + 1), wd));
    TRY(check_line(ny, ilog2(ny), wc));
    TRY(check_line(ny + 1, ilog2(ny + 1), wd));
    auto* h = new whhd_recover2d_s;
    h->nx = nx; h->ny = ny; h->lx = ilog2(nx); h->ly = ilog2(ny);
    h->bc = bc; h->bd = bd; h->wc = wc; h->wd = wd;
    int st = project2d_create(nx, ny, dx, dy, wd, bd, &h->proj);
    if (st != WHHD_OK) {
        delete h;
        return st;
    }
    const size_t sz[2] = {(size_t)(nx + 1) * ny, (size_t)nx * (ny + 1)};
    for (int i = 0; i < 2; ++i)
        if (hipMalloc(&h->tmp[i], sz[i] * sizeof(float)) != hipSuccess) {
            for (int j = 0; j < i; ++j) hipFree(h->tmp[j]);
            whhd_project2d_destroy(h->proj);
            delete h;
            return fail(WHHD_OOM, "recover2d: scratch allocation failed");
        }
    *out = h;
    return WHHD_OK;
}

int recover2d_solve(whhd_recover2d_s* h, float* qz, const float* xv, const float* yv, hipStream_t s) {
    if (!h || !qz || !xv || !yv) return fail(WHHD_INVALID_ARG, "recover2d_solve: null argument");
    const int X = h->nx, Y = h->ny;
    float* tx = h->tmp[0];
    float* ty = h->tmp[1];
    TRY(transform_lines<float>(tx, xv, X + 1, Y, 1, h->lx, 'x', h->bd, h->wd, true, s));
    TRY(transform_lines<float>(tx, tx, X + 1, Y, 1, h->ly, 'y', h->bc, h->wc, true, s));
    TRY(transform_lines<float>(ty, yv, X, Y + 1, 1, h->lx, 'x', h->bc, h->wc, true, s));
    TRY(transform_lines<float>(ty, ty, X, Y + 1, 1, h->ly, 'y', h->bd, h->wd, true, s));
    TRY(project2d_run(h->proj, qz, tx, ty, s));
    TRY(transform_lines<float>(qz, qz, X + 1, Y + 1, 1, h->lx, 'x', h->bd, h->wd, false, s));
    TRY(transform_lines<float>(qz, qz, X + 1, Y + 1, 1, h->ly, 'y', h->bd, h->wd, false, s));
    return WHHD_OK;
}

}  // namespace whhd

extern "C" {

WHHD_API int whhd_recover3d_create(int nx, int ny, int nz, double dx, double dy, double dz, char bc,
                                   int wavelet_curl, int wavelet_div, whhd_recover3d_t* out) {
    return whhd::recover3d_create(nx, ny, nz, dx, dy, dz, bc, wavelet_curl, wavelet_div, out, 1);
}

WHHD_API int whhd_recover3d_solve(whhd_recover3d_t h, float* qx, float* qy, float* qz,
                                  const float* xv, const float* yv, const float* zv, void* stream) {
    return whhd::recover3d_solve(h, qx, qy, qz, nullptr, nullptr, nullptr, xv, yv, zv,
                                 (hipStream_t)stream);
}

WHHD_API int whhd_recover3d_solve_proj(whhd_recover3d_t h, float* qx, float* qy, float* qz,
                                       float* xv_out, float* yv_out, float* zv_out, const float* xv,
                                       const float* yv, const float* zv, void* stream) {
    if (!xv_out || !yv_out || !zv_out)
        return whhd::fail(WHHD_INVALID_ARG, "recover3d_solve_proj: v_out arrays are required");
    return whhd::recover3d_solve(h, qx, qy, qz, xv_out, yv_out, zv_out, xv, yv, zv,
                                 (hipStream_t)stream);
}

WHHD_API int whhd_recover3d_cut(whhd_recover3d_t h, float* qx_out, float* qy_out, float* qz_out,
                                const float* qx, const float* qy, const float* qz, int level,
                                void* stream) {
    return whhd::recover3d_cut(h, qx_out, qy_out, qz_out, qx, qy, qz, level, (hipStream_t)stream);
}

WHHD_API int whhd_recover3d_scratch_bytes(int nx, int ny, int nz, unsigned long long* bytes) {
    if (!bytes || nx < 1 || ny < 1 || nz < 1)
        return whhd::fail(WHHD_INVALID_ARG, "recover3d_scratch_bytes: bad argument");
    size_t t[3], q[3];
    whhd::recover3d_scratch_elems(nx, ny, nz, t, q);
    *bytes = (unsigned long long)(t[0] + t[1] + t[2] + q[0] + q[1] + q[2]) * sizeof(float);
    return WHHD_OK;
}

WHHD_API int whhd_recover3d_trace_begin(whhd_recover3d_t h, int max_solves) {
    if (!h) return whhd::fail(WHHD_INVALID_ARG, "trace_begin: null handle");
    return h->trace.begin(max_solves);
}

WHHD_API int whhd_recover3d_trace_end(whhd_recover3d_t h, float* stage_us, int max_stages,
                                      int* n_stages, int* n_solves) {
    if (!h) return whhd::fail(WHHD_INVALID_ARG, "trace_end: null handle");
    return h->trace.end(stage_us, max_stages, n_stages, n_solves);
}

WHHD_API const char* whhd_recover3d_trace_label(whhd_recover3d_t h, int stage) {
    if (!h || stage < 0 || stage >= (int)h->trace.labels.size()) return nullptr;
    return h->trace.labels[stage].c_str();
}

WHHD_API void whhd_recover3d_destroy(whhd_recover3d_t h) {
    if (!h) return;
    for (float* p : h->tmp) hipFree(p);
    for (float* p : h->sq) hipFree(p);
    whhd_project3d_destroy(h->proj);
    delete h;
}

WHHD_API int whhd_recover2d_create(int nx, int ny, double dx, double dy, char bc, int wavelet_curl,
                                   int wavelet_div, whhd_recover2d_t* out) {
    return whhd::recover2d_create(nx, ny, dx, dy, bc, wavelet_curl, wavelet_div, out);
}

WHHD_API int whhd_recover2d_solve(whhd_recover2d_t h, float* qz, const float* xv, const float* yv,
                                  void* stream) {
    return whhd::recover2d_solve(h, qz, xv, yv, (hipStream_t)stream);
}

WHHD_API void whhd_recover2d_destroy(whhd_recover2d_t h) {
    if (!h) return;
    for (float* p : h->tmp) hipFree(p);
    whhd_project2d_destroy(h->proj);
    delete h;
}

}  // extern "C"
