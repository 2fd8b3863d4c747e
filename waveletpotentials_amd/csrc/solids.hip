// Solid geometry on the cell grid: sphere level sets, face fractions and the
// eikonal relaxation (WaveletPotentials/cuSolidLevelSet3D.cu:11-90, 319-383;
// WaveletPotentials/cuFastSweeping.cu:98-257).
//
// All three are streaming stencils over y-fastest arrays, HBM-bound: one
// thread per cell (or face), a (64, 4) block over (y, x) and one z-plane per
// grid row, so consecutive lanes touch consecutive floats.
#include <hip/hip_runtime.h>

#include <cmath>

#include "whhd_internal.h"

namespace whhd {

namespace {

constexpr int kMaxSpheres = 64;

struct Spheres {
    float c[kMaxSpheres][4];  // centre x, y, z and radius
    int count;
};

// cuInitLsSphere (WaveletPotentials/cuSolidLevelSet3D.cu:11-26) for a union of
// spheres: the signed distance at cell centres, negative inside a solid.
__global__ void levelset_spheres_kernel(float* __restrict__ ls, int nx, int ny, int nz, float dx,
                                        float dy, float dz, Spheres s) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= ny || x >= nx) return;
    const float px = (x + 0.5f) * dx, py = (y + 0.5f) * dy, pz = (z + 0.5f) * dz;
    float v = INFINITY;
    for (int i = 0; i < s.count; ++i) {
        const float ex = px - s.c[i][0], ey = py - s.c[i][1], ez = pz - s.c[i][2];
        v = fminf(v, sqrtf(ex * ex + ey * ey + ez * ez) - s.c[i][3]);
    }
    ls[y + (int64_t)x * ny + (int64_t)z * nx * ny] = v;
}

// cuComputeFrac (WaveletPotentials/cuSolidLevelSet3D.cu:28-34): wet fraction of
// a face of width h between cells with level-set values l and r.
__device__ __forceinline__ float face_frac(float h, float l, float r) {
    const float dif = r - l;
    const float d = sqrtf(fmaxf(h * h - dif * dif, 1e-14f));
    return 1.f - fminf(fmaxf((1.f - (r + l) / d) * 0.5f, 0.f), 1.f);
}

// cuGetAx / cuGetAy / cuGetAz (WaveletPotentials/cuSolidLevelSet3D.cu:36-90) in
// one launch over the (nx+1, ny+1, nz+1) nodes: node (x, y, z) owns x-face
// (x, y, z), y-face (x, y, z) and z-face (x, y, z) where they exist.  Faces on
// the domain boundary see the boundary cell on both sides.
__global__ void frac_kernel(float* __restrict__ ax, float* __restrict__ ay, float* __restrict__ az,
                            const float* __restrict__ ls, int nx, int ny, int nz, float dx,
                            float dy, float dz) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y > ny || x > nx) return;
    const int64_t sx = ny, sz = (int64_t)nx * ny;
    const int xc = x < nx ? x : nx - 1, yc = y < ny ? y : ny - 1, zc = z < nz ? z : nz - 1;
    const int xl = x > 0 ? x - 1 : 0, yl = y > 0 ? y - 1 : 0, zl = z > 0 ? z - 1 : 0;
    if (y < ny && z < nz)
        ax[y + (int64_t)x * ny + (int64_t)z * (nx + 1) * ny] =
            face_frac(dx, ls[y + xl * sx + z * sz], ls[y + xc * sx + z * sz]);
    if (x < nx && z < nz)
        ay[y + (int64_t)x * (ny + 1) + (int64_t)z * nx * (ny + 1)] =
            face_frac(dy, ls[yl + x * sx + z * sz], ls[yc + x * sx + z * sz]);
    if (x < nx && y < ny)
        az[y + (int64_t)x * ny + (int64_t)z * nx * ny] =
            face_frac(dz, ls[y + x * sx + zl * sz], ls[y + x * sx + zc * sz]);
}

// solve_eikonal_nomin (WaveletPotentials/cuFastSweeping.cu:105-146) for equal
// spacing h: the first-order upwind (Godunov) update from the smallest
// neighbour value along each axis, using one, two or three axes.  With equal
// spacings the reference's expanded quadratics reduce to the forms below.
__device__ __forceinline__ float eikonal_update(float a, float b, float c, float h) {
    // sort a <= b <= c
    float t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    float u = a + h;
    if (u > b) {
        const float ab = a - b;
        u = 0.5f * (a + b + sqrtf(fmaxf(2.f * h * h - ab * ab, 0.f)));
        if (u > c) {
            const float ac = a - c, bc = b - c;
            u = (a + b + c + sqrtf(fmaxf(3.f * h * h - ab * ab - ac * ac - bc * bc, 0.f))) *
                (1.f / 3.f);
        }
    }
    return u;
}

// One Jacobi relaxation of fast_sweep_kernel_pos
// (WaveletPotentials/cuFastSweeping.cu:215-248): cells with ls > -1e-12 take the
// eikonal update of their neighbours, solid cells keep their value; neighbours
// outside the grid count as 1e8.  Reads src, writes dst, so the result does
// not depend on thread scheduling (the reference updates in place).
__global__ void eikonal_kernel(float* __restrict__ dst, const float* __restrict__ src, int nx,
                               int ny, int nz, float h) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= ny || x >= nx) return;
    const int64_t sx = ny, sz = (int64_t)nx * ny;
    const int64_t i = y + x * sx + z * sz;
    const float c = src[i];
    if (!(c > -1e-12f)) {
        dst[i] = c;
        return;
    }
    constexpr float kFar = 1e8f;
    const float l = x > 0 ? src[i - sx] : kFar, r = x < nx - 1 ? src[i + sx] : kFar;
    const float d = y > 0 ? src[i - 1] : kFar, u = y < ny - 1 ? src[i + 1] : kFar;
    const float f = z > 0 ? src[i - sz] : kFar, b = z < nz - 1 ? src[i + sz] : kFar;
    dst[i] = eikonal_update(fminf(l, r), fminf(d, u), fminf(f, b), h);
}

int check_grid(int nx, int ny, int nz, const char* who) {
    if (nx < 1 || ny < 1 || nz < 1 || (int64_t)(nx + 1) * (ny + 1) > (int64_t)1 << 30 ||
        nz + 1 > 65535)
        return fail(WHHD_INVALID_ARG, std::string(who) + ": grid extents out of range");
    return WHHD_OK;
}

dim3 cell_grid(int nx, int ny, int nz) { return dim3((ny + 63) / 64, (nx + 3) / 4, nz); }

}  // namespace

int levelset_spheres(float* ls, int nx, int ny, int nz, double dx, double dy, double dz,
                     const double* centers, const double* radii, int count, hipStream_t s) {
    if (!ls || !centers || !radii) return fail(WHHD_INVALID_ARG, "levelset_spheres: null argument");
    if (count < 1 || count > kMaxSpheres)
        return fail(WHHD_INVALID_ARG, "levelset_spheres: count must be in [1, 64]");
    int st = check_grid(nx, ny, nz, "levelset_spheres");
    if (st != WHHD_OK) return st;
    Spheres sp{};
    sp.count = count;
    for (int i = 0; i < count; ++i) {
        for (int k = 0; k < 3; ++k) sp.c[i][k] = (float)centers[3 * i + k];
        sp.c[i][3] = (float)radii[i];
    }
    hipLaunchKernelGGL(levelset_spheres_kernel, cell_grid(nx, ny, nz), dim3(64, 4), 0, s, ls, nx,
                       ny, nz, (float)dx, (float)dy, (float)dz, sp);
    set_last_kernel("levelset_spheres_kernel");
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "levelset_spheres kernel");
}

int frac_from_ls(float* ax, float* ay, float* az, const float* ls, int nx, int ny, int nz,
                 double dx, double dy, double dz, hipStream_t s) {
    if (!ax || !ay || !az || !ls) return fail(WHHD_INVALID_ARG, "frac_from_ls: null argument");
    int st = check_grid(nx, ny, nz, "frac_from_ls");
    if (st != WHHD_OK) return st;
    hipLaunchKernelGGL(frac_kernel, cell_grid(nx + 1, ny + 1, nz + 1), dim3(64, 4), 0, s, ax, ay, az,
                       ls, nx, ny, nz, (float)dx, (float)dy, (float)dz);
    set_last_kernel("frac_kernel");
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "frac kernel");
}

int eikonal_sweep(float* ls, float* scratch, int nx, int ny, int nz, double h, int n_iters,
                  hipStream_t s) {
    if (!ls || !scratch) return fail(WHHD_INVALID_ARG, "eikonal_sweep: null argument");
    if (ls == scratch) return fail(WHHD_INVALID_ARG, "eikonal_sweep: scratch must differ from ls");
    if (n_iters < 0) return fail(WHHD_INVALID_ARG, "eikonal_sweep: n_iters < 0");
    if (!(h > 0.0)) return fail(WHHD_INVALID_ARG, "eikonal_sweep: spacing must be positive");
    int st = check_grid(nx, ny, nz, "eikonal_sweep");
    if (st != WHHD_OK) return st;
    float* a = ls;
    float* b = scratch;
    for (int i = 0; i < n_iters; ++i) {
        hipLaunchKernelGGL(eikonal_kernel, cell_grid(nx, ny, nz), dim3(64, 4), 0, s, b, a, nx, ny,
                           nz, (float)h);
        float* t = a;
        a = b;
        b = t;
    }
    set_last_kernel("eikonal_kernel");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "eikonal kernel");
    if (a != ls) {
        e = hipMemcpyAsync(ls, a, sizeof(float) * (size_t)nx * ny * nz, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "eikonal_sweep copy-back");
    }
    return WHHD_OK;
}

}  // namespace whhd

extern "C" {

WHHD_API int whhd_levelset_spheres(float* ls, int nx, int ny, int nz, double dx, double dy,
                                   double dz, const double* centers, const double* radii,
                                   int count, void* stream) {
    return whhd::levelset_spheres(ls, nx, ny, nz, dx, dy, dz, centers, radii, count,
                                  (hipStream_t)stream);
}

WHHD_API int whhd_frac_from_ls(float* ax, float* ay, float* az, const float* ls, int nx, int ny,
                               int nz, double dx, double dy, double dz, void* stream) {
    return whhd::frac_from_ls(ax, ay, az, ls, nx, ny, nz, dx, dy, dz, (hipStream_t)stream);
}

WHHD_API int whhd_eikonal_sweep(float* ls, float* scratch, int nx, int ny, int nz, double h,
                                int n_iters, void* stream) {
    return whhd::eikonal_sweep(ls, scratch, nx, ny, nz, h, n_iters, (hipStream_t)stream);
}

}  // extern "C"
