// Solid geometry on the cell grid: sphere level sets, face fractions (the
// two-point formula and the supersampled one) and the eikonal relaxation
// (WaveletPotentials/cuSolidLevelSet3D.cu:11-383;
// WaveletPotentials/cuFastSweeping.cu:98-257).
//
// The level set, two-point fractions and eikonal relaxation are streaming
// stencils over y-fastest arrays, HBM-bound: one thread per cell (or face), a
// (64, 4) block over (y, x) and one z-plane per grid row, so consecutive lanes
// touch consecutive floats.  The supersampled fractions are VALU-bound on the
// faces near the interface (16384 samples each) and streaming elsewhere.
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
// spacings the reference's expanded quadratics reduce to the forms below.  The
// three values are ordered by a min/max network and the square roots are the
// hardware's (v_sqrt_f32, within 1 ulp): the relaxation is VALU-bound, and the
// correctly rounded sqrtf sequence and branchy swaps cost twice the rest.
__device__ __forceinline__ float eikonal_update(float a, float b, float c, float h) {
    const float lo = fminf(a, b), hi = fmaxf(a, b);
    const float t = fmaxf(lo, c);
    const float m0 = fminf(lo, c), m1 = fminf(hi, t), m2 = fmaxf(hi, t);  // m0 <= m1 <= m2
    float u = m0 + h;
    if (u > m1) {
        const float d01 = m0 - m1;
        u = 0.5f * (m0 + m1 + __builtin_amdgcn_sqrtf(fmaxf(2.f * h * h - d01 * d01, 0.f)));
        if (u > m2) {
            const float d02 = m0 - m2, d12 = m1 - m2;
            u = (m0 + m1 + m2 +
                 __builtin_amdgcn_sqrtf(fmaxf(3.f * h * h - d01 * d01 - d02 * d02 - d12 * d12, 0.f))) *
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
//
// One thread per (x, y) column walks kEikZ consecutive z cells, carrying the
// z - 1, z and z + 1 values in registers, so every value is loaded once per
// column walk (the in-plane neighbours come from the caches) instead of up to
// three times.
constexpr int kEikZ = 16;  // 4, 8 and 16 time alike at 192^3 (~20 us); 32: 30 us

__global__ void __launch_bounds__(256) eikonal_kernel(float* __restrict__ dst, const float* __restrict__ src,
                                                      int nx, int ny, int nz, float h) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z0 = blockIdx.z * kEikZ;
    if (y >= ny || x >= nx) return;
    const int64_t sx = ny, sz = (int64_t)nx * ny;
    constexpr float kFar = 1e8f;
    int64_t i = y + x * sx + z0 * sz;
    float below = z0 > 0 ? src[i - sz] : kFar;
    float c = src[i];
    const int zend = min(z0 + kEikZ, nz);
    for (int z = z0; z < zend; ++z, i += sz) {
        const float above = z < nz - 1 ? src[i + sz] : kFar;
        // neighbours through clamped (always valid) indices, replaced by kFar
        // outside the grid: no branch per load
        const float l0 = src[i - (x > 0 ? sx : 0)], r0 = src[i + (x < nx - 1 ? sx : 0)];
        const float d0 = src[i - (y > 0 ? 1 : 0)], u0 = src[i + (y < ny - 1 ? 1 : 0)];
        const float l = x > 0 ? l0 : kFar, r = x < nx - 1 ? r0 : kFar;
        const float d = y > 0 ? d0 : kFar, u = y < ny - 1 ? u0 : kFar;
        const float upd = eikonal_update(fminf(l, r), fminf(d, u), fminf(below, above), h);
        dst[i] = c > -1e-12f ? upd : c;
        below = c;
        c = above;
    }
}

// ---- supersampled face fractions ------------------------------------------
// cuGetAx/Ay/Az_divide (WaveletPotentials/cuSolidLevelSet3D.cu:164-317), the
// kernels of the static GetFrac (:359-372).  Grid coordinates are cell-centred
// (world position / spacing - 1/2; the reference passes c_min = 0 and
// dx_c = dx, so they do not depend on the spacing).  A face of the plane
// normal to axis A with lower corner g0 (grid coordinates) is 1 when the
// quadratic B-spline interpolant of the level set at its centre exceeds h (the
// spacing along A), 0 when it is <= -h, and otherwise the share of its 128 x 128
// midpoint sub-samples where the trilinear interpolant is positive.

constexpr int kSub = 128;
constexpr float kClampL = 1e-5f;  // CLAMP_L of the reference's FP32 build

// cuGetIxFx (WaveletPotentials/Interpolation.cuh:16-29)
__device__ __forceinline__ void ixfx(int& ix, float& fx, float x, int n) {
    x = fmaxf(kClampL, x);
    ix = (int)x;
    if (ix >= n - 1) {
        ix = n - 2;
        fx = 1.f - kClampL;
    } else {
        fx = x - (float)ix;
    }
}

// InterpolateQuadratic3D (WaveletPotentials/Interpolation.cuh:119-157)
__device__ float interp_quadratic(const float* __restrict__ v, float x, float y, float z, int nx,
                                  int ny, int nz) {
    const int ix = (int)floorf(x - 0.5f), iy = (int)floorf(y - 0.5f), iz = (int)floorf(z - 0.5f);
    const float fx = x - ix, fy = y - iy, fz = z - iz;
    auto sq = [](float a) { return a * a; };
    const float wx[3] = {0.5f * sq(1.5f - fx), 0.75f - sq(fx - 1.f), 0.5f * sq(fx - 0.5f)};
    const float wy[3] = {0.5f * sq(1.5f - fy), 0.75f - sq(fy - 1.f), 0.5f * sq(fy - 0.5f)};
    const float wz[3] = {0.5f * sq(1.5f - fz), 0.75f - sq(fz - 1.f), 0.5f * sq(fz - 0.5f)};
    float r = 0.f;
    for (int k = 0; k < 3; ++k) {
        const int64_t Z = min(max(iz + k, 0), nz - 1);
        for (int j = 0; j < 3; ++j) {
            const int64_t X = min(max(ix + j, 0), nx - 1);
            for (int i = 0; i < 3; ++i) {
                const int64_t Y = min(max(iy + i, 0), ny - 1);
                r += wx[j] * wy[i] * wz[k] * v[Y + X * ny + Z * nx * ny];
            }
        }
    }
    return r;
}

// One launch per face array (A: normal axis).  A workgroup of 256 threads
// owns 256 consecutive faces: each thread classifies its face by the band
// test and writes 0 or 1 directly; faces in the band go to an LDS list, and
// the whole workgroup then counts the samples of each listed face.  For a
// listed face, thread t takes in-plane column i = t % 128 along axis P and the
// 64 rows k of half t / 128 along axis Q.  The trilinear interpolant on the
// face plane at fixed column is linear in the Q coordinate between Q nodes,
// so each thread folds the A and P weights into the values of three
// consecutive Q nodes (12 loads) and evaluates each sample with one lerp.
// The cell index along Q changes at most once inside a half (it spans less
// than half a cell), which the three nodes cover.
template <int A>
__global__ __launch_bounds__(256) void frac_divide_kernel(float* __restrict__ out,
                                                          const float* __restrict__ ls, int nx,
                                                          int ny, int nz, float h,
                                                          int64_t nfaces) {
    constexpr int P = A == 0 ? 1 : 0;  // in-plane axis of the thread's column
    constexpr int Q = A == 2 ? 1 : 2;  // in-plane axis walked per sample
    __shared__ int list[256];
    __shared__ int nlist;
    __shared__ int wsum[4];
    const int t = threadIdx.x;
    if (t == 0) nlist = 0;
    __syncthreads();
    const int ex = nx + (A == 0), ey = ny + (A == 1);
    const int n3[3] = {nx, ny, nz};
    const int64_t base = (int64_t)blockIdx.x * 256;
    auto decode = [&](int64_t f, int g[3]) {
        const int64_t r = f / ey;
        g[1] = (int)(f - r * ey);
        g[0] = (int)(r % ex);
        g[2] = (int)(r / ex);
    };
    if (base + t < nfaces) {
        int g[3];
        decode(base + t, g);
        float c[3];
        for (int a = 0; a < 3; ++a) c[a] = a == A ? g[a] - 0.5f : (float)g[a];
        const float lc = interp_quadratic(ls, c[0], c[1], c[2], nx, ny, nz);
        if (lc > h)
            out[base + t] = 1.f;
        else if (!(lc > -h))
            out[base + t] = 0.f;
        else
            list[atomicAdd(&nlist, 1)] = t;
    }
    __syncthreads();
    const int cnt = nlist;
    const int i = t & (kSub - 1), half = t >> 7;
    const int64_t sx = ny, sz = (int64_t)nx * ny;
    auto at = [&](const int id[3]) { return ls[id[1] + id[0] * sx + id[2] * sz]; };
    for (int e = 0; e < cnt; ++e) {
        const int64_t f = base + list[e];
        int g[3];
        decode(f, g);
        int iA, iP, iQ0;
        float fA, fP, fQ;
        ixfx(iA, fA, g[A] - 0.5f, n3[A]);
        ixfx(iP, fP, g[P] - 0.5f + (i + 0.5f) * (1.f / kSub), n3[P]);
        const int k0 = half * (kSub / 2);
        ixfx(iQ0, fQ, g[Q] - 0.5f + (k0 + 0.5f) * (1.f / kSub), n3[Q]);
        const float wA[2] = {1.f - fA, fA}, wP[2] = {1.f - fP, fP};
        float col[3];
        for (int j = 0; j < 3; ++j) {
            float acc = 0.f;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    int id[3];
                    id[A] = iA + a;
                    id[P] = iP + b;
                    id[Q] = min(iQ0 + j, n3[Q] - 1);
                    acc += at(id) * wA[a] * wP[b];
                }
            col[j] = acc;
        }
        int pos = 0;
        for (int k = k0; k < k0 + kSub / 2; ++k) {
            int iQ;
            ixfx(iQ, fQ, g[Q] - 0.5f + (k + 0.5f) * (1.f / kSub), n3[Q]);
            const bool up = iQ != iQ0;
            const float lo = up ? col[1] : col[0], hi = up ? col[2] : col[1];
            pos += lo * (1.f - fQ) + hi * fQ > 0.f;
        }
        for (int o = 32; o > 0; o >>= 1) pos += __shfl_xor(pos, o);
        if ((t & 63) == 0) wsum[t >> 6] = pos;
        __syncthreads();
        if (t == 0) {
            const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
            out[f] = fminf(fmaxf(tot * (1.f / (kSub * kSub)), 0.f), 1.f);
        }
        __syncthreads();
    }
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

int frac_divide(float* ax, float* ay, float* az, const float* ls, int nx, int ny, int nz,
                double dx, double dy, double dz, hipStream_t s) {
    if (!ax || !ay || !az || !ls) return fail(WHHD_INVALID_ARG, "frac_divide: null argument");
    int st = check_grid(nx, ny, nz, "frac_divide");
    if (st != WHHD_OK) return st;
    if (nx < 2 || ny < 2 || nz < 2)
        return fail(WHHD_INVALID_ARG, "frac_divide: needs at least 2 cells along every axis");
    const int64_t X = nx, Y = ny, Z = nz;
    const int64_t n[3] = {(X + 1) * Y * Z, X * (Y + 1) * Z, X * Y * (Z + 1)};
    const dim3 b(256);
    hipLaunchKernelGGL(frac_divide_kernel<0>, dim3((unsigned)((n[0] + 255) / 256)), b, 0, s, ax, ls,
                       nx, ny, nz, (float)dx, n[0]);
    hipLaunchKernelGGL(frac_divide_kernel<1>, dim3((unsigned)((n[1] + 255) / 256)), b, 0, s, ay, ls,
                       nx, ny, nz, (float)dy, n[1]);
    hipLaunchKernelGGL(frac_divide_kernel<2>, dim3((unsigned)((n[2] + 255) / 256)), b, 0, s, az, ls,
                       nx, ny, nz, (float)dz, n[2]);
    set_last_kernel("frac_divide_kernel");
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "frac_divide kernel");
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
        hipLaunchKernelGGL(eikonal_kernel, dim3((ny + 63) / 64, (nx + 3) / 4, (nz + kEikZ - 1) / kEikZ),
                           dim3(64, 4), 0, s, b, a, nx, ny, nz, (float)h);
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

WHHD_API int whhd_frac_divide(float* ax, float* ay, float* az, const float* ls, int nx, int ny,
                              int nz, double dx, double dy, double dz, void* stream) {
    return whhd::frac_divide(ax, ay, az, ls, nx, ny, nz, dx, dy, dz, (hipStream_t)stream);
}

WHHD_API int whhd_eikonal_sweep(float* ls, float* scratch, int nx, int ny, int nz, double h,
                                int n_iters, void* stream) {
    return whhd::eikonal_sweep(ls, scratch, nx, ny, nz, h, n_iters, (hipStream_t)stream);
}

}  // extern "C"
