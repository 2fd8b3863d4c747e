// Face/cell field operators around the pressure solve and device reductions:
//   gradient with fractions  CuGradient::Gradient3D_Frac_P
//                            (WaveletPotentials/cuGradient.cu:1841-1951, 2514-2545)
//   divergence of a*u        CuConvergence::GetB_3d_frac_vel
//                            (WaveletPotentials/cuConvergence.cu:40-124, 435-447)
//   u += s*v, u = l*r        CudaMatAdd / CudaMatMul (WaveletPotentials/cudaMath.cu:110-168)
//   sums and max |x|         CudaArraySum / CudaArrayNormSum / CudaArrayAbsSum /
//                            CudaFindMaxValue (WaveletPotentials/cudaMath.cu:35-108, 1230-1280)
// plus one fused pass, u <- a * (u - grad p), for the two element-wise passes
// the reference runs after the gradient (WaveletPotentials/TestLocalProjection.cpp:690-701).
//
// Stencils run one thread per node (x <= nx, y <= ny, z <= nz) of a (64, 4)
// block over (y, x): node (x, y, z) owns x-face (x, y, z), y-face (x, y, z) and
// z-face (x, y, z) where they exist, so each face is written once.  The
// reductions accumulate in FP64 in a grid-stride pass and one final block.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "whhd_internal.h"

namespace whhd {
namespace {

constexpr float kFracEps = 1e-12f;  // eps<DTYPE>, WaveletPotentials/Utils.h:52-53

struct FaceGrid {
    int nx, ny, nz;
    float hx, hy, hz;   // cell spacings
    float ex, ey, ez;   // distance to the wall value: h/2 for 'd', h for 'z'; <0: closed wall
};

__device__ __forceinline__ int64_t ixf(int x, int y, int z, int nx, int ny) {  // x-face
    return y + (int64_t)x * ny + (int64_t)z * (nx + 1) * ny;
}
__device__ __forceinline__ int64_t iyf(int x, int y, int z, int nx, int ny) {  // y-face
    return y + (int64_t)x * (ny + 1) + (int64_t)z * nx * (ny + 1);
}
__device__ __forceinline__ int64_t izf(int x, int y, int z, int nx, int ny) {  // z-face
    return y + (int64_t)x * ny + (int64_t)z * nx * ny;
}
__device__ __forceinline__ int64_t icell(int x, int y, int z, int nx, int ny) {
    return y + (int64_t)x * ny + (int64_t)z * nx * ny;
}

// Gradient component on the face between cells (lo) and (hi) along one axis
// with c cells: interior faces (p_hi - p_lo)/h; wall faces use the wall
// distance e (zero wall value) or are zero when closed (e < 0).  Faces whose
// fraction is not above eps get 0.  Both loads are issued from the (clamped,
// always valid) indices and the case is selected afterwards: a branch per
// load would make each face wait out its own memory latency.
__device__ __forceinline__ float face_grad(const float* __restrict__ p, float a, int k, int c,
                                          int64_t i_hi, int64_t i_lo, float h, float e) {
    const float ph = p[i_hi], pl = p[i_lo];
    const float inner = (ph - pl) / h;
    const float wall = k == 0 ? ph / e : -pl / e;
    const float g = (k > 0 && k < c) ? inner : (e < 0.f ? 0.f : wall);
    return a > kFracEps ? g : 0.f;
}

// FUSE = false: out = grad p (Gradient3D_Frac_P); FUSE = true: u <- a * (u - grad p),
// where p = p + p2 when p2 is given (a potential held as a sum of two floats,
// whose gradients are taken separately so the low part is not rounded away).
template <bool FUSE>
__global__ void gradient_kernel(float* __restrict__ ux, float* __restrict__ uy,
                                float* __restrict__ uz, const float* __restrict__ p,
                                const float* __restrict__ ax, const float* __restrict__ ay,
                                const float* __restrict__ az, FaceGrid g,
                                const float* __restrict__ p2 = nullptr) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y > g.ny || x > g.nx) return;
    const int nx = g.nx, ny = g.ny;
    if (y < ny && z < g.nz) {
        const int64_t f = ixf(x, y, z, nx, ny);
        const float a = ax[f];
        const int64_t hi = icell(x < nx ? x : nx - 1, y, z, nx, ny), lo = icell(x > 0 ? x - 1 : 0, y, z, nx, ny);
        float gr = face_grad(p, a, x, nx, hi, lo, g.hx, g.ex);
        if (FUSE && p2) gr += face_grad(p2, a, x, nx, hi, lo, g.hx, g.ex);
        ux[f] = FUSE ? a * (ux[f] - gr) : gr;
    }
    if (x < nx && z < g.nz) {
        const int64_t f = iyf(x, y, z, nx, ny);
        const float a = ay[f];
        const int64_t hi = icell(x, y < ny ? y : ny - 1, z, nx, ny), lo = icell(x, y > 0 ? y - 1 : 0, z, nx, ny);
        float gr = face_grad(p, a, y, ny, hi, lo, g.hy, g.ey);
        if (FUSE && p2) gr += face_grad(p2, a, y, ny, hi, lo, g.hy, g.ey);
        uy[f] = FUSE ? a * (uy[f] - gr) : gr;
    }
    if (x < nx && y < ny) {
        const int64_t f = izf(x, y, z, nx, ny);
        const float a = az[f];
        const int64_t hi = icell(x, y, z < g.nz ? z : g.nz - 1, nx, ny), lo = icell(x, y, z > 0 ? z - 1 : 0, nx, ny);
        float gr = face_grad(p, a, z, g.nz, hi, lo, g.hz, g.ez);
        if (FUSE && p2) gr += face_grad(p2, a, z, g.nz, hi, lo, g.hz, g.ez);
        uz[f] = FUSE ? a * (uz[f] - gr) : gr;
    }
}

// b = -div(a u) over every face of the cell (cuGetB_vel_frac_3d(_n)).
__global__ void divergence_kernel(float* __restrict__ b, const float* __restrict__ ux,
                                  const float* __restrict__ uy, const float* __restrict__ uz,
                                  const float* __restrict__ ax, const float* __restrict__ ay,
                                  const float* __restrict__ az, int nx, int ny, int nz, float ihx,
                                  float ihy, float ihz) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= ny || x >= nx) return;
    const int64_t fx = ixf(x, y, z, nx, ny), fy = iyf(x, y, z, nx, ny), fz = izf(x, y, z, nx, ny);
    const int64_t sx = ny, sz = (int64_t)nx * ny;
    const float v = (ux[fx + sx] * ax[fx + sx] - ux[fx] * ax[fx]) * ihx +
                    (uy[fy + 1] * ay[fy + 1] - uy[fy] * ay[fy]) * ihy +
                    (uz[fz + sz] * az[fz + sz] - uz[fz] * az[fz]) * ihz;
    b[icell(x, y, z, nx, ny)] = -v;
}

__global__ void axpy_kernel(float* __restrict__ dst, const float* __restrict__ src, float s,
                            int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] += src[i] * s;
}

__global__ void mul_kernel(float* __restrict__ dst, const float* __restrict__ l,
                           const float* __restrict__ r, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = l[i] * r[i];
}

constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double red_op(double a, double b, int op) {
    return op == WHHD_REDUCE_MAX_ABS ? fmax(a, b) : a + b;
}

__global__ void reduce_partial_kernel(double* __restrict__ part, const float* __restrict__ x,
                                      int64_t n, int op) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = x[i];
        const double t = op == WHHD_REDUCE_SUM ? v : op == WHHD_REDUCE_SUM_SQ ? v * v : fabs(v);
        acc = red_op(acc, t, op);
    }
    for (int o = 32; o > 0; o >>= 1) acc = red_op(acc, __shfl_xor(acc, o), op);
    __shared__ double l[4];
    if ((threadIdx.x & 63) == 0) l[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red_op(red_op(l[0], l[1], op), red_op(l[2], l[3], op), op);
}

__global__ void reduce_final_kernel(double* __restrict__ out, const double* __restrict__ part,
                                    int nparts, int op) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc = red_op(acc, part[i], op);
    for (int o = 32; o > 0; o >>= 1) acc = red_op(acc, __shfl_xor(acc, o), op);
    __shared__ double l[4];
    if ((threadIdx.x & 63) == 0) l[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = red_op(red_op(l[0], l[1], op), red_op(l[2], l[3], op), op);
}

// Device scratch of whhd_reduce_noalloc: kRedBlocks FP64 partials and the
// result (8 KiB + 8 B per GPU, held until the process ends).  Each GPU holds its
// own copy at its own address; host code uses only the address that
// hipGetSymbolAddress returns for the current device, never the symbol itself.
static __device__ __align__(256) double scratch[kRedBlocks + 1];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// The current device's address of `scratch`, looked up once per device.
int reduce_static_scratch(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "reduce_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxScratchDevices)
        return fail(WHHD_INVALID_ARG, "reduce_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "reduce_noalloc: hipGetSymbolAddress");
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return WHHD_OK;
}

int check_face_args(int nx, int ny, int nz, double dx, double dy, double dz, char bc,
                    const char* who) {
    if (nx < 1 || ny < 1 || nz < 1 || nz + 1 > 65535 ||
        (int64_t)(nx + 1) * (ny + 1) * (nz + 1) > (1LL << 31))
        return fail(WHHD_INVALID_ARG, std::string(who) + ": grid extents out of range");
    if (!(dx > 0 && dy > 0 && dz > 0))
        return fail(WHHD_INVALID_ARG, std::string(who) + ": spacing must be positive");
    if (bc != 'n' && bc != 'd' && bc != 'z')
        return fail(WHHD_INVALID_ARG, std::string(who) + ": bc must be 'n', 'd' or 'z'");
    return WHHD_OK;
}

// Wall distances: 'n' closed, 'd' the wall value sits on the face (h/2,
// the reference's zero exterior spacing), 'z' one cell outside (h).
FaceGrid face_grid(int nx, int ny, int nz, double dx, double dy, double dz, char bc) {
    FaceGrid g{nx, ny, nz, (float)dx, (float)dy, (float)dz, -1.f, -1.f, -1.f};
    if (bc == 'd') {
        g.ex = 0.5f * (float)dx;
        g.ey = 0.5f * (float)dy;
        g.ez = 0.5f * (float)dz;
    } else if (bc == 'z') {
        g.ex = (float)dx;
        g.ey = (float)dy;
        g.ez = (float)dz;
    }
    return g;
}

hipError_t done(const char* name) {
    set_last_kernel(name);
    return hipGetLastError();
}

dim3 node_grid(int nx, int ny, int nz) { return dim3((ny + 1 + 63) / 64, (nx + 1 + 3) / 4, nz + 1); }

}  // namespace
}  // namespace whhd

extern "C" {

WHHD_API int whhd_gradient3d_frac(float* xv, float* yv, float* zv, const float* p, const float* ax,
                                  const float* ay, const float* az, int nx, int ny, int nz,
                                  double dx, double dy, double dz, char bc, void* stream) {
    using namespace whhd;
    if (!xv || !yv || !zv || !p || !ax || !ay || !az)
        return fail(WHHD_INVALID_ARG, "gradient3d_frac: null argument");
    int st = check_face_args(nx, ny, nz, dx, dy, dz, bc, "gradient3d_frac");
    if (st != WHHD_OK) return st;
    hipLaunchKernelGGL(gradient_kernel<false>, node_grid(nx, ny, nz), dim3(64, 4), 0,
                       (hipStream_t)stream, xv, yv, zv, p, ax, ay, az,
                       face_grid(nx, ny, nz, dx, dy, dz, bc));
    hipError_t e = done("gradient_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "gradient kernel");
}

WHHD_API int whhd_subtract_gradient_frac(float* xv, float* yv, float* zv, const float* p,
                                         const float* ax, const float* ay, const float* az, int nx,
                                         int ny, int nz, double dx, double dy, double dz, char bc,
                                         void* stream) {
    using namespace whhd;
    if (!xv || !yv || !zv || !p || !ax || !ay || !az)
        return fail(WHHD_INVALID_ARG, "subtract_gradient_frac: null argument");
    int st = check_face_args(nx, ny, nz, dx, dy, dz, bc, "subtract_gradient_frac");
    if (st != WHHD_OK) return st;
    hipLaunchKernelGGL(gradient_kernel<true>, node_grid(nx, ny, nz), dim3(64, 4), 0,
                       (hipStream_t)stream, xv, yv, zv, p, ax, ay, az,
                       face_grid(nx, ny, nz, dx, dy, dz, bc));
    hipError_t e = done("subtract_gradient_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "subtract gradient kernel");
}

WHHD_API int whhd_subtract_gradient_frac2(float* xv, float* yv, float* zv, const float* p_hi,
                                          const float* p_lo, const float* ax, const float* ay,
                                          const float* az, int nx, int ny, int nz, double dx,
                                          double dy, double dz, char bc, void* stream) {
    using namespace whhd;
    if (!xv || !yv || !zv || !p_hi || !p_lo || !ax || !ay || !az)
        return fail(WHHD_INVALID_ARG, "subtract_gradient_frac2: null argument");
    int st = check_face_args(nx, ny, nz, dx, dy, dz, bc, "subtract_gradient_frac2");
    if (st != WHHD_OK) return st;
    hipLaunchKernelGGL(gradient_kernel<true>, node_grid(nx, ny, nz), dim3(64, 4), 0,
                       (hipStream_t)stream, xv, yv, zv, p_hi, ax, ay, az,
                       face_grid(nx, ny, nz, dx, dy, dz, bc), p_lo);
    hipError_t e = done("subtract_gradient_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "subtract gradient kernel");
}

WHHD_API int whhd_divergence3d_frac(float* b, const float* xv, const float* yv, const float* zv,
                                    const float* ax, const float* ay, const float* az, int nx,
                                    int ny, int nz, double dx, double dy, double dz, void* stream) {
    using namespace whhd;
    if (!b || !xv || !yv || !zv || !ax || !ay || !az)
        return fail(WHHD_INVALID_ARG, "divergence3d_frac: null argument");
    int st = check_face_args(nx, ny, nz, dx, dy, dz, 'n', "divergence3d_frac");
    if (st != WHHD_OK) return st;
    hipLaunchKernelGGL(divergence_kernel, dim3((ny + 63) / 64, (nx + 3) / 4, nz), dim3(64, 4), 0,
                       (hipStream_t)stream, b, xv, yv, zv, ax, ay, az, nx, ny, nz,
                       (float)(1.0 / dx), (float)(1.0 / dy), (float)(1.0 / dz));
    hipError_t e = done("divergence_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "divergence kernel");
}

WHHD_API int whhd_axpy(float* dst, const float* src, long long n, float scale, void* stream) {
    using namespace whhd;
    if (!dst || !src || n < 0) return fail(WHHD_INVALID_ARG, "axpy: bad argument");
    if (n == 0) return WHHD_OK;
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(axpy_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, dst, src,
                       scale, (int64_t)n);
    hipError_t e = done("axpy_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "axpy kernel");
}

WHHD_API int whhd_mul(float* dst, const float* l, const float* r, long long n, void* stream) {
    using namespace whhd;
    if (!dst || !l || !r || n < 0) return fail(WHHD_INVALID_ARG, "mul: bad argument");
    if (n == 0) return WHHD_OK;
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(mul_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, dst, l, r,
                       (int64_t)n);
    hipError_t e = done("mul_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "mul kernel");
}

WHHD_API int whhd_reduce(const float* x, long long n, int op, double* scratch, double* out,
                         void* stream) {
    using namespace whhd;
    if (!x || !scratch || !out || n < 0 || op < WHHD_REDUCE_SUM || op > WHHD_REDUCE_MAX_ABS)
        return fail(WHHD_INVALID_ARG, "reduce: bad argument");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(reduce_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, scratch, x,
                       (int64_t)n, op);
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(256), 0, s, scratch + kRedBlocks, scratch,
                       kRedBlocks, op);
    hipError_t e = done("reduce_kernels");
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, scratch + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "reduce");
}

WHHD_API int whhd_reduce_scratch_doubles(void) { return whhd::kRedBlocks + 1; }

WHHD_API int whhd_reduce_noalloc(const float* x, long long n, int op, double* out, void* stream) {
    using namespace whhd;
    if (!x || !out || n < 0 || op < WHHD_REDUCE_SUM || op > WHHD_REDUCE_MAX_ABS)
        return fail(WHHD_INVALID_ARG, "reduce_noalloc: bad argument");
    double* sc = nullptr;
    const int st = reduce_static_scratch(&sc);
    if (st != WHHD_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    // every partial slot is written by its block and the result slot by the
    // final kernel, so nothing left by an earlier call is read
    hipLaunchKernelGGL(reduce_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, sc, x,
                       (int64_t)n, op);
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(256), 0, s, sc + kRedBlocks, sc,
                       kRedBlocks, op);
    hipError_t e = done("reduce_kernels");
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, sc + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "reduce_noalloc");
}

}  // extern "C"
