// Fractional-area multigrid for the pressure Poisson problem with solids:
// CuMultigridFrac3D (WaveletPotentials/cuMultigridFrac3D.cu:9-327) with the
// weighted Jacobi smoother, residual, restriction and prolongation of
// WaveletPotentials/cuWeightedJacobi3D.cu:620-1722, cuRestrict3D.cu:524-1140 and
// cuProlongate3D.cu:353-616.
//
// Every level is a y-fastest cell grid with face fractions ax (nx+1, ny, nz),
// ay (nx, ny+1, nz), az (nx, ny, nz+1) and per-axis cell spacings.  All
// kernels are streaming stencils (HBM-bound): one thread per output cell, a
// (64, 4) block over (y, x) and one z-plane per grid row.  The handle owns
// the hierarchy; solves enqueue on the caller's stream without host
// synchronization (the residual queries read back two FP64 sums).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "whhd_internal.h"

namespace whhd {
namespace {

struct Dim3 {
    int x, y, z;
};

int half_up(int n) { return (n & 1) ? (n >> 1) + 1 : n >> 1; }

dim3 grid_of(int nx, int ny, int nz) { return dim3((ny + 63) / 64, (nx + 3) / 4, nz); }

// Coefficient of one face, cuWjFrac3D_n / _z: A * 2 / ((h_nbr + h) * h).  The
// spacing factor 2 / ((h_nbr + h) h) of every face position is tabulated per
// level (face_factor_kernel, FP64 then rounded), so a coefficient is one
// multiply: the smoothing kernels were bound by the twelve IEEE divisions a
// cell's faces cost otherwise.  Table layout per level: y- (ny), y+ (ny),
// x- (nx), x+ (nx), z- (nz), z+ (nz).
__global__ void face_factor_kernel(float* __restrict__ kf, const float* __restrict__ hx,
                                   const float* __restrict__ hy, const float* __restrict__ hz, int nx,
                                   int ny, int nz, float ex, float ey, float ez) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * (nx + ny + nz)) return;
    const float* hs;
    int n, j, up;
    float e;
    if (i < 2 * ny) { hs = hy; n = ny; j = i % ny; up = i >= ny; e = ey; }
    else if (i < 2 * (ny + nx)) { hs = hx; n = nx; j = (i - 2 * ny) % nx; up = i >= 2 * ny + nx; e = ex; }
    else { hs = hz; n = nz; j = (i - 2 * (ny + nx)) % nz; up = i >= 2 * (ny + nx) + nz; e = ez; }
    const double h = hs[j];
    const double hn = up ? (j < n - 1 ? hs[j + 1] : e) : (j > 0 ? hs[j - 1] : e);
    kf[i] = (float)(2.0 / ((hn + h) * h));
}

struct Level {
    Dim3 d;
    float *f = nullptr, *v = nullptr, *t = nullptr;  // rhs, iterate, Jacobi partner
    float *ax = nullptr, *ay = nullptr, *az = nullptr;
    float *hx = nullptr, *hy = nullptr, *hz = nullptr;
    float* kf = nullptr;  // face spacing factors (face_factor_kernel)
    __half *axh = nullptr, *ayh = nullptr, *azh = nullptr;  // finest level: FP16 copies
};

struct LevelArgs {
    const float *ax, *ay, *az, *hx, *hy, *hz;
    int nx, ny, nz;
    float ex, ey, ez;  // exterior spacing at the walls ('z')
    const float* kf;   // face spacing factors
    // FP16 copies of the fractions (the finest level inside the tolerance
    // solve's V-cycles), or null
    const __half *axh, *ayh, *azh;
};

// Fraction j of a face array, from the FP16 copy when H16.
template <bool H16>
__device__ __forceinline__ float frac_at(const float* f32, const __half* f16, int64_t j) {
    if constexpr (H16) return __half2float(f16[j]);
    else return f32[j];
}

// Spacing factors of the six faces of cell (x, y, z), in stencil()'s order.
struct FaceK {
    const float* kf;
    int nx, ny, nz;
    __device__ float ym(int y) const { return kf[y]; }
    __device__ float yp(int y) const { return kf[ny + y]; }
    __device__ float xm(int x) const { return kf[2 * ny + x]; }
    __device__ float xp(int x) const { return kf[2 * ny + nx + x]; }
    __device__ float zm(int z) const { return kf[2 * (ny + nx) + z]; }
    __device__ float zp(int z) const { return kf[2 * (ny + nx) + nz + z]; }
};

// Diagonal and off-diagonal sum of cell (x, y, z) applied to the iterate,
// read through src(j, dy, dx, dz) (a neighbour's index and its offset from the
// cell), accumulated in T; the face coefficients are FP32 (the
// operator every level's kernels share).  Z: walls carry a Dirichlet
// coefficient with the exterior spacing; otherwise closed.
template <bool Z, class T, class Src, bool H16 = false>
__device__ __forceinline__ void stencil(const LevelArgs& a, Src src, int x, int y, int z, int64_t i,
                                        T& diag, T& off) {
    const int64_t sx = a.ny, sz = (int64_t)a.nx * a.ny;
    const FaceK k{a.kf, a.nx, a.ny, a.nz};
    // Every load is issued from a valid index whether or not the cell has that
    // face or neighbour (no branch per load); a face the cell does not have
    // enters with coefficient 0 and a neighbour outside the grid with value 0,
    // which leaves both sums exactly as if the term were skipped.
    const int64_t iay = y + (int64_t)x * (a.ny + 1) + (int64_t)z * a.nx * (a.ny + 1);
    const int64_t iax = y + (int64_t)x * a.ny + (int64_t)z * (a.nx + 1) * a.ny;
    const bool ym = y > 0, yp = y < a.ny - 1, xm = x > 0, xp = x < a.nx - 1, zm = z > 0, zp = z < a.nz - 1;
    const float f0 = frac_at<H16>(a.ay, a.ayh, iay), f1 = frac_at<H16>(a.ay, a.ayh, iay + 1),
                f2 = frac_at<H16>(a.ax, a.axh, iax), f3 = frac_at<H16>(a.ax, a.axh, iax + a.ny),
                f4 = frac_at<H16>(a.az, a.azh, i), f5 = frac_at<H16>(a.az, a.azh, i + sz);
    const T n0 = src(ym ? i - 1 : i, ym ? -1 : 0, 0, 0), n1 = src(yp ? i + 1 : i, yp ? 1 : 0, 0, 0);
    const T n2 = src(xm ? i - sx : i, 0, xm ? -1 : 0, 0), n3 = src(xp ? i + sx : i, 0, xp ? 1 : 0, 0);
    const T n4 = src(zm ? i - sz : i, 0, 0, zm ? -1 : 0), n5 = src(zp ? i + sz : i, 0, 0, zp ? 1 : 0);
    const float k0 = k.ym(y), k1 = k.yp(y), k2 = k.xm(x), k3 = k.xp(x), k4 = k.zm(z), k5 = k.zp(z);
    const T c0 = (ym || Z) ? T(f0 * k0) : T(0), c1 = (yp || Z) ? T(f1 * k1) : T(0);
    const T c2 = (xm || Z) ? T(f2 * k2) : T(0), c3 = (xp || Z) ? T(f3 * k3) : T(0);
    const T c4 = (zm || Z) ? T(f4 * k4) : T(0), c5 = (zp || Z) ? T(f5 * k5) : T(0);
    diag = T(0);
    diag += c0;
    diag += c1;
    diag += c2;
    diag += c3;
    diag += c4;
    diag += c5;
    off = T(0);
    off -= c0 * (ym ? n0 : T(0));
    off -= c1 * (yp ? n1 : T(0));
    off -= c2 * (xm ? n2 : T(0));
    off -= c3 * (xp ? n3 : T(0));
    off -= c4 * (zm ? n4 : T(0));
    off -= c5 * (zp ? n5 : T(0));
}

// One weighted Jacobi sweep dst <- src (cuWjFrac3D_n / _z); cells whose
// faces are all closed get 0.  ZERO: src is zero (the first sweep of a V-cycle
// on a correction) and is not read; the arithmetic is the same with zeros.
// PRO: the sweep's input is src plus the prolonged coarse correction vc
// (SolveAddP1RN, cuProlongate3D.cu:594-616: every cell adds its parent), formed
// per read as prolong_add_kernel would store it, so the post-smoothing's first
// sweep needs no separate prolongation pass.
template <bool Z, bool ZERO = false, bool PRO = false, bool H16 = false>
__global__ void jacobi_kernel(float* __restrict__ dst, const float* __restrict__ src,
                              const float* __restrict__ f, LevelArgs a, float w,
                              const float* __restrict__ vc = nullptr, int cx = 0, int cy = 0) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= a.ny || x >= a.nx) return;
    const int64_t i = y + (int64_t)x * a.ny + (int64_t)z * a.nx * a.ny;
    const int64_t cs = (int64_t)cx * cy;
    auto in = [&](int64_t j, int dy, int dx, int dz) {
        if constexpr (ZERO) return 0.f;
        else if constexpr (PRO)
            return src[j] + vc[((y + dy) >> 1) + (int64_t)((x + dx) >> 1) * cy + (int64_t)((z + dz) >> 1) * cs];
        else return src[j];
    };
    float diag, off;
    stencil<Z, float, decltype(in), H16>(a, in, x, y, z, i, diag, off);
    const float v = in(i, 0, 0, 0);
    dst[i] = diag > 1e-12f ? (1.f - w) * v + w * (f[i] - off) / diag : 0.f;
}

// Coarsest-level sweeps in FP64 for at most 256 cells: one cell per thread,
// its coefficients in registers, the iterate in LDS.  All threads of the
// workgroup must call it.
template <bool Z>
__device__ void coarsest_small(float* __restrict__ v, const float* __restrict__ f, const LevelArgs& a,
                               int sweeps, double w) {
    const int64_t n = (int64_t)a.nx * a.ny * a.nz;
    __shared__ double S[2][256];
    const int i = threadIdx.x;
    double cf[6] = {0, 0, 0, 0, 0, 0}, diag = 0.0, fi = 0.0;
    int nb[6] = {-1, -1, -1, -1, -1, -1};
    if (i < n) {
        const int y = i % a.ny, x = (i / a.ny) % a.nx, z = i / (a.ny * a.nx);
        const int sx = a.ny, sz = a.nx * a.ny;
        const int iay = y + x * (a.ny + 1) + z * a.nx * (a.ny + 1);
        const int iax = y + x * a.ny + z * (a.nx + 1) * a.ny;
        // faces in the order of stencil(): y-, y+, x-, x+, z-, z+
        const bool has[6] = {y > 0 || Z, y < a.ny - 1 || Z, x > 0 || Z, x < a.nx - 1 || Z,
                             z > 0 || Z, z < a.nz - 1 || Z};
        const bool in[6] = {y > 0, y < a.ny - 1, x > 0, x < a.nx - 1, z > 0, z < a.nz - 1};
        const float fr[6] = {a.ay[iay], a.ay[iay + 1], a.ax[iax], a.ax[iax + a.ny], a.az[i],
                             a.az[i + sz]};
        const FaceK fk{a.kf, a.nx, a.ny, a.nz};
        const float kk[6] = {fk.ym(y), fk.yp(y), fk.xm(x), fk.xp(x), fk.zm(z), fk.zp(z)};
        const int off[6] = {-1, 1, -sx, sx, -sz, sz};
        for (int k = 0; k < 6; ++k)
            if (has[k]) {
                cf[k] = fr[k] * kk[k];
                diag += cf[k];
                if (in[k]) nb[k] = i + off[k];
            }
        fi = f[i];
        S[0][i] = v[i];
    }
    __syncthreads();
    int cur = 0;
    for (int it = 0; it < sweeps; ++it) {
        if (i < n) {
            double o = 0.0;
            for (int k = 0; k < 6; ++k)
                if (nb[k] >= 0) o -= cf[k] * S[cur][nb[k]];
            S[cur ^ 1][i] = diag > 1e-12 ? (1.0 - w) * S[cur][i] + w * (fi - o) / diag : 0.0;
        }
        __syncthreads();
        cur ^= 1;
    }
    if (i < n) v[i] = (float)S[cur][i];
}

// The coarsest level's sweeps in FP64 (the reference solves it with one block,
// cuWeightedJacobi3D.cu:1473-1722): one workgroup, FP64 ping-pong in `buf`
// (2 * cells doubles), starting from and ending in v.
template <bool Z>
__global__ void __launch_bounds__(256) coarsest_kernel(float* __restrict__ v,
                                                       const float* __restrict__ f, LevelArgs a,
                                                       double* __restrict__ buf, int sweeps,
                                                       double w) {
    const int64_t n = (int64_t)a.nx * a.ny * a.nz;
    if (n <= 256) {
        coarsest_small<Z>(v, f, a, sweeps, w);
        return;
    }
    double* A = buf;
    double* B = buf + n;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) A[i] = v[i];
    __syncthreads();
    for (int it = 0; it < sweeps; ++it) {
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            const int y = (int)(i % a.ny);
            const int x = (int)((i / a.ny) % a.nx);
            const int z = (int)(i / ((int64_t)a.ny * a.nx));
            double diag, off;
            stencil<Z, double>(a, [&](int64_t j, auto...) { return A[j]; }, x, y, z, i, diag, off);
            B[i] = diag > 1e-12 ? (1.0 - w) * A[i] + w * ((double)f[i] - off) / diag : 0.0;
        }
        __syncthreads();
        double* t = A;
        A = B;
        B = t;
    }
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v[i] = (float)A[i];
}

// FP64 face factors of the finest level, 2 / ((h_nbr + h) h) for the low and
// high face of each cell along each axis (exterior spacing at the walls).
struct Factors64 {
    const double *ylo, *yhi, *xlo, *xhi, *zlo, *zhi;
};

// Residual of the two-float iterate hi + lo in FP64, r = f - A (hi + lo), with
// the operator's coefficients in FP64 (fraction times factor), stored as FP32,
// with per-workgroup FP64 sums of r^2 and f^2 in part.
// The FP64 residual and the conjugate-gradient direction: a workgroup takes
// kZPlanes consecutive z planes of its 64 x 4 tile, so their per-workgroup
// partial sums are kZPlanes times fewer for the one-block reductions after
// them.
constexpr int kZPlanes = 4;

template <bool Z>
__global__ void __launch_bounds__(256) residual64_kernel(float* __restrict__ r,
                                                         const float* __restrict__ hi,
                                                         const float* __restrict__ lo,
                                                         const float* __restrict__ f, LevelArgs a,
                                                         Factors64 k, double* __restrict__ part) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    double sr = 0.0, sf = 0.0;
    const int z1 = min((int)(blockIdx.z + 1) * kZPlanes, a.nz);
    for (int z = blockIdx.z * kZPlanes; z < z1 && y < a.ny && x < a.nx; ++z) {
        const int64_t i = y + (int64_t)x * a.ny + (int64_t)z * a.nx * a.ny;
        const int64_t sx = a.ny, sz = (int64_t)a.nx * a.ny;
        auto val = [&](int64_t j) { return (double)hi[j] + (double)lo[j]; };
        const int64_t iay = y + (int64_t)x * (a.ny + 1) + (int64_t)z * a.nx * (a.ny + 1);
        const int64_t iax = y + (int64_t)x * a.ny + (int64_t)z * (a.nx + 1) * a.ny;
        // every load from a valid index, absent terms zeroed (as stencil())
        const bool ym = y > 0, yp = y < a.ny - 1, xm = x > 0, xp = x < a.nx - 1, zm = z > 0,
                   zp = z < a.nz - 1;
        const double f0 = a.ay[iay], f1 = a.ay[iay + 1], f2 = a.ax[iax], f3 = a.ax[iax + a.ny],
                     f4 = a.az[i], f5 = a.az[i + sz];
        const double k0 = k.ylo[y], k1 = k.yhi[y], k2 = k.xlo[x], k3 = k.xhi[x], k4 = k.zlo[z],
                     k5 = k.zhi[z];
        const double n0 = val(ym ? i - 1 : i), n1 = val(yp ? i + 1 : i), n2 = val(xm ? i - sx : i),
                     n3 = val(xp ? i + sx : i), n4 = val(zm ? i - sz : i), n5 = val(zp ? i + sz : i);
        const double c0 = (ym || Z) ? f0 * k0 : 0.0, c1 = (yp || Z) ? f1 * k1 : 0.0;
        const double c2 = (xm || Z) ? f2 * k2 : 0.0, c3 = (xp || Z) ? f3 * k3 : 0.0;
        const double c4 = (zm || Z) ? f4 * k4 : 0.0, c5 = (zp || Z) ? f5 * k5 : 0.0;
        double diag = 0.0, off = 0.0;
        diag += c0;
        diag += c1;
        diag += c2;
        diag += c3;
        diag += c4;
        diag += c5;
        off -= c0 * (ym ? n0 : 0.0);
        off -= c1 * (yp ? n1 : 0.0);
        off -= c2 * (xm ? n2 : 0.0);
        off -= c3 * (xp ? n3 : 0.0);
        off -= c4 * (zm ? n4 : 0.0);
        off -= c5 * (zp ? n5 : 0.0);
        const double fi = f[i];
        const double ri = diag > 1e-12 ? fi - (diag * val(i) + off) : 0.0;
        r[i] = (float)ri;
        sr += ri * ri;
        sf += fi * fi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        sf += __shfl_xor(sf, o);
    }
    __shared__ double lr[4], lf[4];
    const int t = threadIdx.y * 64 + threadIdx.x;
    if ((t & 63) == 0) {
        lr[t >> 6] = sr;
        lf[t >> 6] = sf;
    }
    __syncthreads();
    if (t == 0) {
        const int64_t b = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
        part[2 * b] = lr[0] + lr[1] + lr[2] + lr[3];
        part[2 * b + 1] = lf[0] + lf[1] + lf[2] + lf[3];
    }
}

// hi + lo <- hi + lo + e as an unevaluated pair (TwoSum of hi and lo + e).
__global__ void pair_add_kernel(float* __restrict__ hi, float* __restrict__ lo,
                                const float* __restrict__ e, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = hi[i], t = lo[i] + e[i];
    const float s = a + t;
    const float bb = s - a;
    hi[i] = s;
    lo[i] = (a - (s - bb)) + (t - bb);
}

// The last sweep of a refinement cycle fused with the pair update: the
// correction e = J(src) of each cell (jacobi_kernel's arithmetic) goes
// straight into hi + lo (pair_add_kernel's TwoSum) and is not stored.
template <bool Z, bool H16 = false>
__global__ void jacobi_pair_kernel(float* __restrict__ hi, float* __restrict__ lo,
                                   const float* __restrict__ src, const float* __restrict__ f, LevelArgs a,
                                   float w) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= a.ny || x >= a.nx) return;
    const int64_t i = y + (int64_t)x * a.ny + (int64_t)z * a.nx * a.ny;
    float diag, off;
    auto rd = [&](int64_t j, auto...) { return src[j]; };
    stencil<Z, float, decltype(rd), H16>(a, rd, x, y, z, i, diag, off);
    const float e = diag > 1e-12f ? (1.f - w) * src[i] + w * (f[i] - off) / diag : 0.f;
    const float hv = hi[i], t = lo[i] + e;
    const float s = hv + t;
    const float bb = s - hv;
    hi[i] = s;
    lo[i] = (hv - (s - bb)) + (t - bb);
}

// Residual f - A v (cuWjGetFracRes_3d_n / _z).
template <bool Z>
__global__ void residual_kernel(float* __restrict__ r, const float* __restrict__ v,
                                const float* __restrict__ f, LevelArgs a) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= a.ny || x >= a.nx) return;
    const int64_t i = y + (int64_t)x * a.ny + (int64_t)z * a.nx * a.ny;
    float diag, off;
    stencil<Z, float>(a, [&](int64_t j, auto...) { return v[j]; }, x, y, z, i, diag, off);
    r[i] = diag > 1e-12f ? f[i] - (diag * v[i] + off) : 0.f;
}

// The residual (residual_kernel) and its restriction to the coarse rhs,
// SolveP1RN with spacings (cuRestrict3D.cu:939-970: the spacing-weighted
// average of the up-to-2x2x2 children, axis by axis, y then x then z), in one
// pass, so the V-cycle's residual never goes to memory.  A (64, 4, 2)
// workgroup forms the residuals of its fine cells (one per thread, as
// residual_kernel) in LDS; 64 of its threads then restrict them into the
// 32 x 2 coarse cells they cover.  Same FP32 values and order as the two
// separate kernels.
template <bool Z, bool H16 = false>
__global__ void __launch_bounds__(512) residual_restrict_kernel(float* __restrict__ dst, const float* __restrict__ v,
                                                                const float* __restrict__ f, LevelArgs a,
                                                                int cx, int cy) {
    __shared__ float rs[2][4][64];
    const int ty = threadIdx.x, tx = threadIdx.y, tz = threadIdx.z;
    const int y = blockIdx.x * 64 + ty, x = blockIdx.y * 4 + tx, z = blockIdx.z * 2 + tz;
    float r = 0.f;
    if (y < a.ny && x < a.nx && z < a.nz) {
        const int64_t i = y + (int64_t)x * a.ny + (int64_t)z * a.nx * a.ny;
        float diag, off;
        auto rd = [&](int64_t j, auto...) { return v[j]; };
        stencil<Z, float, decltype(rd), H16>(a, rd, x, y, z, i, diag, off);
        r = diag > 1e-12f ? f[i] - (diag * v[i] + off) : 0.f;
    }
    rs[tz][tx][ty] = r;
    __syncthreads();
    const int t = ty + 64 * (tx + 4 * tz);
    if (t >= 64) return;
    const int cyl = t & 31, cxl = t >> 5;  // coarse cell of this block
    const int yc = blockIdx.x * 32 + cyl, xc = blockIdx.y * 2 + cxl, zc = blockIdx.z;
    if (yc >= cy || xc >= cx) return;
    const int y0 = 2 * yc, x0 = 2 * xc, z0 = 2 * zc;
    const bool hy1 = y0 + 1 < a.ny, hx1 = x0 + 1 < a.nx, hz1 = z0 + 1 < a.nz;
    float accz = 0.f, wz = 0.f;
    for (int kz = 0; kz < 2; ++kz) {
        if (kz && !hz1) continue;
        float accx = 0.f, wx = 0.f;
        for (int kx = 0; kx < 2; ++kx) {
            if (kx && !hx1) continue;
            const float* row = rs[kz][2 * cxl + kx];
            float accy = a.hy[y0] * row[2 * cyl], wy = a.hy[y0];
            if (hy1) {
                accy += a.hy[y0 + 1] * row[2 * cyl + 1];
                wy += a.hy[y0 + 1];
            }
            accx += a.hx[x0 + kx] * (accy / wy);
            wx += a.hx[x0 + kx];
        }
        accz += a.hz[z0 + kz] * (accx / wx);
        wz += a.hz[z0 + kz];
    }
    dst[yc + (int64_t)xc * cy + (int64_t)zc * cx * cy] = accz / wz;
}


// SolveAddP1RN (cuProlongate3D.cu:594-616): every fine cell adds its parent.
__global__ void prolong_add_kernel(float* __restrict__ v, const float* __restrict__ vc, int nx,
                                   int ny, int nz, int cx, int cy) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= ny || x >= nx) return;
    v[y + (int64_t)x * ny + (int64_t)z * nx * ny] +=
        vc[(y >> 1) + (int64_t)(x >> 1) * cy + (int64_t)(z >> 1) * cx * cy];
}

// SolveP1Ax/Ay/Az (cuRestrict3D.cu:972-1140) for one face array: along the face
// normal the coarse face X takes the fine face min(2X, F-1); along the other
// axes it averages the one or two children.  (fx, fy, fz) are the fine face
// array's extents, (cx, cy, cz) the coarse ones; `normal` is 0, 1 or 2.
__global__ void restrict_faces_kernel(float* __restrict__ dst, const float* __restrict__ src,
                                      int fx, int fy, int fz, int cx, int cy, int cz, int normal) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= cy || x >= cx) return;
    int lo[3], hi[3];
    const int c[3] = {x, y, z}, F[3] = {fx, fy, fz};
    for (int k = 0; k < 3; ++k) {
        if (k == normal) {
            lo[k] = hi[k] = 2 * c[k] < F[k] - 1 ? 2 * c[k] : F[k] - 1;
        } else {
            lo[k] = 2 * c[k];
            hi[k] = 2 * c[k] + 1 < F[k] ? 2 * c[k] + 1 : 2 * c[k];
        }
    }
    // separable average: y innermost, then x, then z (the reference's order)
    float accz = 0.f;
    for (int zz = lo[2]; zz <= hi[2]; ++zz) {
        float accx = 0.f;
        for (int xx = lo[0]; xx <= hi[0]; ++xx) {
            float accy = 0.f;
            for (int yy = lo[1]; yy <= hi[1]; ++yy)
                accy += src[yy + (int64_t)xx * fy + (int64_t)zz * fx * fy];
            accx += accy / (hi[1] - lo[1] + 1);
        }
        accz += accx / (hi[0] - lo[0] + 1);
    }
    dst[y + (int64_t)x * cy + (int64_t)z * cx * cy] = accz / (hi[2] - lo[2] + 1);
}

// cuRestrictDx (WaveletPotentials/cuMultigridFrac2D.cu:4-17).
__global__ void restrict_h_kernel(float* __restrict__ dst, const float* __restrict__ src, int nc,
                                  int nf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    dst[i] = src[2 * i] + (2 * i + 1 < nf ? src[2 * i + 1] : 0.f);
}

__global__ void to_half_kernel(__half* __restrict__ dst, const float* __restrict__ src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = __float2half_rn(src[i]);
}

__global__ void fill_kernel(float* __restrict__ p, float v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void copy_kernel(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Sums of squares of the residual and of f in FP64: per-block partial sums
// (wavefront shuffles, then LDS), added into out[0], out[1] by one block each
// in a second pass.
constexpr int kNormBlocks = 1024;

__global__ void norm_partial_kernel(double* __restrict__ part, const float* __restrict__ r,
                                    const float* __restrict__ f, int64_t n) {
    double sr = 0.0, sf = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double a = r[i], b = f[i];
        sr += a * a;
        sf += b * b;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        sf += __shfl_xor(sf, o);
    }
    __shared__ double lr[4], lf[4];
    if ((threadIdx.x & 63) == 0) {
        lr[threadIdx.x >> 6] = sr;
        lf[threadIdx.x >> 6] = sf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lr[0] + lr[1] + lr[2] + lr[3];
        part[2 * blockIdx.x + 1] = lf[0] + lf[1] + lf[2] + lf[3];
    }
}

__global__ void __launch_bounds__(1024) norm_final_kernel(double* __restrict__ out,
                                                          const double* __restrict__ part,
                                                          int nparts) {
    double sr = 0.0, sf = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        sr += part[2 * i];
        sf += part[2 * i + 1];
    }
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        sf += __shfl_xor(sf, o);
    }
    __shared__ double lr[16], lf[16];
    if ((threadIdx.x & 63) == 0) {
        lr[threadIdx.x >> 6] = sr;
        lf[threadIdx.x >> 6] = sf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            a += lr[k];
            b += lf[k];
        }
        out[0] = a;
        out[1] = b;
    }
}

// ---- preconditioned conjugate gradients ------------------------------------
// Scalars stay on the device: coef[0] = r.z of the current iteration, coef[1]
// = beta, coef[2] = alpha.

// Sum of a 1024-thread block's values (FP64), valid in thread 0.
__device__ double block_sum_1024(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __shared__ double l[16];
    if ((threadIdx.x & 63) == 0) l[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < 16; ++k) t += l[k];
    return t;
}

// part[block] = the block's share of sum a[i] * b[i] in FP64.
__global__ void dot_partial_kernel(double* __restrict__ part, const float* __restrict__ a,
                                   const float* __restrict__ b, int64_t n) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ double l[4];
    if ((threadIdx.x & 63) == 0) l[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = l[0] + l[1] + l[2] + l[3];
}

// beta = (r.z) / (previous r.z), 0 at the first iteration; r.z kept in coef[0].
__global__ void __launch_bounds__(1024) pcg_beta_kernel(double* __restrict__ coef,
                                                        const double* __restrict__ part, int nparts,
                                                        int first) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
    const double rz = block_sum_1024(s);
    if (threadIdx.x == 0) {
        coef[1] = first || !(coef[0] != 0.0) ? 0.0 : rz / coef[0];
        coef[0] = rz;
    }
}

// alpha = (r.z) / (p.Ap); 0 when p.Ap is not positive (a zero direction).
__global__ void __launch_bounds__(1024) pcg_alpha_kernel(double* __restrict__ coef,
                                                         const double* __restrict__ part, int nparts) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
    const double pap = block_sum_1024(s);
    if (threadIdx.x == 0) coef[2] = pap > 0.0 ? coef[0] / pap : 0.0;
}

// The new direction p = z + beta p_old, formed per read, its product Ap with
// the finest operator (FP32 fractions; stencil()), and each workgroup's share
// of p.Ap in part.  pn and po are different arrays (neighbours read p_old).
template <bool Z>
__global__ void __launch_bounds__(256) pcg_direction_kernel(float* __restrict__ pn, float* __restrict__ ap,
                                                            const float* __restrict__ z,
                                                            const float* __restrict__ po, LevelArgs a,
                                                            const double* __restrict__ coef,
                                                            double* __restrict__ part) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const float beta = (float)coef[1];
    double s = 0.0;
    const int z1 = min((int)(blockIdx.z + 1) * kZPlanes, a.nz);
    for (int zc = blockIdx.z * kZPlanes; zc < z1 && y < a.ny && x < a.nx; ++zc) {
        const int64_t i = y + (int64_t)x * a.ny + (int64_t)zc * a.nx * a.ny;
        // p in FP32 (as stored), A p summed in FP64: the residual recurrence
        // r - alpha A p stays close to the true residual only if A p is
        // accurate, and the stencil's terms cancel for smooth p
        auto in = [&](int64_t j, int, int, int) { return z[j] + beta * po[j]; };
        auto in64 = [&](int64_t j, int dy, int dx, int dz) { return (double)in(j, dy, dx, dz); };
        double diag, off;
        stencil<Z, double>(a, in64, x, y, zc, i, diag, off);
        const float v = in(i, 0, 0, 0);
        const double av = diag > 1e-12 ? diag * (double)v + off : 0.0;
        pn[i] = v;
        ap[i] = (float)av;
        s += (double)v * av;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ double l[4];
    const int t = threadIdx.y * 64 + threadIdx.x;
    if ((t & 63) == 0) l[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
        const int64_t b = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
        part[b] = l[0] + l[1] + l[2] + l[3];
    }
}

// hi + lo <- hi + lo + alpha p (the pair rounded from the FP64 sum) and the
// recurrence residual r <- r - alpha Ap, with each block's share of r.r.
__global__ void pcg_update_kernel(float* __restrict__ hi, float* __restrict__ lo, float* __restrict__ r,
                                  const float* __restrict__ p, const float* __restrict__ ap,
                                  const double* __restrict__ coef, double* __restrict__ part, int64_t n) {
    const double alpha = coef[2];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = (double)hi[i] + (double)lo[i] + alpha * (double)p[i];
        const float h = (float)v;
        hi[i] = h;
        lo[i] = (float)(v - (double)h);
        const float ri = (float)((double)r[i] - alpha * (double)ap[i]);
        r[i] = ri;
        s += (double)ri * (double)ri;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ double l[4];
    if ((threadIdx.x & 63) == 0) l[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = l[0] + l[1] + l[2] + l[3];
}

// out[0] = the sum of nparts partials (one block).
__global__ void __launch_bounds__(1024) sum_final_kernel(double* __restrict__ out,
                                                         const double* __restrict__ part, int nparts) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
    const double t = block_sum_1024(s);
    if (threadIdx.x == 0) out[0] = t;
}


// ---- the coarse levels of a V-cycle in one workgroup -----------------------
// Below a few thousand cells a level's launches are latency-bound (a few
// microseconds each, six per level per cycle).  This kernel runs the V-cycle
// from the first small level to the coarsest and back in one workgroup, with
// every small level's iterate, partner and right-hand side in LDS and each
// cell's face coefficients precomputed at create (the same FP32 values and
// summation order as the per-level kernels), so a sweep is seven coefficient
// loads, seven LDS reads and a few FMAs per cell.  The coarsest level's sweeps
// run in FP64, on one wavefront when it has at most 64 cells.
constexpr int kSmallThreads = 1024;
constexpr int kMaxSmall = 16;

struct SmallLevel {
    int nx, ny, nz, n;
    const float* hx;
    const float* hy;
    const float* hz;
    const float* coef;    // 7 x n: y-, y+, x-, x+, z-, z+ (0 without a neighbour), diagonal
    const double* diag64; // the coarsest level: FP64 diagonal (walls included)
    int ov, ot, of;       // LDS float offsets of v, t, f
};

struct SmallCycle {
    SmallLevel lv[kMaxSmall];
    int nlev;             // small levels, the coarsest included
    int o64;              // LDS float offset of the coarsest level's FP64 ping-pong
    float* v0;            // the first small level's iterate (global)
    const float* f0;      // and its right-hand side
};

// Face coefficients of every cell of one level, in stencil()'s order.
template <bool Z>
__global__ void small_coef_kernel(float* __restrict__ coef, double* __restrict__ diag64, LevelArgs a) {
    const int n = a.nx * a.ny * a.nz;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int y = i % a.ny, x = (i / a.ny) % a.nx, z = i / (a.ny * a.nx);
    const int iay = y + x * (a.ny + 1) + z * a.nx * (a.ny + 1);
    const int iax = y + x * a.ny + z * (a.nx + 1) * a.ny;
    const int sz = a.nx * a.ny;
    const bool in[6] = {y > 0, y < a.ny - 1, x > 0, x < a.nx - 1, z > 0, z < a.nz - 1};
    const float fr[6] = {a.ay[iay], a.ay[iay + 1], a.ax[iax], a.ax[iax + a.ny], a.az[i], a.az[i + sz]};
    const FaceK fk{a.kf, a.nx, a.ny, a.nz};
    const float kk[6] = {fk.ym(y), fk.yp(y), fk.xm(x), fk.xp(x), fk.zm(z), fk.zp(z)};
    float diag = 0.f;
    double d64 = 0.0;
    for (int k = 0; k < 6; ++k) {
        float c = 0.f;
        if (in[k] || Z) {
            c = fr[k] * kk[k];
            diag += c;
            d64 += (double)c;
        }
        coef[k * n + i] = in[k] ? c : 0.f;
    }
    coef[6 * n + i] = diag;
    if (diag64) diag64[i] = d64;
}

// Neighbour indices of cell i (clamped into the level; a missing neighbour
// has a zero coefficient).
struct Nb {
    int k[6];
};
__device__ __forceinline__ Nb small_nb(int i, int ny, int sz, int n) {
    Nb r;
    r.k[0] = i > 0 ? i - 1 : i;
    r.k[1] = i + 1 < n ? i + 1 : i;
    r.k[2] = i - ny >= 0 ? i - ny : i;
    r.k[3] = i + ny < n ? i + ny : i;
    r.k[4] = i - sz >= 0 ? i - sz : i;
    r.k[5] = i + sz < n ? i + sz : i;
    return r;
}

// off = -sum_k c_k src[nb_k] in stencil()'s order
__device__ __forceinline__ float small_off(const SmallLevel& L, const float* src, int i, float& diag) {
    const Nb nb = small_nb(i, L.ny, L.nx * L.ny, L.n);
    float off = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) off -= L.coef[k * L.n + i] * src[nb.k[k]];
    diag = L.coef[6 * L.n + i];
    return off;
}

__device__ void small_smooth(const SmallLevel& L, float* sm, int nu, float w) {
    float* a = sm + L.ov;
    float* b = sm + L.ot;
    const float* f = sm + L.of;
    for (int it = 0; it < nu; ++it) {
        for (int i = threadIdx.x; i < L.n; i += kSmallThreads) {
            float diag;
            const float off = small_off(L, a, i, diag);
            b[i] = diag > 1e-12f ? (1.f - w) * a[i] + w * (f[i] - off) / diag : 0.f;
        }
        __syncthreads();
        float* t = a;
        a = b;
        b = t;
    }
    if (a != sm + L.ov) {
        for (int i = threadIdx.x; i < L.n; i += kSmallThreads) sm[L.ov + i] = a[i];
        __syncthreads();
    }
}

template <bool Z>
__global__ void __launch_bounds__(kSmallThreads) vcycle_small_kernel(SmallCycle c, int nu1, int nu2,
                                                                    float w, int csweeps) {
    extern __shared__ double small_lds[];
    float* sm = (float*)small_lds;
    const int tid = threadIdx.x;
    const int last = c.nlev - 1;
    {
        const SmallLevel& L = c.lv[0];
        for (int i = tid; i < L.n; i += kSmallThreads) {
            sm[L.of + i] = c.f0[i];
            sm[L.ov + i] = 0.f;
        }
        __syncthreads();
    }
    for (int l = 0; l < last; ++l) {
        const SmallLevel& L = c.lv[l];
        const SmallLevel& C = c.lv[l + 1];
        small_smooth(L, sm, nu1, w);
        // residual into the partner array
        for (int i = tid; i < L.n; i += kSmallThreads) {
            float diag;
            const float off = small_off(L, sm + L.ov, i, diag);
            sm[L.ot + i] = diag > 1e-12f ? sm[L.of + i] - (diag * sm[L.ov + i] + off) : 0.f;
        }
        __syncthreads();
        // spacing-weighted restriction (residual_restrict_kernel) and a zero coarse iterate
        const float* t = sm + L.ot;
        for (int i = tid; i < C.n; i += kSmallThreads) {
            const int y = i % C.ny, x = (i / C.ny) % C.nx, z = i / (C.ny * C.nx);
            const int x0 = 2 * x, y0 = 2 * y, z0 = 2 * z;
            const int x1 = x0 + 1 < L.nx ? x0 + 1 : -1, y1 = y0 + 1 < L.ny ? y0 + 1 : -1,
                      z1 = z0 + 1 < L.nz ? z0 + 1 : -1;
            const int sx = L.ny, sz = L.nx * L.ny;
            float accz = 0.f, wz = 0.f;
            for (int kz = 0; kz < 2; ++kz) {
                const int zz = kz ? z1 : z0;
                if (zz < 0) continue;
                float accx = 0.f, wx = 0.f;
                for (int kx = 0; kx < 2; ++kx) {
                    const int xx = kx ? x1 : x0;
                    if (xx < 0) continue;
                    const int b = xx * sx + zz * sz;
                    float accy = L.hy[y0] * t[b + y0], wy = L.hy[y0];
                    if (y1 >= 0) {
                        accy += L.hy[y1] * t[b + y1];
                        wy += L.hy[y1];
                    }
                    accx += L.hx[xx] * (accy / wy);
                    wx += L.hx[xx];
                }
                accz += L.hz[zz] * (accx / wx);
                wz += L.hz[zz];
            }
            sm[C.of + i] = accz / wz;
            sm[C.ov + i] = 0.f;
        }
        __syncthreads();
    }
    {
        // coarsest level: FP64 sweeps, one cell per thread, coefficients in registers
        const SmallLevel& L = c.lv[last];
        double* S = (double*)(sm + c.o64);  // 2 x L.n doubles
        const bool one_wave = L.n <= 64;
        if (!one_wave || tid < 64) {
            const int i = tid;
            double cf[6] = {0, 0, 0, 0, 0, 0}, diag = 0.0, fi = 0.0;
            Nb nb{};
            if (i < L.n) {
                nb = small_nb(i, L.ny, L.nx * L.ny, L.n);
                for (int k = 0; k < 6; ++k) cf[k] = (double)L.coef[k * L.n + i];
                diag = L.diag64[i];
                fi = (double)sm[L.of + i];
                S[i] = (double)sm[L.ov + i];
            }
            if (one_wave) {
                __builtin_amdgcn_wave_barrier();
                asm volatile("" ::: "memory");
            } else {
                __syncthreads();
            }
            // w / diag once: a multiply per sweep instead of an FP64 division
            const double wd = diag > 1e-12 ? (double)w / diag : 0.0;
            int cur = 0;
            for (int it = 0; it < csweeps; ++it) {
                if (i < L.n) {
                    const double* a = S + cur * L.n;
                    double o = 0.0;
                    for (int k = 0; k < 6; ++k) o -= cf[k] * a[nb.k[k]];
                    S[(cur ^ 1) * L.n + i] = diag > 1e-12 ? (1.0 - (double)w) * a[i] + wd * (fi - o) : 0.0;
                }
                if (one_wave) {
                    __builtin_amdgcn_wave_barrier();
                    asm volatile("" ::: "memory");
                } else {
                    __syncthreads();
                }
                cur ^= 1;
            }
            if (i < L.n) sm[L.ov + i] = (float)S[cur * L.n + i];
        }
        __syncthreads();
    }
    for (int l = last - 1; l >= 0; --l) {
        const SmallLevel& L = c.lv[l];
        const SmallLevel& C = c.lv[l + 1];
        for (int i = tid; i < L.n; i += kSmallThreads) {
            const int y = i % L.ny, x = (i / L.ny) % L.nx, z = i / (L.ny * L.nx);
            sm[L.ov + i] += sm[C.ov + (y >> 1) + (x >> 1) * C.ny + (z >> 1) * C.nx * C.ny];
        }
        __syncthreads();
        small_smooth(L, sm, nu2, w);
    }
    {
        const SmallLevel& L = c.lv[0];
        for (int i = tid; i < L.n; i += kSmallThreads) c.v0[i] = sm[L.ov + i];
    }
}

}  // namespace
}  // namespace whhd

struct whhd_mgfrac3d_s {
    whhd::Dim3 dim;
    char bc;            // 'n' or 'z' (the reference maps 'd' to 'z')
    float he[3];        // exterior spacing: dx for 'n'/'z', 0 for 'd'
    std::vector<whhd::Level> lv;
    float* r = nullptr;       // fine-level residual
    double* part = nullptr;   // norm partial sums (2 * kNormBlocks) + the two totals
    double* host = nullptr;   // pinned host copy of the totals
    double* cbuf = nullptr;   // FP64 ping-pong of the coarsest level (2 * its cells)
    double* fac64 = nullptr;  // finest level's FP64 face factors: ylo, yhi, xlo, xhi, zlo, zhi
    whhd::SmallCycle small{};        // the small levels' table (one-workgroup V-cycle)
    size_t small_lds = 0;            // its LDS bytes
    int fuse_from = -1;              // levels from here down run in one workgroup
    bool fp16_fine = false;          // V-cycles read the finest level's FP16 fractions
    int coarsest_n = 50;             // coarsest-level sweeps for bc 'n' (the reference's 50)
    double* rpart = nullptr;  // per-workgroup sums of the FP64 residual kernel, + 2 totals
    int64_t rblocks = 0;
    float* pcg_p[2] = {nullptr, nullptr};  // conjugate directions (ping-pong), finest level
    double* coef = nullptr;                // PCG scalars: r.z, beta, alpha
    hipEvent_t norm_ready = nullptr;       // the residual norms' copy to `host` is done
    std::vector<void*> owned;
};

namespace whhd {
namespace {

LevelArgs args_of(const whhd_mgfrac3d_s* h, const Level& L) {
    const bool h16 = h->fp16_fine && L.axh;
    return LevelArgs{L.ax, L.ay, L.az, L.hx, L.hy, L.hz, L.d.x, L.d.y, L.d.z,
                     h->he[0], h->he[1], h->he[2], L.kf,
                     h16 ? L.axh : nullptr, h16 ? L.ayh : nullptr, h16 ? L.azh : nullptr};
}

hipError_t launch_status(const char* name) {
    set_last_kernel(name);
    return hipGetLastError();
}

// One sweep dst <- J(src); zero: src is zero and is not read; C (not with
// zero): the sweep's input is src plus C's correction prolonged.
int jacobi(const whhd_mgfrac3d_s* h, const Level& L, float* dst, const float* src, const float* f,
           float w, hipStream_t s, bool zero = false, const Level* C = nullptr) {
    const LevelArgs a = args_of(h, L);
    const dim3 g = grid_of(L.d.x, L.d.y, L.d.z), b(64, 4);
    const float* vc = C ? C->v : nullptr;
    const int cx = C ? C->d.x : 0, cy = C ? C->d.y : 0;
#define WHHD_JAC(Zb, H)                                                                                 \
    do {                                                                                              \
        if (zero) hipLaunchKernelGGL((jacobi_kernel<Zb, true, false, H>), g, b, 0, s, dst, src, f, a, w, vc, cx, cy);  \
        else if (C) hipLaunchKernelGGL((jacobi_kernel<Zb, false, true, H>), g, b, 0, s, dst, src, f, a, w, vc, cx, cy); \
        else hipLaunchKernelGGL((jacobi_kernel<Zb, false, false, H>), g, b, 0, s, dst, src, f, a, w, vc, cx, cy);      \
    } while (0)
    const bool h16 = a.axh != nullptr;
    if (h->bc == 'z') {
        if (h16) WHHD_JAC(true, true);
        else WHHD_JAC(true, false);
    } else {
        if (h16) WHHD_JAC(false, true);
        else WHHD_JAC(false, false);
    }
#undef WHHD_JAC
    hipError_t e = launch_status("mg_jacobi_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg jacobi kernel");
}

int residual(const whhd_mgfrac3d_s* h, const Level& L, float* r, const float* v, const float* f,
             hipStream_t s) {
    const LevelArgs a = args_of(h, L);
    if (h->bc == 'z')
        hipLaunchKernelGGL(residual_kernel<true>, grid_of(L.d.x, L.d.y, L.d.z), dim3(64, 4), 0, s,
                           r, v, f, a);
    else
        hipLaunchKernelGGL(residual_kernel<false>, grid_of(L.d.x, L.d.y, L.d.z), dim3(64, 4), 0, s,
                           r, v, f, a);
    hipError_t e = launch_status("mg_residual_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg residual kernel");
}

int copy(float* dst, const float* src, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
    hipError_t e = launch_status("mg_copy_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg copy kernel");
}

int fill(float* dst, float v, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, v, n);
    hipError_t e = launch_status("mg_fill_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg fill kernel");
}

int64_t cells(const Dim3& d) { return (int64_t)d.x * d.y * d.z; }

#define TRY(e) do { int s_ = (e); if (s_ != WHHD_OK) return s_; } while (0)

// n Jacobi sweeps on level L (SolveFrac, cuWeightedJacobi3D.cu:1075-1105); the
// result stays in L.v.  zero: L.v holds zero on entry, so the first sweep does
// not read it (and with no sweep it is written).
// hi, lo (n > 0, zero false): the last sweep adds its result to the pair
// hi + lo instead of storing it (L.v is then left as it was).
// C (n >= 2, or no pair): the first sweep reads L.v plus C's prolonged correction.
int smooth(const whhd_mgfrac3d_s* h, const Level& L, int n, float w, hipStream_t s, bool zero,
           float* hi = nullptr, float* lo = nullptr, const Level* C = nullptr) {
    float* a = L.v;
    float* b = L.t;
    if (zero && n == 0) return fill(L.v, 0.f, cells(L.d), s);
    for (int k = 0; k < n; ++k) {
        if (C && k == 0) {
            TRY(jacobi(h, L, b, a, L.f, w, s, false, C));
            std::swap(a, b);
            continue;
        }
        if (hi && k == n - 1) {
            const LevelArgs la = args_of(h, L);
            const dim3 pg = grid_of(L.d.x, L.d.y, L.d.z), pb(64, 4);
            const bool h16 = la.axh != nullptr;
            if (h->bc == 'z') {
                if (h16) hipLaunchKernelGGL((jacobi_pair_kernel<true, true>), pg, pb, 0, s, hi, lo, a, L.f, la, w);
                else hipLaunchKernelGGL((jacobi_pair_kernel<true, false>), pg, pb, 0, s, hi, lo, a, L.f, la, w);
            } else {
                if (h16) hipLaunchKernelGGL((jacobi_pair_kernel<false, true>), pg, pb, 0, s, hi, lo, a, L.f, la, w);
                else hipLaunchKernelGGL((jacobi_pair_kernel<false, false>), pg, pb, 0, s, hi, lo, a, L.f, la, w);
            }
            hipError_t e = launch_status("mg_jacobi_pair_kernel");
            return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg jacobi pair kernel");
        }
        TRY(jacobi(h, L, b, a, L.f, w, s, zero && k == 0));
        std::swap(a, b);
    }
    if (a != L.v) TRY(copy(L.v, a, cells(L.d), s));
    return WHHD_OK;
}

// The reference's coarsest level runs 50 ('n') or 20 ('z') sweeps of the
// volume-scaled system in one block (cuWeightedJacobi3D.cu:1473-1722); scaling
// rows by the cell volume leaves the Jacobi iteration unchanged, so the same
// sweeps run here, in FP64, by one workgroup.
int coarsest(const whhd_mgfrac3d_s* h, const Level& L, float w, hipStream_t s) {
    const LevelArgs a = args_of(h, L);
    const int sweeps = h->bc == 'z' ? 20 : h->coarsest_n;
    if (h->bc == 'z')
        hipLaunchKernelGGL(coarsest_kernel<true>, dim3(1), dim3(256), 0, s, L.v, L.f, a, h->cbuf,
                           sweeps, (double)w);
    else
        hipLaunchKernelGGL(coarsest_kernel<false>, dim3(1), dim3(256), 0, s, L.v, L.f, a, h->cbuf,
                           sweeps, (double)w);
    hipError_t e = launch_status("mg_coarsest_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg coarsest kernel");
}

// One V-cycle from level l down.  zero: the level's iterate L.v is zero on
// entry (every coarse level's correction; the refinement's error) and is not
// read or cleared first.
int vcycle_body(const whhd_mgfrac3d_s* h, size_t l, int nu1, int nu2, float w, hipStream_t s, bool zero,
                float* hi, float* lo);

// hi, lo: the cycle's result is added to the pair hi + lo (the refinement),
// fused into the last post-smoothing sweep when there is one.
int vcycle(const whhd_mgfrac3d_s* h, size_t l, int nu1, int nu2, float w, hipStream_t s, bool zero,
           float* hi = nullptr, float* lo = nullptr) {
    const Level& L = h->lv[l];
    const bool fused = hi && nu2 > 0 && l + 1 < h->lv.size() && !(h->fuse_from >= 0 && (int)l == h->fuse_from);
    TRY(vcycle_body(h, l, nu1, nu2, w, s, zero, fused ? hi : nullptr, fused ? lo : nullptr));
    if (!hi || fused) return WHHD_OK;
    const int64_t n = cells(L.d);
    hipLaunchKernelGGL(pair_add_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, hi, lo, L.v, n);
    hipError_t e = launch_status("mg_pair_add_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg pair update kernel");
}

int vcycle_body(const whhd_mgfrac3d_s* h, size_t l, int nu1, int nu2, float w, hipStream_t s, bool zero,
                float* hi, float* lo) {
    const Level& L = h->lv[l];
    if (h->fuse_from >= 0 && (int)l == h->fuse_from) {
        // the small levels start from a zero iterate in LDS (coarse levels are
        // entered with zero)
        const int cs = h->bc == 'z' ? 20 : h->coarsest_n;
        if (h->bc == 'z')
            hipLaunchKernelGGL(vcycle_small_kernel<true>, dim3(1), dim3(kSmallThreads), h->small_lds, s,
                               h->small, nu1, nu2, w, cs);
        else
            hipLaunchKernelGGL(vcycle_small_kernel<false>, dim3(1), dim3(kSmallThreads), h->small_lds, s,
                               h->small, nu1, nu2, w, cs);
        hipError_t e = launch_status("mg_vcycle_small_kernel");
        return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg small-level V-cycle kernel");
    }
    if (l + 1 >= h->lv.size()) {
        if (zero) TRY(fill(L.v, 0.f, cells(L.d), s));
        return coarsest(h, L, w, s);
    }
    const Level& C = h->lv[l + 1];
    TRY(smooth(h, L, nu1, w, s, zero));
    // residual and its restriction to the coarse rhs in one pass
    const LevelArgs a = args_of(h, L);
    const dim3 rg((L.d.y + 63) / 64, (L.d.x + 3) / 4, (L.d.z + 1) / 2), rb(64, 4, 2);
    const bool h16 = a.axh != nullptr;
    if (h->bc == 'z') {
        if (h16) hipLaunchKernelGGL((residual_restrict_kernel<true, true>), rg, rb, 0, s, C.f, L.v, L.f, a, C.d.x, C.d.y);
        else hipLaunchKernelGGL((residual_restrict_kernel<true, false>), rg, rb, 0, s, C.f, L.v, L.f, a, C.d.x, C.d.y);
    } else {
        if (h16) hipLaunchKernelGGL((residual_restrict_kernel<false, true>), rg, rb, 0, s, C.f, L.v, L.f, a, C.d.x, C.d.y);
        else hipLaunchKernelGGL((residual_restrict_kernel<false, false>), rg, rb, 0, s, C.f, L.v, L.f, a, C.d.x, C.d.y);
    }
    hipError_t e = launch_status("mg_residual_restrict_kernel");
    if (e != hipSuccess) return hip_fail(e, "mg residual-restrict kernel");
    TRY(vcycle(h, l + 1, nu1, nu2, w, s, true));
    // the prolongation rides on the first post-smoothing sweep when that sweep
    // is not also the pair update
    if (nu2 >= 2 || (nu2 == 1 && !hi)) return smooth(h, L, nu2, w, s, false, hi, lo, &C);
    hipLaunchKernelGGL(prolong_add_kernel, grid_of(L.d.x, L.d.y, L.d.z), dim3(64, 4), 0, s, L.v,
                       C.v, L.d.x, L.d.y, L.d.z, C.d.x, C.d.y);
    e = launch_status("mg_prolong_add_kernel");
    if (e != hipSuccess) return hip_fail(e, "mg prolong kernel");
    return smooth(h, L, nu2, w, s, false, hi, lo);
}

void release(whhd_mgfrac3d_s* h) {
    if (!h) return;
    for (void* p : h->owned) (void)hipFree(p);
    if (h->host) (void)hipHostFree(h->host);
    if (h->norm_ready) (void)hipEventDestroy(h->norm_ready);
    delete h;
}

}  // namespace

int mgfrac3d_create(const float* ax, const float* ay, const float* az, int nx, int ny, int nz,
                    double dx, double dy, double dz, char bc, whhd_mgfrac3d_s** out) {
    if (!out) return fail(WHHD_INVALID_ARG, "mgfrac3d_create: null handle pointer");
    *out = nullptr;
    if (!ax || !ay || !az) return fail(WHHD_INVALID_ARG, "mgfrac3d_create: null fraction array");
    if (nx < 2 || ny < 2 || nz < 2 || nz > 65535 || (int64_t)(nx + 1) * (ny + 1) * (nz + 1) > (1LL << 31))
        return fail(WHHD_INVALID_ARG, "mgfrac3d_create: grid extents out of range");
    if (bc != 'n' && bc != 'd' && bc != 'z')
        return fail(WHHD_INVALID_ARG, "mgfrac3d_create: bc must be 'n', 'd' or 'z'");
    if (!(dx > 0 && dy > 0 && dz > 0)) return fail(WHHD_INVALID_ARG, "mgfrac3d_create: spacing <= 0");
    auto* h = new whhd_mgfrac3d_s;
    h->dim = {nx, ny, nz};
    // cuMultigridFrac3D.cu:12-13: 'd' runs as 'z' with a wall on the face
    h->bc = bc == 'n' ? 'n' : 'z';
    h->he[0] = bc == 'd' ? 0.f : (float)dx;
    h->he[1] = bc == 'd' ? 0.f : (float)dy;
    h->he[2] = bc == 'd' ? 0.f : (float)dz;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        h->owned.push_back(p);
        return p;
    };
    hipStream_t s = nullptr;
    Dim3 d = {nx, ny, nz}, last = d;
    int min_dim = std::min(nx, std::min(ny, nz));
    int st = WHHD_OK;
    while (min_dim > 1 && st == WHHD_OK) {
        Level L;
        L.d = d;
        const size_t n = (size_t)cells(d);
        L.f = (float*)alloc(n * 4);
        L.v = (float*)alloc(n * 4);
        L.t = (float*)alloc(n * 4);
        L.ax = (float*)alloc((size_t)(d.x + 1) * d.y * d.z * 4);
        L.ay = (float*)alloc((size_t)d.x * (d.y + 1) * d.z * 4);
        L.az = (float*)alloc((size_t)d.x * d.y * (d.z + 1) * 4);
        L.hx = (float*)alloc((size_t)d.x * 4);
        L.hy = (float*)alloc((size_t)d.y * 4);
        L.hz = (float*)alloc((size_t)d.z * 4);
        if (!L.f || !L.v || !L.t || !L.ax || !L.ay || !L.az || !L.hx || !L.hy || !L.hz) {
            st = fail(WHHD_OOM, "mgfrac3d_create: allocation failed");
            break;
        }
        if (h->lv.empty()) {
            hipError_t e = hipMemcpyAsync(L.ax, ax, (size_t)(d.x + 1) * d.y * d.z * 4,
                                          hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(L.ay, ay, (size_t)d.x * (d.y + 1) * d.z * 4,
                                   hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(L.az, az, (size_t)d.x * d.y * (d.z + 1) * 4,
                                   hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) {
                st = hip_fail(e, "mgfrac3d_create: fraction copy");
                break;
            }
            st = fill(L.hx, (float)dx, d.x, s);
            if (st == WHHD_OK) st = fill(L.hy, (float)dy, d.y, s);
            if (st == WHHD_OK) st = fill(L.hz, (float)dz, d.z, s);
        } else {
            const Level& P = h->lv.back();
            hipLaunchKernelGGL(restrict_faces_kernel, grid_of(d.x + 1, d.y, d.z), dim3(64, 4), 0, s,
                               L.ax, P.ax, last.x + 1, last.y, last.z, d.x + 1, d.y, d.z, 0);
            hipLaunchKernelGGL(restrict_faces_kernel, grid_of(d.x, d.y + 1, d.z), dim3(64, 4), 0, s,
                               L.ay, P.ay, last.x, last.y + 1, last.z, d.x, d.y + 1, d.z, 1);
            hipLaunchKernelGGL(restrict_faces_kernel, grid_of(d.x, d.y, d.z + 1), dim3(64, 4), 0, s,
                               L.az, P.az, last.x, last.y, last.z + 1, d.x, d.y, d.z + 1, 2);
            hipLaunchKernelGGL(restrict_h_kernel, dim3((d.x + 255) / 256), dim3(256), 0, s, L.hx,
                               P.hx, d.x, last.x);
            hipLaunchKernelGGL(restrict_h_kernel, dim3((d.y + 255) / 256), dim3(256), 0, s, L.hy,
                               P.hy, d.y, last.y);
            hipLaunchKernelGGL(restrict_h_kernel, dim3((d.z + 255) / 256), dim3(256), 0, s, L.hz,
                               P.hz, d.z, last.z);
            hipError_t e = launch_status("mg_setup_kernels");
            if (e != hipSuccess) st = hip_fail(e, "mgfrac3d_create: hierarchy kernels");
        }
        h->lv.push_back(L);
        last = d;
        d = {half_up(d.x), half_up(d.y), half_up(d.z)};
        min_dim = std::min(d.x, std::min(d.y, d.z));
    }
    if (st == WHHD_OK && !h->lv.empty()) {
        // FP16 copies of the finest level's fractions for the tolerance solve's
        // V-cycles (its FP64 residual keeps the FP32 fractions)
        Level& L = h->lv[0];
        const int64_t n[3] = {(int64_t)(L.d.x + 1) * L.d.y * L.d.z, (int64_t)L.d.x * (L.d.y + 1) * L.d.z,
                              (int64_t)L.d.x * L.d.y * (L.d.z + 1)};
        __half** dst[3] = {&L.axh, &L.ayh, &L.azh};
        const float* src[3] = {L.ax, L.ay, L.az};
        for (int c = 0; c < 3 && st == WHHD_OK; ++c) {
            *dst[c] = (__half*)alloc((size_t)n[c] * sizeof(__half));
            if (!*dst[c]) {
                st = fail(WHHD_OOM, "mgfrac3d_create: allocation failed");
                break;
            }
            hipLaunchKernelGGL(to_half_kernel, dim3((unsigned)((n[c] + 255) / 256)), dim3(256), 0, s, *dst[c],
                               src[c], n[c]);
            hipError_t e = launch_status("mg_to_half_kernel");
            if (e != hipSuccess) st = hip_fail(e, "mgfrac3d_create: FP16 fractions");
        }
    }
    for (Level& L : h->lv) {
        if (st != WHHD_OK) break;
        const int nk = 2 * (L.d.x + L.d.y + L.d.z);
        L.kf = (float*)alloc((size_t)nk * 4);
        if (!L.kf) {
            st = fail(WHHD_OOM, "mgfrac3d_create: allocation failed");
            break;
        }
        hipLaunchKernelGGL(face_factor_kernel, dim3((nk + 255) / 256), dim3(256), 0, s, L.kf, L.hx, L.hy,
                           L.hz, L.d.x, L.d.y, L.d.z, h->he[0], h->he[1], h->he[2]);
        hipError_t e = launch_status("mg_face_factor_kernel");
        if (e != hipSuccess) st = hip_fail(e, "mgfrac3d_create: face factors");
    }
    if (st == WHHD_OK) {
        const dim3 g = grid_of(nx, ny, nz);
        h->rblocks = (int64_t)g.x * g.y * g.z;
        h->r = (float*)alloc((size_t)cells(h->dim) * 4);
        h->part = (double*)alloc((2 * kNormBlocks + 2) * sizeof(double));
        h->cbuf = (double*)alloc((size_t)cells(h->lv.back().d) * 2 * sizeof(double));
        h->rpart = (double*)alloc((size_t)(2 * h->rblocks + 2) * sizeof(double));
        h->fac64 = (double*)alloc((size_t)2 * (nx + ny + nz) * sizeof(double));
        h->pcg_p[0] = (float*)alloc((size_t)cells(h->dim) * 4);
        h->pcg_p[1] = (float*)alloc((size_t)cells(h->dim) * 4);
        h->coef = (double*)alloc(4 * sizeof(double));
        if (!h->r || !h->part || !h->cbuf || !h->rpart || !h->fac64 || !h->pcg_p[0] || !h->pcg_p[1] ||
            !h->coef)
            st = fail(WHHD_OOM, "mgfrac3d_create: allocation failed");
    }
    if (st == WHHD_OK) {
        // levels of at most kFuseCells cells (not the finest) run in one
        // workgroup, when the coarsest has at most 256 cells and the small
        // levels fit in LDS
        constexpr int64_t kFuseCells = 4096;
        const int nl = (int)h->lv.size();
        int from = -1;
        for (int l = 1; l < nl; ++l)
            if (cells(h->lv[l].d) <= kFuseCells) { from = l; break; }
        if (from >= 0 && from < nl - 1 && nl - from <= kMaxSmall && cells(h->lv[nl - 1].d) <= 256) {
            SmallCycle& c = h->small;
            c.nlev = nl - from;
            int off = 0;
            for (int k = 0; k < c.nlev && st == WHHD_OK; ++k) {
                const Level& L = h->lv[from + k];
                SmallLevel& S = c.lv[k];
                S.nx = L.d.x; S.ny = L.d.y; S.nz = L.d.z;
                S.n = (int)cells(L.d);
                S.hx = L.hx; S.hy = L.hy; S.hz = L.hz;
                float* coef = (float*)alloc((size_t)7 * S.n * sizeof(float));
                double* d64 = k == c.nlev - 1 ? (double*)alloc((size_t)S.n * sizeof(double)) : nullptr;
                if (!coef || (k == c.nlev - 1 && !d64)) {
                    st = fail(WHHD_OOM, "mgfrac3d_create: small-level tables");
                    break;
                }
                S.coef = coef;
                S.diag64 = d64;
                const LevelArgs a = args_of(h, L);
                if (h->bc == 'z')
                    hipLaunchKernelGGL(small_coef_kernel<true>, dim3((S.n + 255) / 256), dim3(256), 0, s,
                                       coef, d64, a);
                else
                    hipLaunchKernelGGL(small_coef_kernel<false>, dim3((S.n + 255) / 256), dim3(256), 0, s,
                                       coef, d64, a);
                if (hipGetLastError() != hipSuccess) st = fail(WHHD_HIP_ERROR, "mgfrac3d_create: small-level tables");
                S.ov = off; off += S.n;
                S.ot = off; off += S.n;
                S.of = off; off += S.n;
            }
            off = (off + 1) & ~1;  // 8-byte alignment for the FP64 ping-pong
            c.o64 = off;
            off += 2 * 2 * (int)cells(h->lv[nl - 1].d);
            c.v0 = h->lv[from].v;
            c.f0 = h->lv[from].f;
            h->small_lds = (size_t)off * sizeof(float);
            if (st == WHHD_OK && h->small_lds <= 60 * 1024) h->fuse_from = from;
        }
    }
    if (st == WHHD_OK) {
        // uniform finest spacing: interior faces 2 / (2 h h), wall faces 2 / ((e + h) h)
        std::vector<double> k;
        const int ns[3] = {ny, nx, nz};
        const double hs[3] = {dy, dx, dz};
        const bool dwall = bc == 'd';  // zero exterior spacing
        for (int a = 0; a < 3; ++a) {
            for (int side = 0; side < 2; ++side)
                for (int c = 0; c < ns[a]; ++c) {
                    const bool wall = side == 0 ? c == 0 : c == ns[a] - 1;
                    const double hn = wall ? (dwall ? 0.0 : hs[a]) : hs[a];
                    k.push_back(2.0 / ((hn + hs[a]) * hs[a]));
                }
        }
        if (hipMemcpy(h->fac64, k.data(), k.size() * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess)
            st = fail(WHHD_HIP_ERROR, "mgfrac3d_create: factor upload");
    }
    if (st == WHHD_OK && hipHostMalloc((void**)&h->host, 2 * sizeof(double)) != hipSuccess) {
        h->host = nullptr;
        st = fail(WHHD_OOM, "mgfrac3d_create: pinned allocation failed");
    }
    if (st == WHHD_OK && hipEventCreateWithFlags(&h->norm_ready, hipEventDisableTiming) != hipSuccess) {
        h->norm_ready = nullptr;
        st = fail(WHHD_HIP_ERROR, "mgfrac3d_create: event");
    }
    if (st == WHHD_OK) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(e, "mgfrac3d_create: setup");
    }
    if (st != WHHD_OK) {
        release(h);
        return st;
    }
    *out = h;
    return WHHD_OK;
}

int mgfrac3d_solve(whhd_mgfrac3d_s* h, float* dst, const float* pre_v, const float* f,
                   int n_iters, int nu1, int nu2, float w, hipStream_t s) {
    if (!h || !dst || !f) return fail(WHHD_INVALID_ARG, "mgfrac3d_solve: null argument");
    if (n_iters < 0 || nu1 < 0 || nu2 < 0)
        return fail(WHHD_INVALID_ARG, "mgfrac3d_solve: negative iteration count");
    Level& L0 = h->lv[0];
    const int64_t n = cells(h->dim);
    // Solve (cuMultigridFrac3D.cu:249-265): the iterate starts from pre_v or zero
    if (pre_v) TRY(copy(L0.v, pre_v, n, s));
    else TRY(fill(L0.v, 0.f, n, s));
    TRY(copy(L0.f, f, n, s));
    for (int i = 0; i < n_iters; ++i) TRY(vcycle(h, 0, nu1, nu2, w, s, false));
    return copy(dst, L0.v, n, s);
}

// Mixed-precision refinement: the iterate is the pair dst + lo of FP32 arrays.
// Each cycle forms r = f - A (dst + lo) in FP64 (into the finest level's rhs)
// with its FP64 norm, stops once ||r|| <= tol ||f||, else runs one V-cycle on
// A e = r from e = 0 and adds e to the pair.
// r = f - A (hi + lo) in FP64 into the finest level's rhs (FP32), its FP64
// sums copied to `host` and norm_ready recorded after the copy.
int residual64_enqueue(whhd_mgfrac3d_s* h, const float* hi, const float* lo, const float* f,
                       hipStream_t s, bool read_back = true) {
    const Level& L0 = h->lv[0];
    const LevelArgs a = args_of(h, L0);
    const dim3 g((L0.d.y + 63) / 64, (L0.d.x + 3) / 4, (L0.d.z + kZPlanes - 1) / kZPlanes);
    const int nparts = (int)(g.x * g.y * g.z);
    const int nx = L0.d.x, ny = L0.d.y;
    const Factors64 k{h->fac64, h->fac64 + ny, h->fac64 + 2 * ny, h->fac64 + 2 * ny + nx,
                      h->fac64 + 2 * (ny + nx), h->fac64 + 2 * (ny + nx) + L0.d.z};
    if (h->bc == 'z')
        hipLaunchKernelGGL(residual64_kernel<true>, g, dim3(64, 4), 0, s, L0.f, hi, lo, f, a, k, h->rpart);
    else
        hipLaunchKernelGGL(residual64_kernel<false>, g, dim3(64, 4), 0, s, L0.f, hi, lo, f, a, k, h->rpart);
    if (read_back)
        hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(1024), 0, s, h->rpart + 2 * h->rblocks,
                           h->rpart, nparts);
    hipError_t e = launch_status("mg_residual64_kernel");
    if (e != hipSuccess) return hip_fail(e, "mg FP64 residual kernel");
    if (!read_back) return WHHD_OK;
    e = hipMemcpyAsync(h->host, h->rpart + 2 * h->rblocks, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(h->norm_ready, s);
    if (e != hipSuccess) return hip_fail(e, "mgfrac3d: residual read-back");
    return WHHD_OK;
}

// ||r|| / ||f|| of the last residual64_enqueue, once its copy is done.
int residual64_wait(whhd_mgfrac3d_s* h, double* rel) {
    const hipError_t e = hipEventSynchronize(h->norm_ready);
    if (e != hipSuccess) return hip_fail(e, "mgfrac3d: residual read-back");
    *rel = h->host[1] > 0 ? std::sqrt(h->host[0]) / std::sqrt(h->host[1]) : 0.0;
    return WHHD_OK;
}

// Both: waits for the norms.
int residual64(whhd_mgfrac3d_s* h, const float* hi, const float* lo, const float* f, hipStream_t s,
               double* rel) {
    TRY(residual64_enqueue(h, hi, lo, f, s));
    return residual64_wait(h, rel);
}

int mgfrac3d_solve_tol(whhd_mgfrac3d_s* h, float* dst, float* lo, const float* pre_v,
                       const float* f, int max_cycles, double tol, int nu1, int nu2, float w,
                       int* cycles, double* rel, hipStream_t s) {
    if (!h || !dst || !lo || !f) return fail(WHHD_INVALID_ARG, "mgfrac3d_solve_tol: null argument");
    if (dst == lo || lo == f || dst == f)
        return fail(WHHD_INVALID_ARG, "mgfrac3d_solve_tol: dst, dst_lo and f must be distinct");
    if (max_cycles < 0 || nu1 < 0 || nu2 < 0 || !(tol >= 0))
        return fail(WHHD_INVALID_ARG, "mgfrac3d_solve_tol: bad iteration limits");
    const int64_t n = cells(h->dim);
    // the V-cycles read the finest level's FP16 fractions while this solve runs
    struct Fp16Fine {
        whhd_mgfrac3d_s* h;
        ~Fp16Fine() { h->fp16_fine = false; }
    } fp16_guard{h};
    h->fp16_fine = true;
    if (pre_v && pre_v != dst) TRY(copy(dst, pre_v, n, s));
    else if (!pre_v) TRY(fill(dst, 0.f, n, s));
    TRY(fill(lo, 0.f, n, s));
    double res = 0.0;
    int c = 0, last_k = 1;
    for (;; ++c) {
        const double prev = res;
        TRY(residual64(h, dst, lo, f, s, &res));
        if (res <= tol || c >= max_cycles) break;
        // Up to kMaxK V-cycles per FP64 residual (all but the first on the
        // FP32 residual of the correction equation, exact enough while e is a
        // correction): as many as the observed per-cycle rate says reach tol,
        // two before a rate is known, never past the budget.
        constexpr int kMaxK = 4;
        int k = 2;
        if (c > 0 && prev > res && res > 0) {
            const double rate = std::pow(res / prev, 1.0 / last_k);
            k = (int)std::ceil(std::log(tol / res) / std::log(rate));
        }
        k = std::max(1, std::min(std::min(k, kMaxK), max_cycles - c));
        for (int j = 0; j < k; ++j)
            TRY(j + 1 < k ? vcycle(h, 0, nu1, nu2, w, s, j == 0) : vcycle(h, 0, nu1, nu2, w, s, j == 0, dst, lo));
        c += k - 1;
        last_k = k;
    }
    if (cycles) *cycles = c;
    if (rel) *rel = res;
    return res <= tol ? WHHD_OK : WHHD_NONCONVERGED;
}

// Conjugate gradients on A x = f preconditioned by one V-cycle from zero
// (symmetric for nu1 = nu2: weighted-Jacobi sweeps around the coarse
// correction, spacing-weighted averaging restriction, injection
// prolongation, a fixed-sweep coarsest solve), x the pair hi + lo, every
// residual formed afresh in FP64.
int mgfrac3d_solve_pcg(whhd_mgfrac3d_s* h, float* dst, float* lo, const float* pre_v, const float* f,
                       int max_iters, double tol, int nu1, int nu2, float w, int* iters, double* rel,
                       hipStream_t s) {
    if (!h || !dst || !lo || !f) return fail(WHHD_INVALID_ARG, "mgfrac3d_solve_pcg: null argument");
    if (dst == lo || lo == f || dst == f)
        return fail(WHHD_INVALID_ARG, "mgfrac3d_solve_pcg: dst, dst_lo and f must be distinct");
    if (max_iters < 0 || nu1 < 0 || nu2 < 0 || std::isnan(tol))
        return fail(WHHD_INVALID_ARG, "mgfrac3d_solve_pcg: bad iteration limits");
    Level& L0 = h->lv[0];
    const int64_t n = cells(h->dim);
    // the preconditioning V-cycles read the finest level's FP16 fractions and
    // run 10 coarsest sweeps instead of 50 for bc 'n' (at 192^3 the same 12
    // iterations, 4.89 -> 4.80 ms): the preconditioner need not be the
    // reference's V-cycle, only a fixed symmetric one
    struct Precond {
        whhd_mgfrac3d_s* h;
        ~Precond() {
            h->fp16_fine = false;
            h->coarsest_n = 50;
        }
    } precond_guard{h};
    h->fp16_fine = true;
    h->coarsest_n = 10;
    if (pre_v && pre_v != dst) TRY(copy(dst, pre_v, n, s));
    else if (!pre_v) TRY(fill(dst, 0.f, n, s));
    TRY(fill(lo, 0.f, n, s));
    TRY(fill(h->pcg_p[0], 0.f, n, s));  // beta * p_old reads it at the first iteration
    // the direction's product with A uses the FP32 fractions (the operator
    // the FP64 residual measures)
    LevelArgs a = args_of(h, L0);
    a.axh = a.ayh = a.azh = nullptr;
    const dim3 gd((L0.d.y + 63) / 64, (L0.d.x + 3) / 4, (L0.d.z + kZPlanes - 1) / kZPlanes);
    const int dparts = (int)(gd.x * gd.y * gd.z);
    // z = M r, the new direction and alpha for iteration c (nothing read
    // back; first: beta = 0, p = z)
    auto direction = [&](int c, bool first) -> int {
        TRY(vcycle(h, 0, nu1, nu2, w, s, true));  // z = M r in L0.v
        hipLaunchKernelGGL(dot_partial_kernel, dim3(kNormBlocks), dim3(256), 0, s, h->part, L0.f, L0.v, n);
        hipLaunchKernelGGL(pcg_beta_kernel, dim3(1), dim3(1024), 0, s, h->coef, h->part, kNormBlocks,
                           first ? 1 : 0);
        float* pn = h->pcg_p[(c + 1) & 1];
        const float* po = h->pcg_p[c & 1];
        if (h->bc == 'z')
            hipLaunchKernelGGL(pcg_direction_kernel<true>, gd, dim3(64, 4), 0, s, pn, h->r, L0.v, po, a,
                               h->coef, h->rpart);
        else
            hipLaunchKernelGGL(pcg_direction_kernel<false>, gd, dim3(64, 4), 0, s, pn, h->r, L0.v, po, a,
                               h->coef, h->rpart);
        hipLaunchKernelGGL(pcg_alpha_kernel, dim3(1), dim3(1024), 0, s, h->coef, h->rpart, dparts);
        const hipError_t e = launch_status("mg_pcg_kernels");
        return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg conjugate-gradient kernels");
    };
    // x += alpha p, r -= alpha Ap (the recurrence residual in L0.f), partial r.r
    auto update = [&](int c) -> int {
        hipLaunchKernelGGL(pcg_update_kernel, dim3(kNormBlocks), dim3(256), 0, s, dst, lo, L0.f,
                           h->pcg_p[(c + 1) & 1], h->r, h->coef, h->part, n);
        const hipError_t e = launch_status("mg_pcg_update_kernel");
        return e == hipSuccess ? WHHD_OK : hip_fail(e, "mg conjugate-gradient update");
    };
    if (tol < 0) {
        // exactly max_iters iterations and nothing read back, so that the
        // whole solve can be captured in a graph
        // (the FP64 residual every iteration: the count may run past
        // convergence, where the recurrence no longer describes the iterate)
        for (int c = 0; c < max_iters; ++c) {
            TRY(residual64_enqueue(h, dst, lo, f, s, false));
            TRY(direction(c, c == 0));
            TRY(update(c));
        }
        if (iters) *iters = max_iters;
        if (rel) *rel = -1.0;
        return WHHD_OK;
    }
    // While the residual is far above tol it follows the recurrence
    // r <- r - alpha Ap, updated with the pair in one pass (its drift from the
    // true residual, ~1e-7 of ||f|| from FP32 rounding, is then negligible);
    // from max(1e-5, 10 tol) on, every iteration forms the FP64 residual of
    // the pair afresh (below its drift the recurrence stops describing the
    // iterate and the iteration would wander), restarting the directions once
    // at the switch.  The FP64 residual always decides convergence.
    const double exact_below = std::max(1e-5, 10.0 * tol);
    double res = 0.0;
    TRY(residual64(h, dst, lo, f, s, &res));
    const double fnorm2 = h->host[1];
    bool exact = res <= exact_below;
    int c = 0, restart = 0;  // restart: an iteration whose direction starts afresh (beta = 0)
    while (res > tol && c < max_iters) {
        TRY(direction(c, c == restart));
        TRY(update(c));
        ++c;
        if (exact) {
            TRY(residual64(h, dst, lo, f, s, &res));
            continue;
        }
        hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(1024), 0, s, h->part + kNormBlocks, h->part,
                           kNormBlocks);
        hipError_t e = launch_status("mg_sum_final_kernel");
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->host, h->part + kNormBlocks, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(h->norm_ready, s);
        if (e == hipSuccess) e = hipEventSynchronize(h->norm_ready);
        if (e != hipSuccess) return hip_fail(e, "mgfrac3d_solve_pcg: read-back");
        const double rec = fnorm2 > 0 ? std::sqrt(h->host[0] / fnorm2) : 0.0;
        if (rec <= exact_below || c >= max_iters) {
            TRY(residual64(h, dst, lo, f, s, &res));
            exact = true;
            restart = c;
        } else {
            res = rec;
        }
    }
    if (iters) *iters = c;
    if (rel) *rel = res;
    return res <= tol ? WHHD_OK : WHHD_NONCONVERGED;
}

int mgfrac3d_residual(whhd_mgfrac3d_s* h, const float* v, const float* f, double* rel,
                      double* sums, hipStream_t s) {
    if (!h || !v || !f) return fail(WHHD_INVALID_ARG, "mgfrac3d_residual: null argument");
    const Level& L0 = h->lv[0];
    const int64_t n = cells(h->dim);
    TRY(residual(h, L0, h->r, v, f, s));
    hipLaunchKernelGGL(norm_partial_kernel, dim3(kNormBlocks), dim3(256), 0, s, h->part, h->r, f, n);
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(1024), 0, s, h->part + 2 * kNormBlocks,
                       h->part, kNormBlocks);
    hipError_t e = launch_status("mg_norm_kernels");
    if (e != hipSuccess) return hip_fail(e, "mg norm kernels");
    e = hipMemcpyAsync(h->host, h->part + 2 * kNormBlocks, 2 * sizeof(double),
                       hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "mgfrac3d_residual: read-back");
    if (sums) {
        sums[0] = h->host[0];
        sums[1] = h->host[1];
    }
    if (rel) *rel = h->host[1] > 0 ? std::sqrt(h->host[0]) / std::sqrt(h->host[1]) : 0.0;
    return WHHD_OK;
}

}  // namespace whhd

extern "C" {

WHHD_API int whhd_mgfrac3d_create(const float* ax, const float* ay, const float* az, int nx,
                                  int ny, int nz, double dx, double dy, double dz, char bc,
                                  whhd_mgfrac3d_t* out) {
    return whhd::mgfrac3d_create(ax, ay, az, nx, ny, nz, dx, dy, dz, bc, out);
}

WHHD_API int whhd_mgfrac3d_levels(whhd_mgfrac3d_t h, int* dims, int max_levels) {
    if (!h) return -1;
    const int n = (int)h->lv.size();
    for (int l = 0; l < n && l < max_levels && dims; ++l) {
        dims[3 * l] = h->lv[l].d.x;
        dims[3 * l + 1] = h->lv[l].d.y;
        dims[3 * l + 2] = h->lv[l].d.z;
    }
    return n;
}

WHHD_API int whhd_mgfrac3d_solve(whhd_mgfrac3d_t h, float* dst, const float* pre_v, const float* f,
                                 int n_iters, int nu1, int nu2, float weight, void* stream) {
    return whhd::mgfrac3d_solve(h, dst, pre_v, f, n_iters, nu1, nu2, weight, (hipStream_t)stream);
}

WHHD_API int whhd_mgfrac3d_solve_tol(whhd_mgfrac3d_t h, float* dst, float* dst_lo,
                                     const float* pre_v, const float* f, int max_cycles,
                                     double tol, int nu1, int nu2, float weight, int* cycles,
                                     double* rel, void* stream) {
    return whhd::mgfrac3d_solve_tol(h, dst, dst_lo, pre_v, f, max_cycles, tol, nu1, nu2, weight,
                                    cycles, rel, (hipStream_t)stream);
}

WHHD_API int whhd_mgfrac3d_solve_pcg(whhd_mgfrac3d_t h, float* dst, float* dst_lo,
                                     const float* pre_v, const float* f, int max_iters,
                                     double tol, int nu1, int nu2, float weight, int* iters,
                                     double* rel, void* stream) {
    return whhd::mgfrac3d_solve_pcg(h, dst, dst_lo, pre_v, f, max_iters, tol, nu1, nu2, weight,
                                    iters, rel, (hipStream_t)stream);
}

WHHD_API int whhd_mgfrac3d_residual(whhd_mgfrac3d_t h, const float* v, const float* f,
                                    double* rel, void* stream) {
    return whhd::mgfrac3d_residual(h, v, f, rel, nullptr, (hipStream_t)stream);
}

WHHD_API void whhd_mgfrac3d_destroy(whhd_mgfrac3d_t h) { whhd::release(h); }

}  // extern "C"
