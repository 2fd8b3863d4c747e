// Sweep + DST comparator: a vector potential by integrating the velocity along
// z, made divergence free by a DST-I Poisson solve.
//   CuParallelSweep3D::Solve   WaveletPotentials/cuParallelSweep3D.cu:9-66, 939-965
//   CuParallelSweep3D::Project WaveletPotentials/cuParallelSweep3D.cu:1054-1079, 1316-1326
//   (velocity bc 'n': ProjectZero on the interior nodes; 'd' and 'z': ProjectD
//   on every node of the potential grid)
//   CuDst3D::Solve             WaveletPotentials/cuDst3D.cu:183-341
// The reference sweeps one column per thread with cuFFT for the DST.  Here the
// sweep is the same column walk (y-fastest, so each z step of a wavefront is a
// coalesced row).  The DST-I along an axis of n samples with n + 1 a power of
// two is a half-length complex FFT over tiles of lines staged in LDS, in place,
// the line axis strided through memory as the array lays it out (no
// transposes); other lengths use a dense product with the orthonormal sine
// matrix (an LDS-tiled kernel over 64 x 64 output tiles).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "whhd_internal.h"

namespace whhd {
namespace {

// One thread per node column (x <= nx, y <= ny): the qx column (x < nx) and qy
// column (y < ny) are running sums along z (SweepingQx / SweepingQy), the qz
// column is cleared.  qy starts from zero, or with qy_base from the bottom
// plane already in qy (velocity bc 'd').
template <class T>
__global__ void sweep_kernel(T* __restrict__ qx, T* __restrict__ qy,
                             T* __restrict__ qz, const T* __restrict__ xv,
                             const T* __restrict__ yv, int nx, int ny, int nz, T dz,
                             int qy_base) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    if (y > ny || x > nx) return;
    if (x < nx) {  // qx (nx, ny+1, nz+1) from yv (nx, ny+1, nz)
        const int64_t s = (int64_t)nx * (ny + 1);
        const int64_t b = y + (int64_t)x * (ny + 1);
        T acc = T(0);
        qx[b] = T(0);
#pragma unroll 8
        for (int z = 1; z <= nz; ++z) {
            acc += yv[b + (z - 1) * s] * dz;
            qx[b + z * s] = acc;
        }
    }
    if (y < ny) {  // qy (nx+1, ny, nz+1) from xv (nx+1, ny, nz)
        const int64_t s = (int64_t)(nx + 1) * ny;
        const int64_t b = y + (int64_t)x * ny;
        T acc = qy_base ? qy[b] : T(0);
        qy[b] = acc;
#pragma unroll 8
        for (int z = 1; z <= nz; ++z) {
            acc -= xv[b + (z - 1) * s] * dz;
            qy[b + z * s] = acc;
        }
    }
    {  // qz (nx+1, ny+1, nz) = 0
        const int64_t s = (int64_t)(nx + 1) * (ny + 1);
        const int64_t b = y + (int64_t)x * (ny + 1);
        for (int z = 0; z < nz; ++z) qz[b + z * s] = T(0);
    }
}

// Top gauge for velocity bc 'n' (cuSweepingQ_x and the clears that follow it):
// per y, qp(x) = sum_{x' < x} qx(x', y, top) * dx; the top planes of qx and qy
// become zero and the top plane of qz takes -qp / dz.  A workgroup takes 64
// consecutive y (one per lane, coalesced rows) and splits x over its eight
// wavefronts: each sums its run of x, the runs' offsets come through LDS, and
// each then walks its run again writing qz.
constexpr int kGaugeWaves = 8;

template <class T>
__global__ void __launch_bounds__(64 * kGaugeWaves) top_gauge_kernel(T* __restrict__ qx, T* __restrict__ qy,
                                                                     T* __restrict__ qz, int nx, int ny, int nz,
                                                                     T dx, T dz) {
    __shared__ T part[kGaugeWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int y = blockIdx.x * 64 + lane;
    const bool ok = y <= ny;
    const int64_t tqx = (int64_t)nz * nx * (ny + 1);              // top plane of qx
    const int64_t tqy = (int64_t)nz * (nx + 1) * ny;              // top plane of qy
    const int64_t tqz = (int64_t)(nz - 1) * (nx + 1) * (ny + 1);  // top plane of qz
    const int run = (nx + kGaugeWaves - 1) / kGaugeWaves;
    const int xa = 1 + w * run, xb = min(nx, (w + 1) * run);      // x in [xa, xb]
    T sum = T(0);
    if (ok) {
#pragma unroll 8
        for (int x = xa; x <= xb; ++x) sum += qx[tqx + y + (int64_t)(x - 1) * (ny + 1)] * dx;
    }
    part[w][lane] = sum;
    __syncthreads();
    T qp = T(0);
    for (int k = 0; k < w; ++k) qp += part[k][lane];
    if (ok) {
#pragma unroll 8
        for (int x = xa; x <= xb; ++x) {
            T* e = qx + tqx + y + (int64_t)(x - 1) * (ny + 1);
            qp += *e * dx;
            *e = T(0);
            qz[tqz + y + (int64_t)x * (ny + 1)] -= qp / dz;
        }
    }
    if (y < ny)
        for (int x = w; x <= nx; x += kGaugeWaves) qy[tqy + y + (int64_t)x * ny] = T(0);
}

// Velocity bc 'd' bottom plane (SweepingQy_x, WaveletPotentials/cuParallelSweep3D.cu:
// 68-85): per y, qy(x, y, 0) = sum_{x' < x} zv(x', y, 0) dx.
template <class T>
__global__ void bottom_qy_kernel(T* __restrict__ qy, const T* __restrict__ zv, int nx,
                                 int ny, T dx) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ny) return;
    T acc = T(0);
    qy[y] = T(0);
    for (int x = 1; x <= nx; ++x) {
        acc += zv[y + (int64_t)(x - 1) * ny] * dx;
        qy[y + (int64_t)x * ny] = acc;
    }
}

// Velocity bc 'd' divergence (cuGetB_q_d_3d, WaveletPotentials/cuConvergence.cu:
// 161-200): b = -div q at every node (nx+1, ny+1, nz+1); at a wall the edge
// index is clamped, so the missing edge repeats the wall edge and that axis
// contributes nothing.
template <class T>
__global__ void node_div_d_kernel(T* __restrict__ b, const T* __restrict__ qx,
                                  const T* __restrict__ qy, const T* __restrict__ qz,
                                  int nx, int ny, int nz, T ihx, T ihy, T ihz) {
    const int i = blockIdx.x * 64 + threadIdx.x;  // y
    const int j = blockIdx.y * 4 + threadIdx.y;   // x
    const int k = blockIdx.z;                     // z
    if (i > ny || j > nx) return;
    const int64_t NX = nx, NY = ny;
    const int jl = max(j - 1, 0), jr = min(j, nx - 1);
    const int ib = max(i - 1, 0), it = min(i, ny - 1);
    const int kf = max(k - 1, 0), kb = min(k, nz - 1);
    auto QX = [&](int x) { return qx[i + x * (NY + 1) + k * NX * (NY + 1)]; };
    auto QY = [&](int yy) { return qy[yy + j * NY + k * (NX + 1) * NY]; };
    auto QZ = [&](int z) { return qz[i + j * (NY + 1) + z * (NX + 1) * (NY + 1)]; };
    const T d = (QX(jr) - QX(jl)) * ihx + (QY(it) - QY(ib)) * ihy + (QZ(kb) - QZ(kf)) * ihz;
    b[i + j * (NY + 1) + k * (NX + 1) * (NY + 1)] = -d;
}

// q -= grad phi over every edge between two nodes (cuGradientMinus3D_d_Q,
// WaveletPotentials/cuGradient.cu:2408-2446); phi is (nx+1, ny+1, nz+1).
template <class T>
__global__ void grad_minus_d_kernel(T* __restrict__ qx, T* __restrict__ qy,
                                    T* __restrict__ qz, const T* __restrict__ phi, int nx,
                                    int ny, int nz, T ihx, T ihy, T ihz) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y > ny || x > nx) return;
    const int64_t NX = nx, NY = ny;
    const int64_t sp = (NX + 1) * (NY + 1);
    const int64_t ip = y + x * (NY + 1) + z * sp;
    const T p = phi[ip];
    if (x < nx) qx[y + x * (NY + 1) + z * NX * (NY + 1)] -= (phi[ip + NY + 1] - p) * ihx;
    if (y < ny) qy[y + x * NY + z * (NX + 1) * NY] -= (phi[ip + 1] - p) * ihy;
    if (z < nz) qz[ip] -= (phi[ip + sp] - p) * ihz;
}

// 2D sweep, CuParallelSweep3D::Solve(qz, xv, yv) (WaveletPotentials/cuParallelSweep3D.cu:
// 982-995): one thread per y walks x, qz(x, y) = qz(x-1, y) - yv(x-1, y) dx
// (SweepingQz_x).  bc 'n' starts from qz(0, y) = 0 (SweepingQz_x_0); bc 'd'
// starts from the integral of xv along y at x = 0 (SweepingQz_y), summed in the
// same order here, with qz(0, 0) = 0 where the reference keeps the array's value.
template <class T>
__global__ void sweep2d_kernel(T* __restrict__ qz, const T* __restrict__ xv,
                               const T* __restrict__ yv, int nx, int ny, T dx, T dy,
                               int dbc) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y > ny) return;
    T acc = T(0);
    if (dbc)
        for (int k = 1; k <= y; ++k) acc = acc + xv[k - 1] * dy;
    qz[y] = acc;
    for (int x = 1; x <= nx; ++x) {
        acc = acc - yv[y + (int64_t)(x - 1) * (ny + 1)] * dx;
        qz[y + (int64_t)x * (ny + 1)] = acc;
    }
}

// b = -div q at the interior nodes (GetB_3d_q_zeros); b is (nx-1, ny-1, nz-1).
template <class T>
__global__ void node_div_kernel(T* __restrict__ b, const T* __restrict__ qx,
                                const T* __restrict__ qy, const T* __restrict__ qz, int nx,
                                int ny, int nz, T ihx, T ihy, T ihz) {
    const int i = blockIdx.x * 64 + threadIdx.x;  // y - 1
    const int j = blockIdx.y * 4 + threadIdx.y;   // x - 1
    const int k = blockIdx.z;                     // z - 1
    const int py = ny - 1, px = nx - 1;
    if (i >= py || j >= px) return;
    const int64_t iqx = (i + 1) + (int64_t)j * (ny + 1) + (int64_t)(k + 1) * nx * (ny + 1);
    const int64_t iqy = i + (int64_t)(j + 1) * ny + (int64_t)(k + 1) * (nx + 1) * ny;
    const int64_t iqz = (i + 1) + (int64_t)(j + 1) * (ny + 1) + (int64_t)k * (nx + 1) * (ny + 1);
    const T d = (qx[iqx + (ny + 1)] - qx[iqx]) * ihx + (qy[iqy + 1] - qy[iqy]) * ihy +
                    (qz[iqz + (int64_t)(nx + 1) * (ny + 1)] - qz[iqz]) * ihz;
    b[i + (int64_t)j * py + (int64_t)k * px * py] = -d;
}

// q -= grad phi with phi = 0 on the boundary nodes (GradientMinus3D_zero_Q):
// one thread per node (x <= nx, y <= ny, z <= nz), owning the qx, qy, qz edges
// that start at it.
template <class T>
__device__ __forceinline__ T phi_at(const T* __restrict__ phi, int x, int y, int z, int nx,
                                        int ny, int nz) {
    if (x < 1 || y < 1 || z < 1 || x > nx - 1 || y > ny - 1 || z > nz - 1) return T(0);
    return phi[(y - 1) + (int64_t)(x - 1) * (ny - 1) + (int64_t)(z - 1) * (nx - 1) * (ny - 1)];
}

template <class T>
__global__ void grad_minus_kernel(T* __restrict__ qx, T* __restrict__ qy,
                                  T* __restrict__ qz, const T* __restrict__ phi, int nx,
                                  int ny, int nz, T ihx, T ihy, T ihz) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y > ny || x > nx) return;
    const T p = phi_at(phi, x, y, z, nx, ny, nz);
    if (x < nx)
        qx[y + (int64_t)x * (ny + 1) + (int64_t)z * nx * (ny + 1)] -=
            (phi_at(phi, x + 1, y, z, nx, ny, nz) - p) * ihx;
    if (y < ny)
        qy[y + (int64_t)x * ny + (int64_t)z * (nx + 1) * ny] -=
            (phi_at(phi, x, y + 1, z, nx, ny, nz) - p) * ihy;
    if (z < nz)
        qz[y + (int64_t)x * (ny + 1) + (int64_t)z * (nx + 1) * (ny + 1)] -=
            (phi_at(phi, x, y, z + 1, nx, ny, nz) - p) * ihz;
}

// out(i, k, b) = sum_j in(i, j, b) * S(k, j) for k, j < n, i < m, b < gridDim.z;
// element (i, j, b) at i*si + j*sj + b*sb in both arrays.  A workgroup computes
// a 64 (i) x 64 (k) tile, each thread 4 x 4, with 16-wide j slices staged in LDS.
constexpr int kTI = 64, kTK = 64, kTJ = 16;

template <class T>
__global__ void __launch_bounds__(256) dst_pass_kernel(T* __restrict__ out,
                                                       const T* __restrict__ in,
                                                       const T* __restrict__ S, int n, int m,
                                                       int64_t si, int64_t sj, int64_t sb) {
    __shared__ T As[kTJ][kTI + 1];
    __shared__ T Bs[kTJ][kTK + 1];
    const int i0 = blockIdx.x * kTI, k0 = blockIdx.y * kTK;
    const int64_t boff = (int64_t)blockIdx.z * sb;
    const int t = threadIdx.x;
    const int ti = (t & 15) * 4, tk = (t >> 4) * 4;
    T acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = T(0);
    const bool jfast = (sj == 1);
    for (int j0 = 0; j0 < n; j0 += kTJ) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = t + e * 256;
            const int ii = jfast ? idx / kTJ : idx % kTI;
            const int jj = jfast ? idx % kTJ : idx / kTI;
            const int gi = i0 + ii, gj = j0 + jj;
            As[jj][ii] = (gi < m && gj < n) ? in[boff + gi * si + gj * sj] : T(0);
            const int kk = idx / kTJ, jb = idx % kTJ;
            const int gk = k0 + kk, gjb = j0 + jb;
            Bs[jb][kk] = (gk < n && gjb < n) ? S[(int64_t)gk * n + gjb] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < kTJ; ++jj) {
            T a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[jj][ti + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = Bs[jj][tk + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * b[c];
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gi = i0 + ti + r, gk = k0 + tk + c;
            if (gi < m && gk < n) out[boff + gi * si + gk * sj] = acc[r][c];
        }
}

// ---- DST-I through a half-length complex FFT ---------------------------------
// For a line of n samples with M = n + 1 a power of two (4 <= M <= 1024), the
// orthonormal DST-I X_j = c sum_i x_i sin(pi (i+1)(j+1) / M), c = sqrt(2/M), is
// F_{j+1} c with F_k = sum_{m=1}^{M-1} f_m sin(pi m k / M), f_m = x_{m-1}, f_0 =
// f_M = 0, computed in O(M log M) (the odd/even split of the sine transform
// through a real DFT of length M):
//   g_m = sin(pi m / M) (f_m + f_{M-m}) + (f_m - f_{M-m}) / 2, m = 0..M-1;
//   G = DFT_M(g) from the P = M/2 point complex FFT of h_t = g_{2t} + i g_{2t+1}:
//       G_k = E_k + e^{-2 pi i k / M} O_k, E_k = (H_k + conj H_{P-k}) / 2,
//       O_k = (H_k - conj H_{P-k}) / (2i);
//   F_{2k} = -Im G_k (k >= 1), F_1 = Re G_0 / 2, F_{2k+1} = F_{2k-1} + Re G_k.
// (the sine-symmetric part of g carries the odd outputs, the antisymmetric
// part the even ones).  A workgroup stages a tile of lines in LDS with
// coalesced accesses (rows of consecutive y for the x and z axes, whole
// contiguous lines for y), each half-wavefront transforms one line at a time
// (an in-place radix-2 FFT over its lanes, a prefix sum for the odd outputs),
// and the tile is written back.  The transform is its own inverse.
//
// Per-axis table (T): sin(pi m / M) for m < M, then e^{-2 pi i t / P} for
// t < P/2 (re, im), then e^{-2 pi i k / M} for k < P (re, im).
constexpr int kFftThreads = 256;

template <class T>
struct Cx {
    T r, i;
};

// inclusive prefix sum over each half (32 lanes) of the wavefront.  FP32: DPP
// row shifts by 1, 2, 4, 8 within each row of 16 lanes (lanes without a source
// add 0), then row_bcast:15 adds the last lane of rows 0 and 2 to rows 1 and 3
// (row mask 0xA keeps the two halves apart); FP64 through lane shuffles.
template <class T>
__device__ __forceinline__ T half_scan(T v, int sl) {
    if constexpr (sizeof(T) == 4) {
        auto step = [&](auto ctrl, auto rmask, auto bound) {
            const int u = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), decltype(ctrl)::value,
                                                      decltype(rmask)::value, 0xF, decltype(bound)::value);
            v += __builtin_bit_cast(float, u);
        };
        step(std::integral_constant<int, 0x111>{}, std::integral_constant<int, 0xF>{}, std::true_type{});
        step(std::integral_constant<int, 0x112>{}, std::integral_constant<int, 0xF>{}, std::true_type{});
        step(std::integral_constant<int, 0x114>{}, std::integral_constant<int, 0xF>{}, std::true_type{});
        step(std::integral_constant<int, 0x118>{}, std::integral_constant<int, 0xF>{}, std::true_type{});
        step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{}, std::false_type{});
        return v;
    } else {
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const T u = __shfl_up(v, o, 32);
            if (sl >= o) v += u;
        }
        return v;
    }
}

// lines: li = inner + outer * n_inner; line li starts at inner * s_inner +
// outer * s_outer and its samples are `step` apart; a tile is `tl` lines
// li = tile * tl + l.  A tile is read whole before it is written, and no other
// tile touches its lines, so `out` may be `in`.
// With `lam_s` set, the outputs are divided by the eigenvalue sum of their
// node: lam_q[li / lam_div] + lam_r[li % lam_div] for line li plus lam_s[j]
// for sample j (the Poisson solve's scaling, fused into the last forward pass).
// reps = 2 transforms each staged line twice (the solve's last forward pass,
// the scaling and its first inverse pass share the axis and the tile).
struct DstScale {
    const void* lam_q;
    const void* lam_r;
    const void* lam_s;
    int64_t lam_div;
};

template <class T, int LOGM>
__global__ void __launch_bounds__(kFftThreads) dst_fft_kernel(T* out, const T* in,
                                                              const T* __restrict__ tab, int n,
                                                              int64_t lines, int64_t n_inner, int64_t s_inner,
                                                              int64_t s_outer, int64_t step, int tl, T c,
                                                              DstScale sc, int reps) {
    extern __shared__ unsigned char dst_smem[];
    constexpr int M = 1 << LOGM, P = M >> 1, rs = M + 1;  // row stride: conflict-free strided rows
    T* rows = (T*)dst_smem;                                  // tl x rs
    Cx<T>* scr = (Cx<T>*)(rows + (size_t)tl * rs);           // 4 waves x 2 lines x P
    int64_t* bases = (int64_t*)(scr + 8 * P);                // tl line starts, -1 past the end
    T* lsum = (T*)(bases + tl);                              // tl eigenvalue sums (scaling)
    T* ltab = lsum + tl;                                     // the axis table (2.5 M values)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // workgroups go to the 8 XCDs round-robin; each XCD takes a contiguous run
    // of tiles, so the cache lines that neighbouring tiles of strided lines
    // share (tl < 32 floats of a row) go through one L2
    const unsigned nwg = gridDim.x, bxr = blockIdx.x % 8u, bxq = blockIdx.x / 8u;
    const unsigned tile = bxr * (nwg / 8u) + (bxr < nwg % 8u ? bxr : nwg % 8u) + bxq;
    const int64_t li0 = (int64_t)tile * tl;
    const T* lam_s = (const T*)sc.lam_s;
    if (tid < tl) {
        const int64_t li = li0 + tid;
        const int64_t o = li / n_inner, i = li - o * n_inner;
        bases[tid] = li < lines ? i * s_inner + o * s_outer : -1;
        if (lam_s && li < lines) {
            const int64_t q = li / sc.lam_div, r = li - q * sc.lam_div;
            lsum[tid] = ((const T*)sc.lam_q)[q] + ((const T*)sc.lam_r)[r];
        }
    }
    for (int e = tid; e < M + 3 * P; e += kFftThreads) ltab[e] = tab[e];
    __syncthreads();
    // ---- stage the tile: row slots 0 and M hold f_0 = f_M = 0, slots 1..n the line.
    // Every load comes from a valid index (slots outside 1..n and lines past the
    // end read the tile's first line and store 0): no branch per load.
    const int64_t b0 = bases[0];
    if (step == 1) {
        for (int l = 0; l < tl; ++l) {
            const int64_t b = bases[l];
            const int64_t bb = b >= 0 ? b : b0;
#pragma unroll 4
            for (int j = tid; j <= M; j += kFftThreads) {
                const T v = in[bb + min(max(j - 1, 0), n - 1)];
                rows[l * rs + j] = (j >= 1 && j <= n && b >= 0) ? v : T(0);
            }
        }
    } else {
        // tl divides the workgroup: each thread keeps one line of the tile
        const int l = tid % tl, jstep = kFftThreads / tl;
        const int64_t b = bases[l];
        const int64_t bb = b >= 0 ? b : b0;
#pragma unroll 8
        for (int j = tid / tl; j <= M; j += jstep) {
            const T v = in[bb + (int64_t)min(max(j - 1, 0), n - 1) * step];
            rows[l * rs + j] = (j >= 1 && j <= n && b >= 0) ? v : T(0);
        }
    }
    __syncthreads();
    const T* sn = ltab;
    const Cx<T>* tw = (const Cx<T>*)(ltab + M);
    const Cx<T>* ek = tw + P / 2;
    // each wavefront transforms two lines at a time, one per half (32 lanes):
    // two independent chains of LDS round trips in flight
    const int half_id = lane >> 5, sl = lane & 31;
    Cx<T>* x = scr + (wave * 2 + half_id) * P;
    // Per-lane constants, the same for every line: the sine weights of the lane's
    // h values, the twiddles of its butterflies and the post-processing factors
    // (registers instead of an LDS read per use).  The radix-2 stages of the
    // in-place FFT run in pairs as radix-4 butterflies: stage sh pairs (j, j+half)
    // and (j+2half, j+3half) with twiddle wa, stage sh+1 pairs (j, j+2half) with
    // wb and (j+half, j+3half) with wb * W^(P/4) = -i wb, the same arithmetic as
    // the two radix-2 stages with half the LDS round trips.  An odd stage count
    // starts with the twiddle-free stage 0.
    constexpr int NS = LOGM - 1;                    // radix-2 stages
    constexpr int S0 = NS & 1;                      // a lone stage 0 first
    constexpr int NP4 = NS / 2;                     // radix-4 stage pairs
    constexpr int NB4 = P / 4 >= 32 ? P / 128 : 1;  // radix-4 butterflies per lane
    constexpr int per = P >= 32 ? P / 32 : 1;      // outputs k per lane
    // the first stage (stage 0 alone, or the first radix-4 pair, whose twiddles
    // are 1 and -i) runs on the h values in registers: a lane forms the G h
    // values that land on bit-reversed positions r .. r + G - 1 (t0 + 0, P/2,
    // and for G = 4 also P/4, 3P/4) and writes the butterflies' outputs
    constexpr int G = S0 ? 2 : 4;
    constexpr int NQ = P / G >= 32 ? P / G / 32 : 1;  // groups per lane
    T snr[NQ][G][2];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int t = sl + 32 * q + (u == 1 ? P / 2 : u == 2 ? P / 4 : u == 3 ? 3 * P / 4 : 0);
            const bool ok = sl + 32 * q < P / G;
            snr[q][u][0] = ok ? sn[2 * t] : T(0);
            snr[q][u][1] = ok ? sn[2 * t + 1] : T(0);
        }
    Cx<T> twr[NP4 > 0 ? NP4 : 1][NB4][2];
#pragma unroll
    for (int ps = 0; ps < NP4; ++ps) {
        const int sh = S0 + 2 * ps, half = 1 << sh;
#pragma unroll
        for (int q = 0; q < NB4; ++q) {
            const int bfl = sl + 32 * q, k = bfl & (half - 1);
            twr[ps][q][0] = bfl < P / 4 ? tw[k * (P >> (sh + 1))] : Cx<T>{T(1), T(0)};
            twr[ps][q][1] = bfl < P / 4 ? tw[k * (P >> (sh + 2))] : Cx<T>{T(1), T(0)};
        }
    }
    Cx<T> ekr[per];
#pragma unroll
    for (int q = 0; q < per; ++q) ekr[q] = sl * per + q < P ? ek[sl * per + q] : Cx<T>{T(0), T(0)};
    for (int l0 = wave * 2; l0 < tl; l0 += 2 * (kFftThreads / 64))
    for (int rep = 0; rep < reps; ++rep) {
        const int l = l0 + half_id;
        const bool live = l < tl;
        T* f = rows + (live ? l : l0) * rs;
        // h_t = g_{2t} + i g_{2t+1} in bit-reversed order, through the first stage
        auto bfly = [](Cx<T>& a, Cx<T>& b, const Cx<T>& w) {
            const T tr = w.r * b.r - w.i * b.i, ti = w.r * b.i + w.i * b.r;
            const Cx<T> s{a.r + tr, a.i + ti}, d{a.r - tr, a.i - ti};
            a = s;
            b = d;
        };
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int t0 = sl + 32 * q;
            if (t0 < P / G) {
                Cx<T> e[G];
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    const int t = t0 + (u == 1 ? P / 2 : u == 2 ? P / 4 : u == 3 ? 3 * P / 4 : 0);
                    {
                        const int m = 2 * t;
                        const T a = f[m], b = f[M - m];
                        e[u].r = snr[q][u][0] * (a + b) + T(0.5) * (a - b);
                    }
                    {
                        const int m = 2 * t + 1;
                        const T a = f[m], b = f[M - m];
                        e[u].i = snr[q][u][1] * (a + b) + T(0.5) * (a - b);
                    }
                }
                const int r = (int)(__brev((unsigned)t0) >> (33 - LOGM));
                if constexpr (G == 2) {
                    x[r] = Cx<T>{e[0].r + e[1].r, e[0].i + e[1].i};
                    x[r + 1] = Cx<T>{e[0].r - e[1].r, e[0].i - e[1].i};
                } else {
                    const Cx<T> one{T(1), T(0)}, mi{T(0), T(-1)};
                    bfly(e[0], e[1], one);
                    bfly(e[2], e[3], one);
                    bfly(e[0], e[2], one);
                    bfly(e[1], e[3], mi);
#pragma unroll
                    for (int u = 0; u < 4; ++u) x[r + u] = e[u];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
#pragma unroll
        for (int ps = S0 ? 0 : 1; ps < NP4; ++ps) {
            const int sh = S0 + 2 * ps, half = 1 << sh;
#pragma unroll
            for (int q = 0; q < NB4; ++q) {
                const int bfl = sl + 32 * q;
                if (bfl < P / 4) {
                    const int g = bfl >> sh, k = bfl & (half - 1);
                    const int j = g * 4 * half + k;
                    Cx<T> e0 = x[j], e1 = x[j + half], e2 = x[j + 2 * half], e3 = x[j + 3 * half];
                    const Cx<T> wa = twr[ps][q][0], wb = twr[ps][q][1];
                    const Cx<T> wc{wb.i, -wb.r};  // -i wb
                    bfly(e0, e1, wa);
                    bfly(e2, e3, wa);
                    bfly(e0, e2, wb);
                    bfly(e1, e3, wc);
                    x[j] = e0;
                    x[j + half] = e1;
                    x[j + 2 * half] = e2;
                    x[j + 3 * half] = e3;
                }
            }
            __builtin_amdgcn_wave_barrier();
            asm volatile("" ::: "memory");
        }
        // G_k -> outputs; lane holds the `per` consecutive k of its block
        const int k0 = sl * per;
        T run = T(0);
        T rk[per], ik[per];
#pragma unroll
        for (int q = 0; q < per; ++q) {
            const int k = k0 + q;
            rk[q] = T(0);
            ik[q] = T(0);
            if (k < P) {
                const Cx<T> hk = x[k], hm = x[(P - k) & (P - 1)];
                const T er = T(0.5) * (hk.r + hm.r), ei = T(0.5) * (hk.i - hm.i);
                const T dr = hk.r - hm.r, di = hk.i + hm.i;
                const T orr = T(0.5) * di, oi = T(-0.5) * dr;
                const Cx<T> e = ekr[q];
                rk[q] = er + e.r * orr - e.i * oi;
                ik[q] = ei + e.r * oi + e.i * orr;
                if (k == 0) rk[q] *= T(0.5);
            }
            run += rk[q];
        }
        T pre = half_scan(run, sl) - run;  // exclusive prefix of the odd-output sums
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        if (live) {
#pragma unroll
            for (int q = 0; q < per; ++q) {
                const int k = k0 + q;
                if (k < P) {
                    pre += rk[q];
                    T xe = c * pre;                         // X_{2k} = c F_{2k+1}
                    T xo = -c * ik[q];                      // X_{2k-1} = c F_{2k}
                    if (lam_s && rep == 0) {
                        if constexpr (sizeof(T) == 4) {
                            // reciprocal (within 1 ulp) instead of two IEEE divisions
                            xe *= __builtin_amdgcn_rcpf(lsum[l] + lam_s[2 * k]);
                            if (k >= 1) xo *= __builtin_amdgcn_rcpf(lsum[l] + lam_s[2 * k - 1]);
                        } else {
                            xe /= lsum[l] + lam_s[2 * k];
                            if (k >= 1) xo /= lsum[l] + lam_s[2 * k - 1];
                        }
                    }
                    f[1 + 2 * k] = xe;
                    if (k >= 1) f[2 * k] = xo;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    }
    __syncthreads();
    // ---- write the tile back (slots 1..n)
    if (step == 1) {
        for (int l = 0; l < tl; ++l) {
            const int64_t b = bases[l];
            if (b < 0) break;
            for (int j = tid; j < n; j += kFftThreads) out[b + j] = rows[l * rs + 1 + j];
        }
    } else {
        const int l = tid % tl, jstep = kFftThreads / tl;
        const int64_t b = bases[l];
        if (b >= 0)
            for (int j = tid / tl; j < n; j += jstep) out[b + (int64_t)j * step] = rows[l * rs + 1 + j];
    }
}

// phi_hat /= lambda_x + lambda_y + lambda_z (cuDstProj_dx)
template <class T>
__global__ void eigen_scale_kernel(T* __restrict__ a, const T* __restrict__ lx,
                                   const T* __restrict__ ly, const T* __restrict__ lz,
                                   int n1, int n2, int n3) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int x = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (y >= n2 || x >= n1) return;
    a[y + (int64_t)x * n2 + (int64_t)z * n1 * n2] /= lx[x] + ly[y] + lz[z];
}

dim3 grid_xyz(int nx, int ny, int nz) { return dim3((ny + 63) / 64, (nx + 3) / 4, nz); }

hipError_t done(const char* name) {
    set_last_kernel(name);
    return hipGetLastError();
}

}  // namespace
}  // namespace whhd

// Handles hold their tables and scratch in FP32 or, for the *_f64 entry
// points (the reference built with USE_DOUBLE), in FP64.
struct whhd_dst3d_s {
    int n1, n2, n3;
    bool planar = false;  // 2D solver (CuDst): no z passes, no z eigenvalue
    bool fp64 = false;
    // per axis: log2(n + 1) when the FFT kernel applies (S holds its table),
    // else 0 (S holds the dense orthonormal sine matrix)
    int logm[3] = {0, 0, 0};
    void* S[3] = {nullptr, nullptr, nullptr};
    void* lam[3] = {nullptr, nullptr, nullptr};
    void* tmp = nullptr;
};

struct whhd_sweep3d_s {
    int nx, ny, nz;       // nz = 0: a 2D sweep
    char bc;              // 'n', or 'd' ('z' behaves as 'd')
    bool fp64 = false;
    double dx, dy, dz;
    whhd_dst3d_s* dst = nullptr;
    void* b = nullptr;
};

namespace whhd {
namespace {

void dst_release(whhd_dst3d_s* h) {
    if (!h) return;
    for (int a = 0; a < 3; ++a) {
        if (h->S[a]) (void)hipFree(h->S[a]);
        if (h->lam[a]) (void)hipFree(h->lam[a]);
    }
    if (h->tmp) (void)hipFree(h->tmp);
    delete h;
}

template <class T>
constexpr bool is64() { return sizeof(T) == sizeof(double); }

// Lines per FFT tile: a divisor of the workgroup size, with the staged rows
// within 20 KB of LDS (16 lines of 255 samples; 32-line tiles, whole 128-byte
// rows for the strided axes, measured 8% slower: fewer workgroups per CU).
int dst_fft_tile(int M, size_t elem) {
    int tl = 64;
    while (tl > 4 && (size_t)tl * (M + 1) * elem > 20 * 1024) tl >>= 1;
    return tl;
}

// one DST-I pass along axis a (0: x, 1: y, 2: z) of an (n1, n2, n3) y-fastest array
template <class T>
int dst_pass(const whhd_dst3d_s* h, T* out, const T* in, int a, hipStream_t s,
             bool scale = false, int reps = 1) {
    const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
    int n, m;
    int64_t si, sj, sb;
    unsigned nb;
    if (a == 0) {
        n = n1; m = n2; si = 1; sj = n2; sb = (int64_t)n1 * n2; nb = (unsigned)n3;
    } else if (a == 1) {
        n = n2; m = n1 * n3; si = n2; sj = 1; sb = 0; nb = 1;
    } else {
        n = n3; m = n1 * n2; si = 1; sj = (int64_t)n1 * n2; sb = 0; nb = 1;
    }
    if (h->logm[a]) {
        const int lm = h->logm[a], M = 1 << lm, P = M / 2;
        int64_t lines, n_inner, s_inner, s_outer, step;
        if (a == 0) {
            lines = (int64_t)n2 * n3; n_inner = n2; s_inner = 1; s_outer = (int64_t)n1 * n2; step = n2;
        } else if (a == 1) {
            lines = (int64_t)n1 * n3; n_inner = lines; s_inner = n2; s_outer = 0; step = 1;
        } else {
            lines = (int64_t)n1 * n2; n_inner = lines; s_inner = 1; s_outer = 0; step = (int64_t)n1 * n2;
        }
        const int tl = dst_fft_tile(M, sizeof(T));
        const size_t lds = (size_t)tl * (M + 1) * sizeof(T) + 8 * (size_t)P * 2 * sizeof(T) +
                           (size_t)tl * sizeof(int64_t) + (size_t)tl * sizeof(T) +
                           (size_t)(M + 3 * P) * sizeof(T);
        const dim3 grid((unsigned)((lines + tl - 1) / tl));
        DstScale sc{nullptr, nullptr, nullptr, 1};
        if (scale) {
            // lines of the z pass are li = y + x n2; of the planar solver's y pass li = x
            if (a == 2) sc = DstScale{h->lam[0], h->lam[1], h->lam[2], n2};
            else sc = DstScale{h->lam[2], h->lam[0], h->lam[1], n1};
        }
        const T c = (T)std::sqrt(2.0 / M);
        const T* tab = (const T*)h->S[a];
#define WHHD_DST_FFT(L) \
    case L: hipLaunchKernelGGL((dst_fft_kernel<T, L>), grid, dim3(kFftThreads), lds, s, out, in, tab, n, \
                               lines, n_inner, s_inner, s_outer, step, tl, c, sc, reps); break;
        switch (lm) {
            WHHD_DST_FFT(2) WHHD_DST_FFT(3) WHHD_DST_FFT(4) WHHD_DST_FFT(5) WHHD_DST_FFT(6)
            WHHD_DST_FFT(7) WHHD_DST_FFT(8) WHHD_DST_FFT(9) WHHD_DST_FFT(10)
            default: return fail(WHHD_INVALID_ARG, "dst: FFT length out of range");
        }
#undef WHHD_DST_FFT
        hipError_t e = done("dst_fft_kernel");
        return e == hipSuccess ? WHHD_OK : hip_fail(e, "dst fft kernel");
    }
    dim3 grid((unsigned)((m + kTI - 1) / kTI), (unsigned)((n + kTK - 1) / kTK), nb);
    hipLaunchKernelGGL(dst_pass_kernel<T>, grid, dim3(256), 0, s, out, in, (const T*)h->S[a], n, m,
                       si, sj, sb);
    hipError_t e = done("dst_pass_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "dst pass kernel");
}

#define TRY(e) do { int s_ = (e); if (s_ != WHHD_OK) return s_; } while (0)

}  // namespace

template <class T>
int dst3d_create_t(int n1, int n2, int n3, double dx, double dy, double dz, bool planar,
                   whhd_dst3d_s** out) {
    if (!out) return fail(WHHD_INVALID_ARG, "dst3d_create: null handle pointer");
    *out = nullptr;
    if (n1 < 1 || n2 < 1 || n3 < 1 || n3 > 65535 || (int64_t)n1 * n2 * n3 >= (1LL << 31))
        return fail(WHHD_INVALID_ARG, "dst3d_create: extents out of range");
    if (!(dx > 0 && dy > 0 && dz > 0)) return fail(WHHD_INVALID_ARG, "dst3d_create: spacing <= 0");
    auto* h = new whhd_dst3d_s;
    h->n1 = n1; h->n2 = n2; h->n3 = n3;
    h->planar = planar;
    h->fp64 = is64<T>();
    const int ns[3] = {n1, n2, n3};
    const double hs[3] = {dx, dy, dz};
    for (int a = 0; a < 3; ++a) {
        const int n = ns[a];
        // orthonormal DST-I, S(k, j) = sqrt(2/(n+1)) sin(pi (k+1)(j+1) / (n+1)), and the
        // eigenvalues (2 - 2 cos(pi (k+1)/(n+1))) / h^2 of the 3-point -d2/dx2 (none
        // on z for the 2D solver)
        const int M = n + 1;
        const bool fft = M >= 4 && M <= 1024 && (M & (M - 1)) == 0;
        std::vector<T> S, lam(n);
        if (fft) {
            // sin(pi m / M), m < M; e^{-2 pi i t / P}, t < P/2; e^{-2 pi i k / M}, k < P
            int lm = 0;
            while ((1 << lm) < M) ++lm;
            h->logm[a] = lm;
            const int P = M / 2;
            for (int m = 0; m < M; ++m) S.push_back((T)std::sin(M_PI * m / M));
            for (int t = 0; t < P / 2; ++t) {
                S.push_back((T)std::cos(2.0 * M_PI * t / P));
                S.push_back((T)-std::sin(2.0 * M_PI * t / P));
            }
            for (int k = 0; k < P; ++k) {
                S.push_back((T)std::cos(2.0 * M_PI * k / M));
                S.push_back((T)-std::sin(2.0 * M_PI * k / M));
            }
        } else {
            S.resize((size_t)n * n);
            const double sc = std::sqrt(2.0 / (n + 1));
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j)
                    S[(size_t)k * n + j] = (T)(sc * std::sin(M_PI * (double)((int64_t)(k + 1) * (j + 1) % (2 * (n + 1))) / (n + 1)));
        }
        for (int k = 0; k < n; ++k)
            lam[k] = (planar && a == 2) ? T(0)
                                        : (T)((2.0 - 2.0 * std::cos(M_PI * (k + 1) / (n + 1))) / (hs[a] * hs[a]));
        if (hipMalloc(&h->S[a], S.size() * sizeof(T)) != hipSuccess ||
            hipMalloc(&h->lam[a], lam.size() * sizeof(T)) != hipSuccess ||
            hipMemcpy(h->S[a], S.data(), S.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->lam[a], lam.data(), lam.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
            dst_release(h);
            return fail(WHHD_OOM, "dst3d_create: table allocation failed");
        }
    }
    if (hipMalloc(&h->tmp, (size_t)n1 * n2 * n3 * sizeof(T)) != hipSuccess) {
        dst_release(h);
        return fail(WHHD_OOM, "dst3d_create: scratch allocation failed");
    }
    *out = h;
    return WHHD_OK;
}

template <class T>
int dst3d_solve_t(const whhd_dst3d_s* h, T* phi, const T* b, hipStream_t s) {
    if (!h || !phi || !b) return fail(WHHD_INVALID_ARG, "dst3d_solve: null argument");
    if (h->fp64 != is64<T>())
        return fail(WHHD_INVALID_ARG, is64<T>() ? "dst solve (FP64): the handle was created for FP32"
                                                : "dst solve (FP32): the handle was created for FP64");
    T* tmp = (T*)h->tmp;
    const T* const* lam = (const T* const*)h->lam;
    if (h->planar) {
        // b -> tmp (x) -> phi (y), scale, phi -> tmp (y) -> phi (x)
        if (h->logm[0] && h->logm[1]) {
            // FFT passes stage whole lines per tile, so they run in place: x, then
            // forward y, scaling and inverse y in one pass, then x
            TRY(dst_pass<T>(h, phi, b, 0, s));
            TRY(dst_pass<T>(h, phi, phi, 1, s, true, 2));
            return dst_pass<T>(h, phi, phi, 0, s);
        }
        TRY(dst_pass<T>(h, tmp, b, 0, s));
        TRY(dst_pass<T>(h, phi, tmp, 1, s));
        hipLaunchKernelGGL(eigen_scale_kernel<T>, grid_xyz(h->n1, h->n2, 1), dim3(64, 4), 0, s, phi,
                           lam[0], lam[1], lam[2], h->n1, h->n2, 1);
        hipError_t e = done("eigen_scale_kernel");
        if (e != hipSuccess) return hip_fail(e, "eigen scale kernel");
        TRY(dst_pass<T>(h, tmp, phi, 1, s));
        return dst_pass<T>(h, phi, tmp, 0, s);
    }
    // b -> tmp (x) -> phi (y) -> tmp (z), scale, tmp -> phi (z) -> tmp (y) -> phi (x)
    if (h->logm[0] && h->logm[1] && h->logm[2]) {
        // FFT passes stage whole lines per tile, so they run in place: x, y, then
        // forward z, scaling and inverse z in one pass, then y, x
        TRY(dst_pass<T>(h, phi, b, 0, s));
        TRY(dst_pass<T>(h, phi, phi, 1, s));
        TRY(dst_pass<T>(h, phi, phi, 2, s, true, 2));
        TRY(dst_pass<T>(h, phi, phi, 1, s));
        return dst_pass<T>(h, phi, phi, 0, s);
    }
    TRY(dst_pass<T>(h, tmp, b, 0, s));
    TRY(dst_pass<T>(h, phi, tmp, 1, s));
    TRY(dst_pass<T>(h, tmp, phi, 2, s));
    hipLaunchKernelGGL(eigen_scale_kernel<T>, grid_xyz(h->n1, h->n2, h->n3), dim3(64, 4), 0, s, tmp,
                       lam[0], lam[1], lam[2], h->n1, h->n2, h->n3);
    hipError_t e = done("eigen_scale_kernel");
    if (e != hipSuccess) return hip_fail(e, "eigen scale kernel");
    TRY(dst_pass<T>(h, phi, tmp, 2, s));
    TRY(dst_pass<T>(h, tmp, phi, 1, s));
    return dst_pass<T>(h, phi, tmp, 0, s);
}

namespace {

template <class T>
int sweep3d_create_t(int nx, int ny, int nz, double dx, double dy, double dz, char bc,
                     whhd_sweep3d_t* out) {
    if (!out) return fail(WHHD_INVALID_ARG, "sweep3d_create: null handle pointer");
    *out = nullptr;
    if (bc != 'n' && bc != 'd' && bc != 'z')
        return fail(WHHD_INVALID_ARG, "sweep3d_create: velocity bc must be 'n', 'd' or 'z'");
    if (nx < 2 || ny < 2 || nz < 2 || nz > 65534 ||
        (int64_t)(nx + 2) * (ny + 2) * (nz + 2) >= (1LL << 31))
        return fail(WHHD_INVALID_ARG, "sweep3d_create: extents out of range");
    if (!(dx > 0 && dy > 0 && dz > 0)) return fail(WHHD_INVALID_ARG, "sweep3d_create: spacing <= 0");
    auto* h = new whhd_sweep3d_s;
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->bc = bc == 'n' ? 'n' : 'd';
    h->fp64 = is64<T>();
    h->dx = dx; h->dy = dy; h->dz = dz;
    // the potential grid's interior nodes ('n'), or all of its nodes ('d')
    const int e = h->bc == 'n' ? -1 : 1;
    int st = dst3d_create_t<T>(nx + e, ny + e, nz + e, dx, dy, dz, false, &h->dst);
    if (st == WHHD_OK &&
        hipMalloc(&h->b, (size_t)(nx + e) * (ny + e) * (nz + e) * sizeof(T)) != hipSuccess)
        st = fail(WHHD_OOM, "sweep3d_create: scratch allocation failed");
    if (st != WHHD_OK) {
        dst_release(h->dst);
        delete h;
        return st;
    }
    *out = h;
    return WHHD_OK;
}

template <class T>
int check_sweep(const whhd_sweep3d_s* h, bool planar, const char* what) {
    if (planar != (h->nz == 0))
        return fail(WHHD_INVALID_ARG, std::string(what) + (planar ? ": the handle is a 3D sweep"
                                                                  : ": the handle is a 2D sweep"));
    if (h->fp64 != is64<T>())
        return fail(WHHD_INVALID_ARG, std::string(what) + (is64<T>() ? ": the handle was created for FP32"
                                                                     : ": the handle was created for FP64"));
    return WHHD_OK;
}

template <class T>
int sweep3d_solve_t(whhd_sweep3d_t h, T* qx, T* qy, T* qz, const T* xv, const T* yv, const T* zv,
                    hipStream_t s) {
    if (!h || !qx || !qy || !qz || !xv || !yv)
        return fail(WHHD_INVALID_ARG, "sweep3d_solve: null argument");
    TRY(check_sweep<T>(h, false, "sweep3d_solve"));
    const bool d = h->bc == 'd';
    if (d && !zv) return fail(WHHD_INVALID_ARG, "sweep3d_solve: velocity bc 'd' needs zv");
    if (d)  // SetQxQyBoundary 'd': qy's bottom plane integrates zv along x
        hipLaunchKernelGGL(bottom_qy_kernel<T>, dim3((h->ny + 255) / 256), dim3(256), 0, s, qy, zv,
                           h->nx, h->ny, (T)h->dx);
    hipLaunchKernelGGL(sweep_kernel<T>, grid_xyz(h->nx + 1, h->ny + 1, 1), dim3(64, 4), 0, s, qx, qy,
                       qz, xv, yv, h->nx, h->ny, h->nz, (T)h->dz, d ? 1 : 0);
    if (!d)
        hipLaunchKernelGGL(top_gauge_kernel<T>, dim3((h->ny + 1 + 63) / 64), dim3(64 * kGaugeWaves), 0, s,
                           qx, qy, qz, h->nx, h->ny, h->nz, (T)h->dx, (T)h->dz);
    hipError_t e = done("sweep_kernels");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "sweep kernels");
}

template <class T>
int sweep3d_project_t(whhd_sweep3d_t h, T* qx, T* qy, T* qz, hipStream_t s) {
    if (!h || !qx || !qy || !qz) return fail(WHHD_INVALID_ARG, "sweep3d_project: null argument");
    TRY(check_sweep<T>(h, false, "sweep3d_project"));
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const T ihx = (T)(1.0 / h->dx), ihy = (T)(1.0 / h->dy), ihz = (T)(1.0 / h->dz);
    T* b = (T*)h->b;
    if (h->bc == 'd') {  // ProjectD, WaveletPotentials/cuParallelSweep3D.cu:1066-1079
        hipLaunchKernelGGL(node_div_d_kernel<T>, grid_xyz(nx + 1, ny + 1, nz + 1), dim3(64, 4), 0, s,
                           b, qx, qy, qz, nx, ny, nz, ihx, ihy, ihz);
        hipError_t e = done("node_div_d_kernel");
        if (e != hipSuccess) return hip_fail(e, "node divergence kernel");
        TRY(dst3d_solve_t<T>(h->dst, b, b, s));
        hipLaunchKernelGGL(grad_minus_d_kernel<T>, grid_xyz(nx + 1, ny + 1, nz + 1), dim3(64, 4), 0,
                           s, qx, qy, qz, b, nx, ny, nz, ihx, ihy, ihz);
        e = done("grad_minus_d_kernel");
        return e == hipSuccess ? WHHD_OK : hip_fail(e, "gradient kernel");
    }
    hipLaunchKernelGGL(node_div_kernel<T>, grid_xyz(nx - 1, ny - 1, nz - 1), dim3(64, 4), 0, s, b,
                       qx, qy, qz, nx, ny, nz, ihx, ihy, ihz);
    hipError_t e = done("node_div_kernel");
    if (e != hipSuccess) return hip_fail(e, "node divergence kernel");
    TRY(dst3d_solve_t<T>(h->dst, b, b, s));
    hipLaunchKernelGGL(grad_minus_kernel<T>, grid_xyz(nx + 1, ny + 1, nz + 1), dim3(64, 4), 0, s, qx,
                       qy, qz, b, nx, ny, nz, ihx, ihy, ihz);
    e = done("grad_minus_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "gradient kernel");
}

template <class T>
int sweep2d_create_t(int nx, int ny, double dx, double dy, char bc, whhd_sweep3d_t* out) {
    if (!out) return fail(WHHD_INVALID_ARG, "sweep2d_create: null handle pointer");
    *out = nullptr;
    if (bc != 'n' && bc != 'd' && bc != 'z')
        return fail(WHHD_INVALID_ARG, "sweep2d_create: velocity bc must be 'n', 'd' or 'z'");
    if (nx < 1 || ny < 1 || (int64_t)(nx + 1) * (ny + 1) >= (1LL << 31))
        return fail(WHHD_INVALID_ARG, "sweep2d_create: extents out of range");
    if (!(dx > 0 && dy > 0)) return fail(WHHD_INVALID_ARG, "sweep2d_create: spacing <= 0");
    auto* h = new whhd_sweep3d_s;
    h->nx = nx; h->ny = ny; h->nz = 0;
    h->bc = bc == 'n' ? 'n' : 'd';
    h->fp64 = is64<T>();
    h->dx = dx; h->dy = dy; h->dz = 1.0;
    *out = h;
    return WHHD_OK;
}

template <class T>
int sweep2d_solve_t(whhd_sweep3d_t h, T* qz, const T* xv, const T* yv, hipStream_t s) {
    if (!h || !qz || !xv || !yv) return fail(WHHD_INVALID_ARG, "sweep2d_solve: null argument");
    TRY(check_sweep<T>(h, true, "sweep2d_solve"));
    hipLaunchKernelGGL(sweep2d_kernel<T>, dim3((h->ny + 1 + 255) / 256), dim3(256), 0, s, qz, xv, yv,
                       h->nx, h->ny, (T)h->dx, (T)h->dy, h->bc == 'd' ? 1 : 0);
    hipError_t e = done("sweep2d_kernel");
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "2D sweep kernel");
}

}  // namespace
}  // namespace whhd

extern "C" {

WHHD_API int whhd_dst3d_create(int n1, int n2, int n3, double dx, double dy, double dz,
                               whhd_dst3d_t* out) {
    return whhd::dst3d_create_t<float>(n1, n2, n3, dx, dy, dz, false, out);
}
WHHD_API int whhd_dst3d_create_f64(int n1, int n2, int n3, double dx, double dy, double dz,
                                   whhd_dst3d_t* out) {
    return whhd::dst3d_create_t<double>(n1, n2, n3, dx, dy, dz, false, out);
}
WHHD_API int whhd_dst2d_create(int n1, int n2, double dx, double dy, whhd_dst3d_t* out) {
    return whhd::dst3d_create_t<float>(n1, n2, 1, dx, dy, 1.0, true, out);
}
WHHD_API int whhd_dst2d_create_f64(int n1, int n2, double dx, double dy, whhd_dst3d_t* out) {
    return whhd::dst3d_create_t<double>(n1, n2, 1, dx, dy, 1.0, true, out);
}
WHHD_API int whhd_dst3d_solve(whhd_dst3d_t h, float* phi, const float* b, void* stream) {
    return whhd::dst3d_solve_t<float>(h, phi, b, (hipStream_t)stream);
}
WHHD_API int whhd_dst3d_solve_f64(whhd_dst3d_t h, double* phi, const double* b, void* stream) {
    return whhd::dst3d_solve_t<double>(h, phi, b, (hipStream_t)stream);
}
WHHD_API void whhd_dst3d_destroy(whhd_dst3d_t h) { whhd::dst_release(h); }

WHHD_API int whhd_sweep3d_create(int nx, int ny, int nz, double dx, double dy, double dz, char bc,
                                 whhd_sweep3d_t* out) {
    return whhd::sweep3d_create_t<float>(nx, ny, nz, dx, dy, dz, bc, out);
}
WHHD_API int whhd_sweep3d_create_f64(int nx, int ny, int nz, double dx, double dy, double dz,
                                     char bc, whhd_sweep3d_t* out) {
    return whhd::sweep3d_create_t<double>(nx, ny, nz, dx, dy, dz, bc, out);
}
WHHD_API int whhd_sweep3d_solve(whhd_sweep3d_t h, float* qx, float* qy, float* qz,
                                const float* xv, const float* yv, const float* zv, void* stream) {
    return whhd::sweep3d_solve_t<float>(h, qx, qy, qz, xv, yv, zv, (hipStream_t)stream);
}
WHHD_API int whhd_sweep3d_solve_f64(whhd_sweep3d_t h, double* qx, double* qy, double* qz,
                                    const double* xv, const double* yv, const double* zv,
                                    void* stream) {
    return whhd::sweep3d_solve_t<double>(h, qx, qy, qz, xv, yv, zv, (hipStream_t)stream);
}
WHHD_API int whhd_sweep3d_project(whhd_sweep3d_t h, float* qx, float* qy, float* qz,
                                  void* stream) {
    return whhd::sweep3d_project_t<float>(h, qx, qy, qz, (hipStream_t)stream);
}
WHHD_API int whhd_sweep3d_project_f64(whhd_sweep3d_t h, double* qx, double* qy, double* qz,
                                      void* stream) {
    return whhd::sweep3d_project_t<double>(h, qx, qy, qz, (hipStream_t)stream);
}
WHHD_API int whhd_sweep2d_create(int nx, int ny, double dx, double dy, char bc,
                                 whhd_sweep3d_t* out) {
    return whhd::sweep2d_create_t<float>(nx, ny, dx, dy, bc, out);
}
WHHD_API int whhd_sweep2d_create_f64(int nx, int ny, double dx, double dy, char bc,
                                     whhd_sweep3d_t* out) {
    return whhd::sweep2d_create_t<double>(nx, ny, dx, dy, bc, out);
}
WHHD_API int whhd_sweep2d_solve(whhd_sweep3d_t h, float* qz, const float* xv, const float* yv,
                                void* stream) {
    return whhd::sweep2d_solve_t<float>(h, qz, xv, yv, (hipStream_t)stream);
}
WHHD_API int whhd_sweep2d_solve_f64(whhd_sweep3d_t h, double* qz, const double* xv,
                                    const double* yv, void* stream) {
    return whhd::sweep2d_solve_t<double>(h, qz, xv, yv, (hipStream_t)stream);
}
WHHD_API void whhd_sweep3d_destroy(whhd_sweep3d_t h) {
    if (!h) return;
    whhd::dst_release(h->dst);
    if (h->b) (void)hipFree(h->b);
    delete h;
}

}  // extern "C"
