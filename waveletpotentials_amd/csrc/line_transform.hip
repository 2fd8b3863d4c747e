// Multi-level lifting wavelet transform of every line of a 3D array along one
// axis (the work of CuWHHDForward::Solve1D, WaveletPotentials/cuWHHDForward.cu:
// 1122-1253, kernels :5-742).
//
// Design for gfx950: a workgroup stages a tile of whole lines in LDS with
// coalesced global accesses, then each wavefront lifts one line at a time in
// windows of 64 pairs (64 - 2*halo pairs written per window), and the tile is
// written back coalesced.  Two tile shapes:
//   * contiguous lines (y axis): TILE consecutive lines form one contiguous
//     block of memory;
//   * strided lines (x and z axes): TILE consecutive y positions of lines that
//     share the outer index, so every sample row is TILE consecutive floats.
// LDS rows are padded to an odd length so that per-line walks and the
// sample-major staging loads are both free of bank conflicts.
#include <hip/hip_runtime.h>

#include <cstdio>

#include <cstdint>

#include "lifting.h"
#include "whhd_internal.h"

namespace whhd {

constexpr int kWave = 64;

// Compiler-level ordering point between LDS writes of one window and reads of
// the next: a wavefront's LDS operations execute in order, so this only has
// to stop the compiler from moving them across each other.
WHHD_DEV void lds_order() { asm volatile("" ::: "memory"); }

// Forward transform of one line staged in `s` (length n); the result is left
// in `r`.  Follows cuWHHDForward_y_ml: levels l = levels-1..2 lift pairs of
// the first 2^(l+1)+ext samples, approximations stay at the front of `s`,
// details of level l go to r[2^l+ext, 2^(l+1)+ext), the last 4+ext
// approximations get the dense L10 matrix.
template <class W, class T>
WHHD_DEV void forward_line(T* s, T* r, int n, int levels, int lane) {
    constexpr int H = W::halo;
    constexpr int WL = kWave - 2 * H;
    const int ext = n & 1;
    T a = T(0), d = T(0);
    if (levels > 2) {
        for (int l = levels - 1; l > 1; --l) {
            const int L = 1 << (l + 1);
            const int Wd = (L >> 1) + ext;
            for (int base = 0; base < Wd; base += WL) {
                const int k = base + lane - H;
                int ai = 2 * k, di = 2 * k + 1;
                T sa, sd;
                W::ix::fwd(ai, di, sa, sd, L);
                di = di < 0 ? 0 : di;
                if (k < Wd + H) {
                    // clamp keeps reads inside the row; valid inputs never need it
                    ai = min(max(ai, 0), n - 1);
                    di = min(di, n - 1);
                    a = s[ai];
                    d = s[di];
                }
                W::up::fwd(a, d, sa, sd);
                lds_order();
                if (lane >= H && lane < H + WL && k < Wd) {
                    s[k] = a;
                    if (k < Wd - ext) r[Wd + k] = d;
                }
                lds_order();
            }
        }
        a = lane_dn(a, H);
    } else {
        if (lane < W::m10) a = s[lane];
    }
    const T v = apply_l10<W::m10, W::l10>(a, lane);
    if (lane < W::m10) r[lane] = v;
    lds_order();
}

// Inverse transform of one line staged in `in`; `other` is the second buffer.
// Follows cuWHHDBackward_y_ml: inverse L10 first, then levels l = 2..levels-1
// reading approximations [0, 2^l+ext) from the previous level and details
// [2^l+ext, 2^(l+1)+ext) from the staged input; buffers alternate so that the
// input buffer is reused once its details have been consumed.  Returns the
// buffer holding the result.
template <class W, class T>
WHHD_DEV T* inverse_line(T* in, T* other, int n, int levels, int lane) {
    constexpr int H = W::halo;
    constexpr int WL = kWave - 2 * H;
    const int ext = n & 1;
    T* dst = other;
    T* src = in;
    if (levels & 1) {
        dst = in;
        src = other;
    }
    T a = T(0), d = T(0);
    if (lane < W::m10) a = in[lane];
    const T v = apply_l10<W::m10, W::il10>(a, lane);
    lds_order();
    if (lane < W::m10) dst[lane] = v;
    lds_order();
    for (int l = 2; l < levels; ++l) {
        const int L = 1 << l;
        const int Wd = L + ext;
        T* t = dst;
        dst = src;
        src = t;
        for (int base = 0; base < Wd; base += WL) {
            const int k = base + lane - H;
            int ai = k, di = k;
            T sa, sd;
            W::ix::inv(ai, di, sa, sd, L);
            if (k < L + H) {
                ai = min(max(ai, 0), n - 1);
                const int dd = min(max(di + L + ext, 0), n - 1);
                d = in[dd];
                a = src[ai];
            }
            W::up::inv(a, d, sa, sd);
            lds_order();
            if (lane >= H && lane < H + WL) {
                if (k < L + ext) dst[2 * k] = a;
                if (k < L) dst[2 * k + 1] = d;
            }
            lds_order();
        }
    }
    return dst;
}

// Tile kernel.  STRIDED == false: lines are contiguous (y axis), tile = `tile`
// consecutive lines.  STRIDED == true: tile = `tile` consecutive y positions
// at outer index blockIdx.y; sample i of line (y, o) sits at
// y + o*outer_stride + i*sample_stride.
template <class W, class T, bool FWD, bool STRIDED>
__global__ void __launch_bounds__(256) line_tile_kernel(T* __restrict__ dst,
                                                        const T* __restrict__ src, LineGeom g) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    const int n = g.n;
    const int ls = g.lds_row;  // padded row length
    const int tile = g.tile;
    T* buf0 = sm;
    T* buf1 = sm + (size_t)tile * ls;
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int nwaves = nthreads / kWave;

    int64_t tile_base;  // element offset of line 0 of the tile
    int lines_here;
    if (!STRIDED) {
        const int64_t first = (int64_t)blockIdx.x * tile;
        lines_here = (int)min<int64_t>((int64_t)tile, g.lines - first);
        tile_base = first * n;
    } else {
        const int y0 = blockIdx.x * tile;
        lines_here = min(tile, g.lines - y0);
        tile_base = (int64_t)y0 + (int64_t)blockIdx.y * g.outer_stride;
    }

    // ---- stage ----
    if (!STRIDED) {
        const int total = lines_here * n;
        for (int e = tid; e < total; e += nthreads) {
            const int li = e / n;
            const int i = e - li * n;
            buf0[li * ls + i] = src[tile_base + e];
        }
    } else {
        const int total = n * tile;
        for (int e = tid; e < total; e += nthreads) {
            const int i = e / tile;
            const int yl = e - i * tile;
            if (yl < lines_here)
                buf0[yl * ls + i] = src[tile_base + yl + (int64_t)i * g.sample_stride];
        }
    }
    __syncthreads();

    // ---- lift, one line per wavefront at a time ----
    // Both directions leave every line's result in buf1: the forward writes its
    // outputs there, and the inverse starts its ping-pong so that the last
    // level lands in the second buffer whatever the level count.
    for (int li = wave; li < lines_here; li += nwaves) {
        T* s = buf0 + li * ls;
        T* r = buf1 + li * ls;
        if (FWD) forward_line<W, T>(s, r, n, g.levels, lane);
        else inverse_line<W, T>(s, r, n, g.levels, lane);
    }
    __syncthreads();

    // ---- write back ----
    const T* res = buf1;
    if (!STRIDED) {
        const int total = lines_here * n;
        for (int e = tid; e < total; e += nthreads) {
            const int li = e / n;
            const int i = e - li * n;
            dst[tile_base + e] = res[li * ls + i];
        }
    } else {
        const int total = n * tile;
        for (int e = tid; e < total; e += nthreads) {
            const int i = e / tile;
            const int yl = e - i * tile;
            if (yl < lines_here) dst[tile_base + yl + (int64_t)i * g.sample_stride] = res[yl * ls + i];
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side dispatch
// ---------------------------------------------------------------------------
template <class W, class T>
static hipError_t launch_w(T* dst, const T* src, const LineGeom& g, bool fwd, bool strided,
                           int threads, hipStream_t stream) {
    const size_t smem = (size_t)2 * g.tile * g.lds_row * sizeof(T);
    dim3 block(threads);
    if (!strided) {
        dim3 grid((unsigned)((g.lines + g.tile - 1) / g.tile));
        if (fwd) hipLaunchKernelGGL((line_tile_kernel<W, T, true, false>), grid, block, smem, stream, dst, src, g);
        else hipLaunchKernelGGL((line_tile_kernel<W, T, false, false>), grid, block, smem, stream, dst, src, g);
    } else {
        dim3 grid((unsigned)((g.lines + g.tile - 1) / g.tile), (unsigned)g.outer);
        if (fwd) hipLaunchKernelGGL((line_tile_kernel<W, T, true, true>), grid, block, smem, stream, dst, src, g);
        else hipLaunchKernelGGL((line_tile_kernel<W, T, false, true>), grid, block, smem, stream, dst, src, g);
    }
    return hipGetLastError();
}

template <class T>
static hipError_t dispatch(int wt, char bc, T* dst, const T* src, const LineGeom& g, bool fwd,
                           bool strided, int threads, hipStream_t st) {
#define L_(WTYPE) return launch_w<WTYPE, T>(dst, src, g, fwd, strided, threads, st)
    switch (wt) {
    case WT_CDF3_5: if (bc == 'n') L_(W_Cdf35N); else L_(W_Cdf35D);
    case WT_CDF3_7: if (bc == 'n') L_(W_Cdf37N); else L_(W_Cdf37D);
    case WT_CDF2_6: if (bc == 'n') L_(W_Cdf26N); else if (bc == 'd') L_(W_Cdf26D); else L_(W_Cdf26Z);
    case WT_CDF2_8: if (bc == 'n') L_(W_Cdf28N); else if (bc == 'd') L_(W_Cdf28D); else L_(W_Cdf28Z);
    case WT_CDF4_4: if (bc == 'd') L_(W_Cdf44D); else L_(W_Cdf44N);
    case WT_CDF4_6: if (bc == 'd') L_(W_Cdf46D); else L_(W_Cdf46N);
    case WT_CDF2_0: if (bc == 'n') L_(W_Cdf20N); else L_(W_Cdf20D);
    case WT_CDF1_1: if (bc == 'n') L_(W_Cdf11N); else L_(W_Cdf11D);
    case WT_CDF1_7: if (bc == 'n') L_(W_Cdf17N); else L_(W_Cdf17D);
    default: return hipErrorInvalidValue;
    }
#undef L_
}

template <class T>
int transform_lines(T* dst, const T* src, int nx, int ny, int nz, int levels, char axis, char bc,
                    int wt, bool fwd, hipStream_t stream, const TileConfig* cfg,
                    const LineBatch* batch) {
    const LineBatch one{};
    const LineBatch& b = batch ? *batch : one;
    if (b.nb < 1 || (b.nb > 1 && b.stride < (int64_t)nx * ny * nz)) return WHHD_INVALID_ARG;
    int st = check_wavelet(wt, bc);
    if (st != WHHD_OK) return st;
    if (nx <= 0 || ny <= 0 || nz <= 0 || !dst || !src) return WHHD_INVALID_ARG;
    int n;
    LineGeom g{};
    bool strided;
    const int64_t slice = (int64_t)nx * ny;
    switch (axis) {
    case 'y': case 'Y':
        n = ny; strided = false; g.lines = (int64_t)nx * nz; break;
    case 'x': case 'X':
        n = nx; strided = true; g.lines = ny; g.outer = nz; g.sample_stride = ny; g.outer_stride = slice; break;
    case 'z': case 'Z':
        n = nz; strided = true; g.lines = ny; g.outer = nx; g.sample_stride = slice; g.outer_stride = ny; break;
    default:
        return WHHD_INVALID_ARG;
    }
    st = check_line(n, levels, wt);
    if (st != WHHD_OK) return st;
    if constexpr (sizeof(T) == sizeof(float)) {
        if ((!cfg || cfg->allow_line_parallel) && g_line_kernel_mode.load() == 0) {
            const int r = lp_transform((float*)dst, (const float*)src, nx, ny, nz, levels, axis, bc,
                                       wt, fwd, stream, b);
            if (r != -1) return r;
        }
    }
    g.n = n;
    g.levels = levels;
    g.lds_row = (n & 1) ? n : n + 1;
    TileConfig c = cfg ? *cfg : default_tile(n, sizeof(T), strided);
    g.tile = strided ? c.strided_tile : c.contig_tile;
    if (g.tile <= 0) return WHHD_INVALID_ARG;
    if ((size_t)2 * g.tile * g.lds_row * sizeof(T) > kMaxLdsBytes) return WHHD_INVALID_ARG;
    hipError_t e = hipSuccess;
    for (int p = 0; p < b.nb && e == hipSuccess; ++p)  // the tile kernels take one problem a launch
        e = dispatch<T>(wt, bc, dst + p * b.stride, src + p * b.stride, g, fwd, strided, c.threads,
                        stream);
    char name[96];
    std::snprintf(name, sizeof(name), "line_tile_kernel<%s %c,L=%d,%s,%s,%s>", wavelet_name(wt), bc,
                  levels, fwd ? "fwd" : "inv", strided ? "strided" : "contiguous",
                  sizeof(T) == sizeof(float) ? "f32" : "f64");
    set_last_kernel(name);
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "line tile kernel");
}

template int transform_lines<float>(float*, const float*, int, int, int, int, char, char, int, bool,
                                    hipStream_t, const TileConfig*, const LineBatch*);
template int transform_lines<double>(double*, const double*, int, int, int, int, char, char, int,
                                     bool, hipStream_t, const TileConfig*, const LineBatch*);

}  // namespace whhd
