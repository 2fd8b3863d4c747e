// Learned initial guess for the fractional multigrid: a 3-level 3D U-Net
// (16/32/64 channels, 3x3x3 convolutions, ReLU, 2x2x2 average pooling,
// nearest upsampling, concatenated skips) mapping (f, solid level set) to the
// pressure.  No reference counterpart: the reference's multigrid takes its
// initial guess as pre_v (WaveletPotentials/cuMultigridFrac3D.cu:249-258),
// which is where the prediction enters; the plain multigrid remains the solver
// and the judge of convergence.
//
// Everything runs on the device with this file's kernels:
//  * convolutions as implicit GEMMs on the BF16 matrix cores
//    (v_mfma_f32_16x16x32_bf16, FP32 accumulation): rows are voxels, columns
//    output channels, the reduction runs over (tap, input channel); the
//    upsampling and the skip concatenation are folded into the operand reads;
//  * the input gradient is the same kernel with the flipped, transposed
//    weights; the weight gradient reduces over voxels with per-wavefront LDS
//    transposes of 64-voxel chunks and FP32 atomics;
//  * pooling, ReLU masks, the loss and Adam are element-wise kernels.
// Activations and their gradients are BF16, channels-last ((z, x, y) voxel
// order, channels innermost); master weights, moments and gradients are FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "whhd_internal.h"

namespace whhd {
namespace unet {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { MODE_RELU = 0, MODE_F32 = 1, MODE_GRAD = 2 };

constexpr int kTaps = 27;
constexpr int kIn = 16;  // input channels (f, level set, 14 zero channels)

__host__ __device__ constexpr int kpad(int cin) { return (kTaps * cin + 31) / 32 * 32; }

// Source of a convolution: channels [0, CA) from `a`, at half resolution when
// the layer upsamples it; channels [CA, CA + CB) from `b`; output extents Z, X, Y.
struct ConvSrc {
    const bf16* a;
    const bf16* b;
    int Z, X, Y;
};

struct ConvOut {
    bf16* o1;            // RELU: all CO channels; GRAD: channels [0, split)
    bf16* o2;            // GRAD: channels [split, CO)
    float* of;           // F32: channel 0
    const float* bias;   // RELU, F32
    int split;
};

__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// y = conv3x3x3(x) on a wavefront tile of 32 voxels x CO channels.
template <int CA, int CB, bool UPA, int CO, int MODE>
__global__ void __launch_bounds__(256) conv3_kernel(ConvSrc s, const bf16* __restrict__ w,
                                                    ConvOut o) {
    constexpr int CIN = CA + CB;
    static_assert(CIN % 8 == 0 && CA % 8 == 0 && CO % 16 == 0, "channel blocking");
    constexpr int KP = kpad(CIN);
    constexpr int NT = CO / 16;
    constexpr int MT = 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t V = (int64_t)s.Z * s.X * s.Y;
    const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * (16 * MT);
    if (base >= V) return;
    const int m = lane & 15, g = lane >> 4;
    int vz[MT], vx[MT], vy[MT];
    bool vok[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int64_t v = base + t * 16 + m;
        vok[t] = v < V;
        const int64_t vv = vok[t] ? v : 0;
        vy[t] = (int)(vv % s.Y);
        const int64_t r = vv / s.Y;
        vx[t] = (int)(r % s.X);
        vz[t] = (int)(r / s.X);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[t][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int Xh = s.X >> 1, Yh = s.Y >> 1;
    // unrolled so the operand loads of the next chunks are in flight while
    // the matrix cores work on the current one (a serial chain of load, wait,
    // MFMA per chunk otherwise: small grids, with few wavefronts, were bound
    // by that latency)
#pragma unroll 4
    for (int kc = 0; kc < KP / 32; ++kc) {
        const int k0 = kc * 32 + 8 * g;
        const int tap = k0 / CIN, ci = k0 - tap * CIN;
        const int dz = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dy = tap % 3 - 1;
        bf16x8 A[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            const int zz = vz[t] + dz, xx = vx[t] + dx, yy = vy[t] + dy;
            const bool in = vok[t] && tap < kTaps && zz >= 0 && zz < s.Z && xx >= 0 &&
                            xx < s.X && yy >= 0 && yy < s.Y;
            A[t] = bf16x8{};
            if (in) {
                if (CB == 0 || ci < CA) {
                    const int64_t idx =
                        UPA ? ((((int64_t)(zz >> 1) * Xh + (xx >> 1)) * Yh + (yy >> 1)) * CA + ci)
                            : ((((int64_t)zz * s.X + xx) * s.Y + yy) * CA + ci);
                    A[t] = *(const bf16x8*)(s.a + idx);
                } else {
                    const int64_t idx = (((int64_t)zz * s.X + xx) * s.Y + yy) * CB + (ci - CA);
                    A[t] = *(const bf16x8*)(s.b + idx);
                }
            }
        }
        bf16x8 B[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) B[n] = *(const bf16x8*)(w + (int64_t)(n * 16 + m) * KP + k0);
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int n = 0; n < NT; ++n)
                acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[t], B[n], acc[t][n], 0, 0, 0);
    }
    // D row 4g + j is voxel base + 16t + 4g + j, column m is channel 16n + m
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t v = base + t * 16 + 4 * g + j;
                const int co = n * 16 + m;
                if (v >= V) continue;
                const float val = acc[t][n][j];
                if constexpr (MODE == MODE_RELU) {
                    o.o1[v * CO + co] = (bf16)fmaxf(val + o.bias[co], 0.f);
                } else if constexpr (MODE == MODE_F32) {
                    if (co == 0) o.of[v] = val + o.bias[0];
                } else {
                    if (co < o.split) o.o1[v * o.split + co] = (bf16)val;
                    else o.o2[v * (CO - o.split) + (co - o.split)] = (bf16)val;
                }
            }
}

// dW[co][tap][ci] += sum_v G[v][co] X[v + tap][ci] for tap = blockIdx.y; each
// wavefront walks 64-voxel chunks, staging G and the shifted X transposed in
// its own LDS rows (voxels contiguous) for the matrix-core reads.
template <int CA, int CB, bool UPA, int CO>
__global__ void __launch_bounds__(256) conv3_wgrad_kernel(ConvSrc s, const bf16* __restrict__ gy,
                                                          float* __restrict__ dw) {
    constexpr int CIN = CA + CB;
    constexpr int MO = CO / 16, NI = CIN / 16;
    static_assert(CIN % 16 == 0 && CO % 16 == 0, "channel blocking");
    constexpr int R = 72;  // LDS row: 64 voxels + 8 (keeps 16-byte rows aligned)
    extern __shared__ bf16 wg_lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bf16* Gt = wg_lds + wave * (CO + CIN) * R;
    bf16* Xt = Gt + CO * R;
    const int tap = blockIdx.y;
    const int dz = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dy = tap % 3 - 1;
    const int64_t V = (int64_t)s.Z * s.X * s.Y;
    const int64_t chunks = (V + 63) / 64;
    const int m = lane & 15, g = lane >> 4;
    const int Xh = s.X >> 1, Yh = s.Y >> 1;
    f32x4 acc[MO][NI];
#pragma unroll
    for (int a = 0; a < MO; ++a)
#pragma unroll
        for (int b = 0; b < NI; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t c = (int64_t)blockIdx.x * 4 + wave; c < chunks; c += (int64_t)gridDim.x * 4) {
        // lane = voxel of the chunk
        const int64_t v = c * 64 + lane;
        const bool ok = v < V;
        const int64_t vv = ok ? v : 0;
        const int y = (int)(vv % s.Y);
        const int64_t r = vv / s.Y;
        const int x = (int)(r % s.X), z = (int)(r / s.X);
#pragma unroll
        for (int q = 0; q < CO / 8; ++q) {
            bf16x8 val = bf16x8{};
            if (ok) val = *(const bf16x8*)(gy + v * CO + 8 * q);
#pragma unroll
            for (int e = 0; e < 8; ++e) Gt[(8 * q + e) * R + lane] = val[e];
        }
        const int zz = z + dz, xx = x + dx, yy = y + dy;
        const bool in = ok && zz >= 0 && zz < s.Z && xx >= 0 && xx < s.X && yy >= 0 && yy < s.Y;
#pragma unroll
        for (int q = 0; q < CIN / 8; ++q) {
            bf16x8 val = bf16x8{};
            if (in) {
                const int ci = 8 * q;
                if (CB == 0 || ci < CA) {
                    const int64_t idx =
                        UPA ? ((((int64_t)(zz >> 1) * Xh + (xx >> 1)) * Yh + (yy >> 1)) * CA + ci)
                            : ((((int64_t)zz * s.X + xx) * s.Y + yy) * CA + ci);
                    val = *(const bf16x8*)(s.a + idx);
                } else {
                    val = *(const bf16x8*)(s.b + (((int64_t)zz * s.X + xx) * s.Y + yy) * CB +
                                           (ci - CA));
                }
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) Xt[(8 * q + e) * R + lane] = val[e];
        }
        lds_order();
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bf16x8 A[MO];
#pragma unroll
            for (int a = 0; a < MO; ++a) A[a] = *(const bf16x8*)(Gt + (a * 16 + m) * R + ks * 32 + 8 * g);
#pragma unroll
            for (int b = 0; b < NI; ++b) {
                const bf16x8 B = *(const bf16x8*)(Xt + (b * 16 + m) * R + ks * 32 + 8 * g);
#pragma unroll
                for (int a = 0; a < MO; ++a)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], B, acc[a][b], 0, 0, 0);
            }
        }
        lds_order();
    }
    // D row 4g + j is output channel 16a + 4g + j, column m input channel 16b + m
#pragma unroll
    for (int a = 0; a < MO; ++a)
#pragma unroll
        for (int b = 0; b < NI; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = a * 16 + 4 * g + j, ci = b * 16 + m;
                atomicAdd(dw + ((int64_t)co * kTaps + tap) * CIN + ci, acc[a][b][j]);
            }
}

// db[c] += sum_v G[v][c]
__global__ void bias_grad_kernel(const bf16* __restrict__ gy, float* __restrict__ db, int C,
                                 int64_t V) {
    extern __shared__ float bg_lds[];
    const int groups = blockDim.x / C;  // blockDim.x is a multiple of C
    const int c = threadIdx.x % C, grp = threadIdx.x / C;
    float sum = 0.f;
    for (int64_t v = (int64_t)blockIdx.x * groups + grp; v < V; v += (int64_t)gridDim.x * groups)
        sum += (float)gy[v * C + c];
    bg_lds[threadIdx.x] = sum;
    __syncthreads();
    if (grp == 0) {
        for (int k = 1; k < groups; ++k) sum += bg_lds[k * C + c];
        atomicAdd(db + c, sum);
    }
}

// coarse[zc, xc, yc][c] = mean (or sum) of the 2x2x2 children
template <bool MEAN>
__global__ void pool_kernel(bf16* __restrict__ out, const bf16* __restrict__ in, int Zc, int Xc,
                            int Yc, int C) {
    const int64_t n = (int64_t)Zc * Xc * Yc * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    int64_t r = i / C;
    const int yc = (int)(r % Yc);
    r /= Yc;
    const int xc = (int)(r % Xc), zc = (int)(r / Xc);
    const int X = 2 * Xc, Y = 2 * Yc;
    float s = 0.f;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d)
                s += (float)in[((((int64_t)(2 * zc + a) * X + 2 * xc + b) * Y) + 2 * yc + d) * C + c];
    out[i] = (bf16)(MEAN ? s * 0.125f : s);
}

// fine[v][c] += coarse[parent(v)][c] / 8 (backward of the average pooling)
__global__ void pool_back_kernel(bf16* __restrict__ fine, const bf16* __restrict__ coarse, int Z,
                                 int X, int Y, int C) {
    const int64_t n = (int64_t)Z * X * Y * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    int64_t r = i / C;
    const int y = (int)(r % Y);
    r /= Y;
    const int x = (int)(r % X), z = (int)(r / X);
    const int64_t p = ((((int64_t)(z >> 1) * (X >> 1) + (x >> 1)) * (Y >> 1)) + (y >> 1)) * C + c;
    fine[i] = (bf16)((float)fine[i] + 0.125f * (float)coarse[p]);
}

// g <- g where y > 0, else 0 (ReLU backward)
__global__ void relu_mask_kernel(bf16* __restrict__ g, const bf16* __restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !((float)y[i] > 0.f)) g[i] = (bf16)0.f;
}

// FP64 partial sums.  DIFF = false: part[2b + 1] = sum a^2.  DIFF = true:
// e = a - b / (rms L^2) with rms from stats (sum f^2 over n), part[2b] = sum e,
// part[2b + 1] = sum e^2.
template <bool DIFF>
__global__ void sums_kernel(double* __restrict__ part, const float* __restrict__ a,
                            const float* __restrict__ b, const double* __restrict__ stats,
                            float len2, int64_t n) {
    double sc = 0.0;
    if (DIFF) {
        const double ms = stats[1] / (double)n;
        sc = ms > 0 ? 1.0 / (sqrt(ms) * (double)len2) : 0.0;
    }
    double s1 = 0.0, s2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (DIFF) {
            const double e = (double)a[i] - sc * (double)b[i];
            s1 += e;
            s2 += e * e;
        } else {
            const double e = a[i];
            s2 += e * e;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    __shared__ double l1[4], l2[4];
    if ((threadIdx.x & 63) == 0) {
        l1[threadIdx.x >> 6] = s1;
        l2[threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = l1[0] + l1[1] + l1[2] + l1[3];
        part[2 * blockIdx.x + 1] = l2[0] + l2[1] + l2[2] + l2[3];
    }
}

__global__ void sums_final_kernel(double* __restrict__ out, const double* __restrict__ part,
                                  int nparts) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        s1 += part[2 * i];
        s2 += part[2 * i + 1];
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    __shared__ double l1[4], l2[4];
    if ((threadIdx.x & 63) == 0) {
        l1[threadIdx.x >> 6] = s1;
        l2[threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = l1[0] + l1[1] + l1[2] + l1[3];
        out[1] = l2[0] + l2[1] + l2[2] + l2[3];
    }
}

// Network input: channel 0 = f / rms(f), channel 1 = level set / L, others 0.
// stats[1] holds sum f^2 (from sums_kernel<false>).
__global__ void input_kernel(bf16* __restrict__ x, const float* __restrict__ f,
                             const float* __restrict__ ls, const double* __restrict__ stats,
                             float inv_len, int64_t V) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double ms = stats[1] / (double)V;
    const float inv_rms = ms > 0 ? (float)(1.0 / sqrt(ms)) : 0.f;
    bf16x8 a = bf16x8{}, b = bf16x8{};
    a[0] = (bf16)(f[v] * inv_rms);
    a[1] = (bf16)(ls[v] * inv_len);
    *(bf16x8*)(x + v * kIn) = a;
    *(bf16x8*)(x + v * kIn + 8) = b;
}

// p0 = network output * rms(f) * L^2
__global__ void output_kernel(float* __restrict__ p, const float* __restrict__ out,
                              const double* __restrict__ stats, float len2, int64_t V) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double ms = stats[1] / (double)V;
    p[v] = (float)((double)out[v] * sqrt(ms) * (double)len2);
}

// Loss gradient: target t = p / (rms(f) L^2); e = out - t; loss = mean
// (e - mean e)^2 (the pressure is defined up to a constant); g = 2 (e - e_bar) / V
// into channel 0 of a 16-channel BF16 gradient.  lsum[0] = sum e, lsum[1] = sum e^2.
__global__ void loss_grad_kernel(bf16* __restrict__ g, const float* __restrict__ out,
                                 const float* __restrict__ p, const double* __restrict__ stats,
                                 const double* __restrict__ lsum, float len2, float gscale,
                                 int64_t V) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double ms = stats[1] / (double)V;
    const double sc = ms > 0 ? 1.0 / (sqrt(ms) * (double)len2) : 0.0;
    const double ebar = lsum[0] / (double)V;
    const double e = (double)out[v] - sc * (double)p[v];
    bf16x8 a = bf16x8{}, b = bf16x8{};
    a[0] = (bf16)((float)(2.0 * (e - ebar) / (double)V) * gscale);
    *(bf16x8*)(g + v * 16) = a;
    *(bf16x8*)(g + v * 16 + 8) = b;
}

// Adam (bias-corrected) over the flat parameter vector; grads are divided by
// gscale first (the loss gradient is scaled up to keep BF16 gradients normal).
__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2,
                            const float* __restrict__ grad, const uint8_t* __restrict__ frozen,
                            int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                            float inv_gscale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || frozen[i]) return;
    const float gi = grad[i] * inv_gscale;
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Forward layout wk[co][tap * CIN + ci] and input-gradient layout
// wt[ci][(26 - tap) * CO + co], both BF16 and zero-padded to 32-wide K.
__global__ void pack_kernel(bf16* __restrict__ wk, bf16* __restrict__ wt,
                            const float* __restrict__ w, int CO, int CIN) {
    const int64_t n = (int64_t)CO * kTaps * CIN;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ci = (int)(i % CIN);
    const int tap = (int)((i / CIN) % kTaps);
    const int co = (int)(i / ((int64_t)CIN * kTaps));
    const bf16 v = (bf16)w[i];
    wk[(int64_t)co * kpad(CIN) + tap * CIN + ci] = v;
    wt[(int64_t)ci * kpad(CO) + (kTaps - 1 - tap) * CO + co] = v;
}

// Network grid from a finer one: the mean of each F^3 block of fine cells.
__global__ void down_kernel(float* __restrict__ out, const float* __restrict__ in, int Zc, int Xc,
                            int Yc, int F) {
    const int64_t n = (int64_t)Zc * Xc * Yc;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int yc = (int)(i % Yc);
    const int64_t r = i / Yc;
    const int xc = (int)(r % Xc), zc = (int)(r / Xc);
    const int64_t X = (int64_t)Xc * F, Y = (int64_t)Yc * F;
    float s = 0.f;
    for (int a = 0; a < F; ++a)
#pragma unroll 4
        for (int b = 0; b < F; ++b) {
            const float* row = in + (((int64_t)zc * F + a) * X + (int64_t)xc * F + b) * Y + (int64_t)yc * F;
#pragma unroll 4
            for (int d = 0; d < F; ++d) s += row[d];
        }
    out[i] = s / (float)(F * F * F);
}

// Fine cell values by trilinear interpolation of the network output at cell
// centres (coarse coordinate (i + 1/2) / F - 1/2, clamped at the walls), times
// rms(f) * L^2 * scale.
// One thread per fine cell on a (64, 4) block over (y, x) and one z plane per
// grid row (no 64-bit index division per cell).
__global__ void up_kernel(float* __restrict__ p, const float* __restrict__ out,
                          const double* __restrict__ stats, float len2, int Zc, int Xc, int Yc,
                          int F) {
    const int X = Xc * F, Y = Yc * F;
    const int y = blockIdx.x * 64 + threadIdx.x, x = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (y >= Y || x >= X) return;
    const int64_t i = y + (int64_t)Y * (x + (int64_t)X * z);
    auto axis = [&](int c, int nc, int& c0, int& c1, float& w) {
        float u = ((float)c + 0.5f) / (float)F - 0.5f;
        u = fminf(fmaxf(u, 0.f), (float)(nc - 1));
        c0 = (int)u;
        c1 = c0 + 1 < nc ? c0 + 1 : c0;
        w = u - (float)c0;
    };
    int z0, z1, x0, x1, y0, y1;
    float wz, wx, wy;
    axis(z, Zc, z0, z1, wz);
    axis(x, Xc, x0, x1, wx);
    axis(y, Yc, y0, y1, wy);
    auto at = [&](int a, int b, int c) { return out[((int64_t)a * Xc + b) * Yc + c]; };
    const float v0 = (1.f - wx) * ((1.f - wy) * at(z0, x0, y0) + wy * at(z0, x0, y1)) +
                     wx * ((1.f - wy) * at(z0, x1, y0) + wy * at(z0, x1, y1));
    const float v1 = (1.f - wx) * ((1.f - wy) * at(z1, x0, y0) + wy * at(z1, x0, y1)) +
                     wx * ((1.f - wy) * at(z1, x1, y0) + wy * at(z1, x1, y1));
    const double Vc = (double)Zc * Xc * Yc;
    const double ms = stats[1] / Vc;
    p[i] = (float)((double)((1.f - wz) * v0 + wz * v1) * sqrt(ms) * (double)len2);
}

__global__ void zero_kernel(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace unet
}  // namespace whhd

// ---------------------------------------------------------------------------
// Network and handle
// ---------------------------------------------------------------------------
namespace {
using whhd::unet::bf16;

// Layer table: (CA, CB, upsample a, CO, level).  CO of the last layer is 16
// with rows 1..15 held at zero.
struct LayerDef {
    int ca, cb;
    bool up;
    int co, level;
};
constexpr int kLayers = 11;
constexpr LayerDef kNet[kLayers] = {
    {16, 0, false, 16, 0},  // e0a = relu(conv(x))
    {16, 0, false, 16, 0},  // e0
    {16, 0, false, 32, 1},  // e1a = relu(conv(pool(e0)))
    {32, 0, false, 32, 1},  // e1
    {32, 0, false, 64, 2},  // ba = relu(conv(pool(e1)))
    {64, 0, false, 64, 2},  // b
    {64, 32, true, 32, 1},  // d1a = relu(conv([up(b), e1]))
    {32, 0, false, 32, 1},  // d1
    {32, 16, true, 16, 0},  // d0a = relu(conv([up(d1), e0]))
    {16, 0, false, 16, 0},  // d0
    {16, 0, false, 16, 0},  // out = conv(d0), channel 0
};
}  // namespace

struct whhd_unet_s {
    int nx, ny, nz;   // network grid
    int factor = 1;   // fine grid = factor * network grid
    float *fc = nullptr, *lsc = nullptr, *pc = nullptr;  // inputs averaged to the network grid
    int64_t V[3];
    int64_t woff[kLayers], boff[kLayers], nparams = 0;
    float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    uint8_t* frozen = nullptr;
    bf16 *wk[kLayers] = {}, *wt[kLayers] = {};
    // activations
    bf16 *x = nullptr, *e0a = nullptr, *e0 = nullptr, *p0 = nullptr, *e1a = nullptr, *e1 = nullptr,
         *p1 = nullptr, *ba = nullptr, *b = nullptr, *d1a = nullptr, *d1 = nullptr, *d0a = nullptr,
         *d0 = nullptr;
    float* out = nullptr;
    // gradients
    bf16 *g0a = nullptr, *g0b = nullptr, *u0 = nullptr, *ge0 = nullptr, *g1a = nullptr,
         *g1b = nullptr, *u1 = nullptr, *ge1 = nullptr, *g1c = nullptr, *g2a = nullptr,
         *g2b = nullptr, *g2c = nullptr;
    double* part = nullptr;   // partial sums (2 * kParts) + stats[2] + lsum[2]
    double* host = nullptr;   // pinned copy of lsum
    int64_t step = 0;
    float gscale = 1.f;  // loss-gradient scale (divided out before Adam)
    double out_scale = 1.0;  // output units: p = output * rms(f) * L^2 * out_scale
    std::vector<void*> owned;
};

namespace whhd {
namespace unet {
namespace {

constexpr int kParts = 512;

#define TRY(e) do { int s_ = (e); if (s_ != WHHD_OK) return s_; } while (0)

int check_launch(const char* what) {
    set_last_kernel(what);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WHHD_OK : hip_fail(e, what);
}

dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <int CA, int CB, bool UP, int CO, int MODE>
void conv_launch(const ConvSrc& s, const bf16* w, const ConvOut& o, hipStream_t st) {
    const int64_t V = (int64_t)s.Z * s.X * s.Y;
    hipLaunchKernelGGL((conv3_kernel<CA, CB, UP, CO, MODE>), dim3((unsigned)((V + 127) / 128)),
                       dim3(256), 0, st, s, w, o);
}

template <int CA, int CB, bool UP, int CO>
void wgrad_launch(const ConvSrc& s, const bf16* gy, float* dw, hipStream_t st) {
    const int64_t V = (int64_t)s.Z * s.X * s.Y;
    const int64_t chunks = (V + 63) / 64;
    const unsigned bx = (unsigned)std::min<int64_t>((chunks + 3) / 4, 128);
    const size_t lds = (size_t)4 * (CO + CA + CB) * 72 * sizeof(bf16);
    hipLaunchKernelGGL((conv3_wgrad_kernel<CA, CB, UP, CO>), dim3(bx, 27), dim3(256), lds, st, s,
                       gy, dw);
}

// forward of layer l (RELU into o1, or F32 for the last)
int fwd_layer(whhd_unet_s* h, int l, const ConvSrc& s, bf16* out_bf, hipStream_t st) {
    ConvOut o{out_bf, nullptr, h->out, h->w + h->boff[l], 0};
    const bf16* w = h->wk[l];
    switch (l) {
    case 0: case 1: case 9: conv_launch<16, 0, false, 16, MODE_RELU>(s, w, o, st); break;
    case 2: conv_launch<16, 0, false, 32, MODE_RELU>(s, w, o, st); break;
    case 3: case 7: conv_launch<32, 0, false, 32, MODE_RELU>(s, w, o, st); break;
    case 4: conv_launch<32, 0, false, 64, MODE_RELU>(s, w, o, st); break;
    case 5: conv_launch<64, 0, false, 64, MODE_RELU>(s, w, o, st); break;
    case 6: conv_launch<64, 32, true, 32, MODE_RELU>(s, w, o, st); break;
    case 8: conv_launch<32, 16, true, 16, MODE_RELU>(s, w, o, st); break;
    default: conv_launch<16, 0, false, 16, MODE_F32>(s, w, o, st); break;
    }
    return check_launch("unet_conv3_kernel");
}

// input gradient of layer l from its output gradient G (CO_l channels):
// channels [0, split) into o1, [split, CIN) into o2
int bwd_data(whhd_unet_s* h, int l, const bf16* G, int Z, int X, int Y, bf16* o1, bf16* o2,
             int split, hipStream_t st) {
    const ConvSrc s{G, nullptr, Z, X, Y};
    const ConvOut o{o1, o2, nullptr, nullptr, split};
    const bf16* w = h->wt[l];
    switch (l) {
    case 1: case 9: case 10: conv_launch<16, 0, false, 16, MODE_GRAD>(s, w, o, st); break;
    case 2: conv_launch<32, 0, false, 16, MODE_GRAD>(s, w, o, st); break;
    case 3: case 7: conv_launch<32, 0, false, 32, MODE_GRAD>(s, w, o, st); break;
    case 4: conv_launch<64, 0, false, 32, MODE_GRAD>(s, w, o, st); break;
    case 5: conv_launch<64, 0, false, 64, MODE_GRAD>(s, w, o, st); break;
    case 6: conv_launch<32, 0, false, 96, MODE_GRAD>(s, w, o, st); break;
    case 8: conv_launch<16, 0, false, 48, MODE_GRAD>(s, w, o, st); break;
    default: return fail(WHHD_INVALID_ARG, "unet: no input gradient for the first layer");
    }
    return check_launch("unet_conv3_kernel(grad)");
}

int bwd_weight(whhd_unet_s* h, int l, const ConvSrc& s, const bf16* G, hipStream_t st) {
    float* dw = h->grad + h->woff[l];
    switch (l) {
    case 0: case 1: case 9: case 10: wgrad_launch<16, 0, false, 16>(s, G, dw, st); break;
    case 2: wgrad_launch<16, 0, false, 32>(s, G, dw, st); break;
    case 3: case 7: wgrad_launch<32, 0, false, 32>(s, G, dw, st); break;
    case 4: wgrad_launch<32, 0, false, 64>(s, G, dw, st); break;
    case 5: wgrad_launch<64, 0, false, 64>(s, G, dw, st); break;
    case 6: wgrad_launch<64, 32, true, 32>(s, G, dw, st); break;
    default: wgrad_launch<32, 16, true, 16>(s, G, dw, st); break;
    }
    TRY(check_launch("unet_conv3_wgrad_kernel"));
    const int C = kNet[l].co;
    const int64_t V = (int64_t)s.Z * s.X * s.Y;
    hipLaunchKernelGGL(bias_grad_kernel, dim3(256), dim3(256), 256 * sizeof(float), st, G,
                       h->grad + h->boff[l], C, V);
    return check_launch("unet_bias_grad_kernel");
}

int pool(bf16* out, const bf16* in, int Zc, int Xc, int Yc, int C, bool mean, hipStream_t st) {
    const int64_t n = (int64_t)Zc * Xc * Yc * C;
    if (mean) hipLaunchKernelGGL(pool_kernel<true>, g1(n), dim3(256), 0, st, out, in, Zc, Xc, Yc, C);
    else hipLaunchKernelGGL(pool_kernel<false>, g1(n), dim3(256), 0, st, out, in, Zc, Xc, Yc, C);
    return check_launch("unet_pool_kernel");
}

int mask(bf16* g, const bf16* y, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(relu_mask_kernel, g1(n), dim3(256), 0, st, g, y, n);
    return check_launch("unet_relu_mask_kernel");
}

int pool_back(bf16* fine, const bf16* coarse, int Z, int X, int Y, int C, hipStream_t st) {
    hipLaunchKernelGGL(pool_back_kernel, g1((int64_t)Z * X * Y * C), dim3(256), 0, st, fine, coarse,
                       Z, X, Y, C);
    return check_launch("unet_pool_back_kernel");
}

double* stats(whhd_unet_s* h) { return h->part + 2 * kParts; }
double* lsum(whhd_unet_s* h) { return h->part + 2 * kParts + 2; }

int forward(whhd_unet_s* h, const float* f, const float* ls, float len, hipStream_t st) {
    const int Z = h->nz, X = h->nx, Y = h->ny;
    const int64_t V0 = h->V[0];
    hipLaunchKernelGGL(sums_kernel<false>, dim3(kParts), dim3(256), 0, st, h->part, f, f,
                       (const double*)nullptr, 0.f, V0);
    hipLaunchKernelGGL(sums_final_kernel, dim3(1), dim3(256), 0, st, stats(h), h->part, kParts);
    hipLaunchKernelGGL(input_kernel, g1(V0), dim3(256), 0, st, h->x, f, ls, stats(h), 1.f / len, V0);
    TRY(check_launch("unet_input_kernel"));
    TRY(fwd_layer(h, 0, {h->x, nullptr, Z, X, Y}, h->e0a, st));
    TRY(fwd_layer(h, 1, {h->e0a, nullptr, Z, X, Y}, h->e0, st));
    TRY(pool(h->p0, h->e0, Z / 2, X / 2, Y / 2, 16, true, st));
    TRY(fwd_layer(h, 2, {h->p0, nullptr, Z / 2, X / 2, Y / 2}, h->e1a, st));
    TRY(fwd_layer(h, 3, {h->e1a, nullptr, Z / 2, X / 2, Y / 2}, h->e1, st));
    TRY(pool(h->p1, h->e1, Z / 4, X / 4, Y / 4, 32, true, st));
    TRY(fwd_layer(h, 4, {h->p1, nullptr, Z / 4, X / 4, Y / 4}, h->ba, st));
    TRY(fwd_layer(h, 5, {h->ba, nullptr, Z / 4, X / 4, Y / 4}, h->b, st));
    TRY(fwd_layer(h, 6, {h->b, h->e1, Z / 2, X / 2, Y / 2}, h->d1a, st));
    TRY(fwd_layer(h, 7, {h->d1a, nullptr, Z / 2, X / 2, Y / 2}, h->d1, st));
    TRY(fwd_layer(h, 8, {h->d1, h->e0, Z, X, Y}, h->d0a, st));
    TRY(fwd_layer(h, 9, {h->d0a, nullptr, Z, X, Y}, h->d0, st));
    return fwd_layer(h, 10, {h->d0, nullptr, Z, X, Y}, nullptr, st);
}

// loss = mean((e - e_bar)^2) into lsum (sum e, sum e^2), its gradient into
// h->grad (zeroed first)
int backward(whhd_unet_s* h, const float* p, float len, hipStream_t st) {
    const int Z = h->nz, X = h->nx, Y = h->ny;
    const int Z1 = Z / 2, X1 = X / 2, Y1 = Y / 2, Z2 = Z / 4, X2 = X / 4, Y2 = Y / 4;
    const int64_t V0 = h->V[0], V1 = h->V[1], V2 = h->V[2];
    const float len2 = (float)((double)len * len * h->out_scale);
    hipLaunchKernelGGL(zero_kernel, g1(h->nparams), dim3(256), 0, st, h->grad, h->nparams);
    hipLaunchKernelGGL(sums_kernel<true>, dim3(kParts), dim3(256), 0, st, h->part, h->out, p,
                       stats(h), len2, V0);
    hipLaunchKernelGGL(sums_final_kernel, dim3(1), dim3(256), 0, st, lsum(h), h->part, kParts);
    hipLaunchKernelGGL(loss_grad_kernel, g1(V0), dim3(256), 0, st, h->g0a, h->out, p, stats(h),
                       lsum(h), len2, h->gscale, V0);
    TRY(check_launch("unet_loss_kernels"));
    // out = conv(d0)
    TRY(bwd_weight(h, 10, {h->d0, nullptr, Z, X, Y}, h->g0a, st));
    TRY(bwd_data(h, 10, h->g0a, Z, X, Y, h->g0b, nullptr, 16, st));
    // d0 = relu(conv(d0a))
    TRY(mask(h->g0b, h->d0, V0 * 16, st));
    TRY(bwd_weight(h, 9, {h->d0a, nullptr, Z, X, Y}, h->g0b, st));
    TRY(bwd_data(h, 9, h->g0b, Z, X, Y, h->g0a, nullptr, 16, st));
    // d0a = relu(conv([up(d1), e0]))
    TRY(mask(h->g0a, h->d0a, V0 * 16, st));
    TRY(bwd_weight(h, 8, {h->d1, h->e0, Z, X, Y}, h->g0a, st));
    TRY(bwd_data(h, 8, h->g0a, Z, X, Y, h->u0, h->ge0, 32, st));
    TRY(pool(h->g1a, h->u0, Z1, X1, Y1, 32, false, st));  // up backward: sum of the children
    // d1 = relu(conv(d1a))
    TRY(mask(h->g1a, h->d1, V1 * 32, st));
    TRY(bwd_weight(h, 7, {h->d1a, nullptr, Z1, X1, Y1}, h->g1a, st));
    TRY(bwd_data(h, 7, h->g1a, Z1, X1, Y1, h->g1b, nullptr, 32, st));
    // d1a = relu(conv([up(b), e1]))
    TRY(mask(h->g1b, h->d1a, V1 * 32, st));
    TRY(bwd_weight(h, 6, {h->b, h->e1, Z1, X1, Y1}, h->g1b, st));
    TRY(bwd_data(h, 6, h->g1b, Z1, X1, Y1, h->u1, h->ge1, 64, st));
    TRY(pool(h->g2a, h->u1, Z2, X2, Y2, 64, false, st));
    // b = relu(conv(ba))
    TRY(mask(h->g2a, h->b, V2 * 64, st));
    TRY(bwd_weight(h, 5, {h->ba, nullptr, Z2, X2, Y2}, h->g2a, st));
    TRY(bwd_data(h, 5, h->g2a, Z2, X2, Y2, h->g2b, nullptr, 64, st));
    // ba = relu(conv(pool(e1)))
    TRY(mask(h->g2b, h->ba, V2 * 64, st));
    TRY(bwd_weight(h, 4, {h->p1, nullptr, Z2, X2, Y2}, h->g2b, st));
    TRY(bwd_data(h, 4, h->g2b, Z2, X2, Y2, h->g2c, nullptr, 32, st));
    TRY(pool_back(h->ge1, h->g2c, Z1, X1, Y1, 32, st));
    // e1 = relu(conv(e1a))
    TRY(mask(h->ge1, h->e1, V1 * 32, st));
    TRY(bwd_weight(h, 3, {h->e1a, nullptr, Z1, X1, Y1}, h->ge1, st));
    TRY(bwd_data(h, 3, h->ge1, Z1, X1, Y1, h->g1a, nullptr, 32, st));
    // e1a = relu(conv(pool(e0)))
    TRY(mask(h->g1a, h->e1a, V1 * 32, st));
    TRY(bwd_weight(h, 2, {h->p0, nullptr, Z1, X1, Y1}, h->g1a, st));
    TRY(bwd_data(h, 2, h->g1a, Z1, X1, Y1, h->g1c, nullptr, 16, st));
    TRY(pool_back(h->ge0, h->g1c, Z, X, Y, 16, st));
    // e0 = relu(conv(e0a))
    TRY(mask(h->ge0, h->e0, V0 * 16, st));
    TRY(bwd_weight(h, 1, {h->e0a, nullptr, Z, X, Y}, h->ge0, st));
    TRY(bwd_data(h, 1, h->ge0, Z, X, Y, h->g0b, nullptr, 16, st));
    // e0a = relu(conv(x))
    TRY(mask(h->g0b, h->e0a, V0 * 16, st));
    return bwd_weight(h, 0, {h->x, nullptr, Z, X, Y}, h->g0b, st);
}

// `in` on the network grid: itself, or its F^3 block means in `buf`.
const float* coarse(whhd_unet_s* h, float* buf, const float* in, hipStream_t st) {
    if (h->factor == 1) return in;
    hipLaunchKernelGGL(down_kernel, g1(h->V[0]), dim3(256), 0, st, buf, in, h->nz, h->nx, h->ny,
                       h->factor);
    return buf;
}

int pack_all(whhd_unet_s* h, hipStream_t st) {
    for (int l = 0; l < kLayers; ++l) {
        const int cin = kNet[l].ca + kNet[l].cb, co = kNet[l].co;
        hipLaunchKernelGGL(pack_kernel, g1((int64_t)co * kTaps * cin), dim3(256), 0, st, h->wk[l],
                           h->wt[l], h->w + h->woff[l], co, cin);
    }
    return check_launch("unet_pack_kernel");
}

// deterministic He-normal initial weights (splitmix64 + Box-Muller), zero biases
void init_weights(std::vector<float>& w, const whhd_unet_s* h, uint64_t seed) {
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto unif = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
    w.assign(h->nparams, 0.f);
    for (int l = 0; l < kLayers; ++l) {
        const int cin = kNet[l].ca + kNet[l].cb;
        const int co_real = l == kLayers - 1 ? 1 : kNet[l].co;
        const int cin_real = l == 0 ? 2 : cin;
        const double sd = std::sqrt(2.0 / (kTaps * cin_real));
        for (int co = 0; co < co_real; ++co)
            for (int t = 0; t < kTaps; ++t)
                for (int ci = 0; ci < cin_real; ++ci) {
                    const double u1 = unif(), u2 = unif();
                    const double n = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                    w[h->woff[l] + ((int64_t)co * kTaps + t) * cin + ci] = (float)(sd * n);
                }
    }
}

}  // namespace
}  // namespace unet
}  // namespace whhd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {
void unet_release(whhd_unet_s* h) {
    if (!h) return;
    for (void* p : h->owned) (void)hipFree(p);
    if (h->host) (void)hipHostFree(h->host);
    delete h;
}
}  // namespace

extern "C" {

WHHD_API int whhd_unet_create(int nx, int ny, int nz, int factor, unsigned long long seed,
                              whhd_unet_t* out) {
    using namespace whhd;
    using namespace whhd::unet;
    if (!out) return fail(WHHD_INVALID_ARG, "unet_create: null handle pointer");
    *out = nullptr;
    if (factor < 1 || factor > 16) return fail(WHHD_INVALID_ARG, "unet_create: factor out of range");
    const int q = 4 * factor;
    if (nx < q || ny < q || nz < q || nx % q || ny % q || nz % q)
        return fail(WHHD_INVALID_ARG, "unet_create: extents must be multiples of 4 * factor");
    if ((int64_t)nx * ny * nz * 96 >= ((int64_t)1 << 40))
        return fail(WHHD_INVALID_ARG, "unet_create: grid too large");
    auto* h = new whhd_unet_s;
    h->factor = factor;
    nx /= factor;
    ny /= factor;
    nz /= factor;
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->V[0] = (int64_t)nx * ny * nz;
    h->V[1] = h->V[0] / 8;
    h->V[2] = h->V[0] / 64;
    int64_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
        const int cin = kNet[l].ca + kNet[l].cb;
        h->woff[l] = off;
        off += (int64_t)kNet[l].co * kTaps * cin;
        h->boff[l] = off;
        off += kNet[l].co;
    }
    h->nparams = off;
    bool ok = true;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (!ok || hipMalloc(&p, bytes) != hipSuccess) {
            ok = false;
            return nullptr;
        }
        h->owned.push_back(p);
        return p;
    };
    const size_t P = (size_t)h->nparams;
    h->w = (float*)alloc(P * 4);
    h->m1 = (float*)alloc(P * 4);
    h->m2 = (float*)alloc(P * 4);
    h->grad = (float*)alloc(P * 4);
    h->frozen = (uint8_t*)alloc(P);
    for (int l = 0; l < kLayers; ++l) {
        const int cin = kNet[l].ca + kNet[l].cb, co = kNet[l].co;
        h->wk[l] = (bf16*)alloc((size_t)co * kpad(cin) * 2);
        h->wt[l] = (bf16*)alloc((size_t)cin * kpad(co) * 2);
    }
    const int64_t V0 = h->V[0], V1 = h->V[1], V2 = h->V[2];
    auto act = [&](int64_t v, int c) { return (bf16*)alloc((size_t)v * c * 2); };
    h->x = act(V0, 16); h->e0a = act(V0, 16); h->e0 = act(V0, 16); h->p0 = act(V1, 16);
    h->e1a = act(V1, 32); h->e1 = act(V1, 32); h->p1 = act(V2, 32); h->ba = act(V2, 64);
    h->b = act(V2, 64); h->d1a = act(V1, 32); h->d1 = act(V1, 32); h->d0a = act(V0, 16);
    h->d0 = act(V0, 16);
    h->out = (float*)alloc((size_t)V0 * 4);
    h->g0a = act(V0, 16); h->g0b = act(V0, 16); h->u0 = act(V0, 32); h->ge0 = act(V0, 16);
    h->g1a = act(V1, 32); h->g1b = act(V1, 32); h->u1 = act(V1, 64); h->ge1 = act(V1, 32);
    h->g1c = act(V1, 16); h->g2a = act(V2, 64); h->g2b = act(V2, 64); h->g2c = act(V2, 32);
    h->part = (double*)alloc((2 * kParts + 4) * sizeof(double));
    if (factor > 1) {
        h->fc = (float*)alloc((size_t)V0 * 4);
        h->lsc = (float*)alloc((size_t)V0 * 4);
        h->pc = (float*)alloc((size_t)V0 * 4);
    }
    if (ok && hipHostMalloc((void**)&h->host, 4 * sizeof(double)) != hipSuccess) {
        h->host = nullptr;
        ok = false;
    }
    if (!ok) {
        unet_release(h);
        return fail(WHHD_OOM, "unet_create: allocation failed");
    }
    // zero everything the kernels do not always write (padded weight columns),
    // set the initial weights and the frozen mask (the last layer's rows 1..15)
    std::vector<float> w;
    init_weights(w, h, seed);
    std::vector<uint8_t> fr(P, 0);
    {
        const int l = kLayers - 1, cin = kNet[l].ca + kNet[l].cb;
        for (int co = 1; co < kNet[l].co; ++co) {
            for (int64_t k = 0; k < (int64_t)kTaps * cin; ++k) fr[h->woff[l] + co * kTaps * cin + k] = 1;
            fr[h->boff[l] + co] = 1;
        }
    }
    hipError_t e = hipSuccess;
    for (int l = 0; l < kLayers && e == hipSuccess; ++l) {
        const int cin = kNet[l].ca + kNet[l].cb, co = kNet[l].co;
        e = hipMemset(h->wk[l], 0, (size_t)co * kpad(cin) * 2);
        if (e == hipSuccess) e = hipMemset(h->wt[l], 0, (size_t)cin * kpad(co) * 2);
    }
    if (e == hipSuccess) e = hipMemcpy(h->w, w.data(), P * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->frozen, fr.data(), P, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->m1, 0, P * 4);
    if (e == hipSuccess) e = hipMemset(h->m2, 0, P * 4);
    if (e == hipSuccess) e = hipMemset(h->grad, 0, P * 4);
    if (e != hipSuccess) {
        unet_release(h);
        return hip_fail(e, "unet_create: initialisation");
    }
    int st = pack_all(h, nullptr);
    if (st == WHHD_OK && hipDeviceSynchronize() != hipSuccess) st = fail(WHHD_HIP_ERROR, "unet_create: sync");
    if (st != WHHD_OK) {
        unet_release(h);
        return st;
    }
    *out = h;
    return WHHD_OK;
}

WHHD_API long long whhd_unet_num_params(whhd_unet_t h) { return h ? h->nparams : -1; }

WHHD_API int whhd_unet_get_params(whhd_unet_t h, float* dst, int which) {
    using namespace whhd;
    if (!h || !dst) return fail(WHHD_INVALID_ARG, "unet_get_params: null argument");
    const float* src = which == 0 ? h->w : which == 1 ? h->grad : nullptr;
    if (!src) return fail(WHHD_INVALID_ARG, "unet_get_params: which is 0 (weights) or 1 (gradients)");
    hipError_t e = hipMemcpy(dst, src, (size_t)h->nparams * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "unet_get_params");
}

WHHD_API int whhd_unet_set_params(whhd_unet_t h, const float* src) {
    using namespace whhd;
    if (!h || !src) return fail(WHHD_INVALID_ARG, "unet_set_params: null argument");
    hipError_t e = hipMemcpy(h->w, src, (size_t)h->nparams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->m1, 0, (size_t)h->nparams * 4);
    if (e == hipSuccess) e = hipMemset(h->m2, 0, (size_t)h->nparams * 4);
    if (e != hipSuccess) return hip_fail(e, "unet_set_params");
    h->step = 0;
    int st = unet::pack_all(h, nullptr);
    if (st == WHHD_OK && hipDeviceSynchronize() != hipSuccess)
        st = fail(WHHD_HIP_ERROR, "unet_set_params: sync");
    return st;
}

WHHD_API int whhd_unet_predict(whhd_unet_t h, float* p0, const float* f, const float* ls,
                               double length, void* stream) {
    using namespace whhd;
    if (!h || !p0 || !f || !ls || !(length > 0)) return fail(WHHD_INVALID_ARG, "unet_predict: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const float* fin = unet::coarse(h, h->fc, f, st);
    const float* lin = unet::coarse(h, h->lsc, ls, st);
    int r = unet::forward(h, fin, lin, (float)length, st);
    if (r != WHHD_OK) return r;
    const float len2 = (float)(length * length * h->out_scale);
    if (h->factor == 1)
        hipLaunchKernelGGL(unet::output_kernel, unet::g1(h->V[0]), dim3(256), 0, st, p0, h->out,
                           unet::stats(h), len2, h->V[0]);
    else
        hipLaunchKernelGGL(unet::up_kernel,
                           dim3((h->ny * h->factor + 63) / 64, (h->nx * h->factor + 3) / 4, h->nz * h->factor),
                           dim3(64, 4), 0, st, p0, h->out, unet::stats(h), len2, h->nz, h->nx, h->ny,
                           h->factor);
    return unet::check_launch("unet_output_kernel");
}

WHHD_API int whhd_unet_loss_grad(whhd_unet_t h, const float* f, const float* ls, const float* p,
                                 double length, double* loss, void* stream) {
    using namespace whhd;
    if (!h || !f || !ls || !p || !(length > 0)) return fail(WHHD_INVALID_ARG, "unet_loss_grad: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const float* fin = unet::coarse(h, h->fc, f, st);
    const float* lin = unet::coarse(h, h->lsc, ls, st);
    const float* pin = unet::coarse(h, h->pc, p, st);
    int r = unet::forward(h, fin, lin, (float)length, st);
    if (r == WHHD_OK) r = unet::backward(h, pin, (float)length, st);
    if (r != WHHD_OK) return r;
    if (loss) {
        hipError_t e = hipMemcpyAsync(h->host, unet::lsum(h), 2 * sizeof(double),
                                      hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "unet_loss_grad: read-back");
        const double V = (double)h->V[0], m = h->host[0] / V;
        *loss = h->host[1] / V - m * m;
    }
    return WHHD_OK;
}

WHHD_API int whhd_unet_adam(whhd_unet_t h, float lr, void* stream) {
    using namespace whhd;
    if (!h || !(lr >= 0)) return fail(WHHD_INVALID_ARG, "unet_adam: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const float b1 = 0.9f, b2 = 0.999f;
    h->step += 1;
    const float c1 = 1.f - std::pow(b1, (float)h->step), c2 = 1.f - std::pow(b2, (float)h->step);
    hipLaunchKernelGGL(unet::adam_kernel, unet::g1(h->nparams), dim3(256), 0, st, h->w, h->m1, h->m2,
                       h->grad, h->frozen, h->nparams, lr, b1, b2, 1e-8f, c1, c2, 1.f / h->gscale);
    int r = unet::check_launch("unet_adam_kernel");
    return r == WHHD_OK ? unet::pack_all(h, st) : r;
}

WHHD_API int whhd_unet_set_output_scale(whhd_unet_t h, double scale) {
    using namespace whhd;
    if (!h || !(scale > 0)) return fail(WHHD_INVALID_ARG, "unet_set_output_scale: bad argument");
    h->out_scale = scale;
    return WHHD_OK;
}

WHHD_API double whhd_unet_output_scale(whhd_unet_t h) { return h ? h->out_scale : 0.0; }

WHHD_API void whhd_unet_destroy(whhd_unet_t h) { unet_release(h); }

}  // extern "C"
