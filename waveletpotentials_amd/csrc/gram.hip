// Gram tables of the synthesis basis, the data of CuGetInnerProduct::GenerateIp
// (WaveletPotentials/cuGetInnerProduct.cu:92-126, row selection :11-77).
//
// The basis functions are the inverse transforms of the unit vectors, made by
// the FP64 line-transform kernel on an identity matrix; the eight selected
// entries of each row of G = B B^T are dot products of two basis rows,
// computed by one wavefront per row with an FP64 cross-lane reduction.  The
// reference forms the whole product through float tiles; this is exact FP64.
#include <hip/hip_runtime.h>

#include "whhd_internal.h"

namespace whhd {

__global__ void identity_kernel(double* m, int n) {
    const int64_t total = (int64_t)n * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / n, c = e - r * n;
        m[e] = (r == c) ? 1.0 : 0.0;
    }
}

__device__ double row_dot(const double* b, int n, int i, int c) {
    double s = 0.0;
    for (int t = threadIdx.x & 63; t < n; t += 64) s += b[(int64_t)i * n + t] * b[(int64_t)c * n + t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// One wavefront per basis row i; writes the 8-row table layout.
__global__ void gram_select_kernel(double* out, const double* b, int n) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int ext = n & 1;
    const int nei = ((i - ext) >> 1) + ext;
    double v[8];
    v[0] = row_dot(b, n, i, i);
    v[1] = i > 1 ? row_dot(b, n, i, i - 2) : 0.0;
    v[2] = i > 0 ? row_dot(b, n, i, i - 1) : 0.0;
    v[3] = i < n - 1 ? row_dot(b, n, i, i + 1) : 0.0;
    v[4] = i < n - 2 ? row_dot(b, n, i, i + 2) : 0.0;
    v[5] = nei > 0 ? row_dot(b, n, i, nei - 1) : 0.0;  // the reference leaves this unwritten
    v[6] = row_dot(b, n, i, nei);
    v[7] = nei + 1 < n ? row_dot(b, n, i, nei + 1) : 0.0;
    if (lane == 0)
        for (int r = 0; r < 8; ++r) out[(int64_t)r * n + i] = v[r];
}

int gram_table_device(double* out, int wt, char bc, int n, hipStream_t stream) {
    int st = check_wavelet(wt, bc);
    if (st != WHHD_OK) return st;
    if (!out) return fail(WHHD_INVALID_ARG, "gram table: null output");
    st = check_line(n, ilog2(n), wt);
    if (st != WHHD_OK) return st;
    double* basis = nullptr;
    hipError_t e = hipMalloc(&basis, sizeof(double) * (size_t)n * n);
    if (e != hipSuccess) return fail(WHHD_OOM, "gram table: basis allocation failed");
    hipLaunchKernelGGL(identity_kernel, dim3(256), dim3(256), 0, stream, basis, n);
    st = transform_lines<double>(basis, basis, n, n, 1, ilog2(n), 'y', bc, wt, false, stream);
    if (st == WHHD_OK) {
        hipLaunchKernelGGL(gram_select_kernel, dim3(n), dim3(64), 0, stream, out, basis, n);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) st = hip_fail(e, "gram table");
    }
    hipFree(basis);
    return st;
}

}  // namespace whhd

extern "C" WHHD_API int whhd_gram_table(double* out, int wavelet, char bc, int n, void* stream) {
    return whhd::gram_table_device(out, wavelet, bc, n, (hipStream_t)stream);
}
