// Status codes, argument checks and small shared helpers of libwhhd.so.
#include <atomic>
#include <cstdio>
#include <string>

#include "whhd_internal.h"

namespace whhd {

namespace {
thread_local std::string g_last_error;
}

std::atomic<int> g_line_kernel_mode{0};

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

int hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return WHHD_HIP_ERROR;
}

int ilog2(int n) {
    int l = 0;
    while ((2LL << l) <= n) ++l;
    return l;
}

char proj_bc(char bc) {
    // WaveletPotentials/WaveletTypes.h:102-118
    if (bc == 'n') return 'd';
    if (bc == 'd') return 'n';
    return bc;
}

int proj_wavelet(int wt) {
    // WaveletPotentials/WaveletTypes.h:120-143
    switch (wt) {
    case WT_CDF3_9: return WT_CDF2_10;
    case WT_CDF3_11: return WT_CDF2_12;
    case WT_CDF3_5: return WT_CDF2_6;
    case WT_CDF4_4: return WT_CDF3_5;
    case WT_CDF4_6: return WT_CDF3_7;
    case WT_CDF3_7: return WT_CDF2_8;
    case WT_CDF4_8: return WT_CDF3_9;
    case WT_CDF4_10: return WT_CDF3_11;
    case WT_CDF2_0: return WT_CDF1_1;
    default: return -1;
    }
}

// Registered (type, bc) keys of WaveletPotentials/cuWHHDForward.cu:744-1080.
int check_wavelet(int wt, char bc) {
    const bool nd = (bc == 'n' || bc == 'd');
    const bool ndz = nd || bc == 'z';
    bool ok;
    switch (wt) {
    case WT_CDF3_5: case WT_CDF3_7: case WT_CDF2_6: case WT_CDF2_8: case WT_CDF4_4: case WT_CDF4_6:
        ok = ndz; break;
    case WT_CDF2_0: case WT_CDF1_1: case WT_CDF1_7:
        ok = nd; break;
    default:
        ok = false;
    }
    if (!ok) {
        char buf[96];
        std::snprintf(buf, sizeof buf, "wavelet %d with bc '%c' has no registered kernel", wt, bc);
        return fail(WHHD_UNSUPPORTED_WAVELET, buf);
    }
    return WHHD_OK;
}

static int halo_of(int wt) {
    // WaveletPotentials/cuWHHDForward.cu:1082-1101
    switch (wt) {
    case WT_CDF3_7: case WT_CDF4_6: case WT_CDF2_8: return 4;
    case WT_CDF2_0: return 1;
    case WT_CDF1_1: return 0;
    default: return 3;
    }
}

// Half-sample symmetric families index lines of 2^L samples; whole-sample
// ones lines of 2^L + 1 samples (the end samples are the boundary nodes).
// The other parity makes the reference read outside the line.
static bool wants_odd(int wt) {
    switch (wt) {
    case WT_CDF2_6: case WT_CDF2_8: case WT_CDF4_4: case WT_CDF4_6: case WT_CDF2_0: return true;
    default: return false;
    }
}

int check_line(int n, int levels, int wt) {
    char buf[160];
    if (n < 4) {
        std::snprintf(buf, sizeof buf, "line length %d: at least 4 samples are needed", n);
        return fail(WHHD_INVALID_ARG, buf);
    }
    const int L = ilog2(n);
    if ((1 << L) != n - (n & 1)) {
        std::snprintf(buf, sizeof buf, "line length %d is neither 2^L nor 2^L+1", n);
        return fail(WHHD_NOT_POW2, buf);
    }
    if (levels != L) {
        std::snprintf(buf, sizeof buf, "level %d does not match line length %d (expected %d)", levels,
                      n, L);
        return fail(WHHD_INVALID_ARG, buf);
    }
    if (wants_odd(wt) != bool(n & 1)) {
        std::snprintf(buf, sizeof buf,
                      "wavelet %d is defined on lines of %s length; got %d samples", wt,
                      wants_odd(wt) ? "odd (2^L+1)" : "even (2^L)", n);
        return fail(WHHD_UNSUPPORTED_WAVELET, buf);
    }
    // The inverse reuses the staged input buffer for the output of a level
    // once its details are consumed; window j of level l may only start
    // writing where the remaining windows no longer read details.
    const int H = halo_of(wt), wl = 64 - 2 * H, ext = n & 1;
    for (int l = 2; l < levels; ++l) {
        const int Lv = 1 << l;
        for (int j = 1; j * wl < Lv + ext; ++j)
            if (j * wl > Lv + ext - H) {
                std::snprintf(buf, sizeof buf, "line length %d unsupported by the in-place inverse", n);
                return fail(WHHD_INVALID_ARG, buf);
            }
    }
    return WHHD_OK;
}

TileConfig default_tile(int n, size_t elem, bool strided) {
    TileConfig c;
    const size_t row = (size_t)((n & 1) ? n : n + 1) * elem * 2;  // two buffers
    const size_t budget = 40 * 1024;
    int t = (int)(budget / row);
    if (t < 1) t = 1;
    if (strided) {
        int s = 64;
        while (s > 1 && (size_t)s * row > budget) s >>= 1;
        c.strided_tile = s;
    }
    c.contig_tile = t > 16 ? 16 : t;
    c.threads = 256;
    return c;
}

}  // namespace whhd

extern "C" {

WHHD_API const char* whhd_status_string(int status) {
    switch (status) {
    case WHHD_OK: return "ok";
    case WHHD_INVALID_ARG: return "invalid argument";
    case WHHD_NOT_POW2: return "size is not a power of two (plus one on staggered axes)";
    case WHHD_UNSUPPORTED_WAVELET: return "unsupported wavelet type / boundary condition";
    case WHHD_HIP_ERROR: return "HIP runtime error";
    case WHHD_OOM: return "out of device memory";
    case WHHD_NONCONVERGED: return "did not converge";
    default: return "unknown status";
    }
}

WHHD_API const char* whhd_last_error(void) { return whhd::g_last_error.c_str(); }

WHHD_API const char* whhd_version(void) { return "whhd-mi355x 0.1 (gfx950)"; }

WHHD_API char whhd_proj_bc(char bc) { return whhd::proj_bc(bc); }

WHHD_API int whhd_proj_wavelet(int wt) { return whhd::proj_wavelet(wt); }

WHHD_API int whhd_set_line_kernel(int mode) {
    if (mode != 0 && mode != 1) return whhd::fail(WHHD_INVALID_ARG, "line kernel mode must be 0 or 1");
    whhd::g_line_kernel_mode.store(mode);
    return WHHD_OK;
}

WHHD_API int whhd_wavelet_supported(int wt, char bc) {
    return whhd::check_wavelet(wt, bc) == WHHD_OK ? 1 : 0;
}

WHHD_API int whhd_transform1d(float* dst, const float* src, int nx, int ny, int nz, int level,
                              char axis, char bc, int wt, int forward, void* stream) {
    return whhd::transform_lines<float>(dst, src, nx, ny, nz, level, axis, bc, wt, forward != 0,
                                        (hipStream_t)stream);
}

WHHD_API int whhd_transform1d_f64(double* dst, const double* src, int nx, int ny, int nz,
                                  int level, char axis, char bc, int wt, int forward,
                                  void* stream) {
    return whhd::transform_lines<double>(dst, src, nx, ny, nz, level, axis, bc, wt, forward != 0,
                                         (hipStream_t)stream);
}

}  // extern "C"
