// Dispatch of the line-parallel kernels (line_parallel.h): the two (wavelet,
// bc) pairs of the recovery path -- curl CDF3.7 'n' on 2^L samples and
// divergence CDF4.6 'd' on 2^L+1 samples -- for L = 5..8, instantiated in
// lp_cdf37n.hip and lp_cdf46d.hip.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "line_parallel.h"
#include "whhd_internal.h"

namespace whhd {

extern template bool lp_launch<W_Cdf37N, 0>(int, bool, bool, float*, const float*,
                                            const lp::LpGeom&, int, int, hipStream_t);
extern template bool lp_launch<W_Cdf46D, 1>(int, bool, bool, float*, const float*,
                                            const lp::LpGeom&, int, int, hipStream_t);

// Returns WHHD_OK when a line-parallel kernel was launched, -1 when none fits
// (the caller then uses the tile kernel).  Needs dst != src: lines are read
// and written in one pass without staging.
int lp_transform(float* dst, const float* src, int nx, int ny, int nz, int levels, char axis,
                 char bc, int wt, bool fwd, hipStream_t stream, const LineBatch& b) {
    if (dst == src) return -1;
    lp::LpGeom g{};
    g.batch_stride = b.stride;
    int n;
    switch (axis) {
    case 'y': case 'Y':
        n = ny; g.total_lines = (int64_t)nx * nz; g.lines_inner = (int)g.total_lines;
        g.inner_stride = ny; g.outer_stride = 0; g.step = 1; break;
    case 'x': case 'X':
        n = nx; g.total_lines = (int64_t)ny * nz; g.lines_inner = ny; g.inner_stride = 1;
        g.outer_stride = (int64_t)nx * ny; g.step = ny; break;
    case 'z': case 'Z':
        n = nz; g.total_lines = (int64_t)ny * nx; g.lines_inner = ny; g.inner_stride = 1;
        g.outer_stride = ny; g.step = (int64_t)nx * ny; break;
    default: return -1;
    }
    g.total_elems = (int64_t)nx * ny * nz;
    // buffer offsets are 32-bit; keep clear of the out-of-range marker
    if (g.total_lines > (int64_t)1 << 31 || g.total_elems * 4 >= (int64_t)lp::kOutOfRange) return -1;
    bool ok = false;
    const bool contig = (axis == 'y' || axis == 'Y');
    if (wt == WT_CDF3_7 && bc == 'n' && n == (1 << levels))
        ok = lp_launch<W_Cdf37N, 0>(levels, fwd, contig, dst, src, g, b.waves, b.nb, stream);
    else if (wt == WT_CDF4_6 && bc == 'd' && n == (1 << levels) + 1)
        ok = lp_launch<W_Cdf46D, 1>(levels, fwd, contig, dst, src, g, b.waves, b.nb, stream);
    if (!ok) return -1;
    char name[96];
    std::snprintf(name, sizeof(name), "lp_kernel<%s %c,L=%d,%s,%s>", wavelet_name(wt), bc, levels,
                  fwd ? "fwd" : "inv", contig ? "contiguous" : "strided");
    set_last_kernel(name);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "line-parallel kernel");
}

}  // namespace whhd
