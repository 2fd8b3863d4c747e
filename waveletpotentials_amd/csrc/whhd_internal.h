// Internal declarations shared by the HIP translation units of libwhhd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/whhd.h"

namespace whhd {

enum {
    WT_CDF3_9 = 0, WT_CDF3_11, WT_CDF3_5, WT_CDF2_10, WT_CDF2_12, WT_CDF2_6, WT_CDF4_4,
    WT_CDF4_6, WT_CDF4_8, WT_CDF4_10, WT_CDF2_8, WT_CDF3_7, WT_CDF26_44, WT_CDF28_46,
    WT_CDF44_26, WT_CDF46_28, WT_CDF2_0, WT_CDF1_1, WT_CDF1_7
};

constexpr size_t kMaxLdsBytes = 160 * 1024;  // LDS per workgroup on gfx950

// Geometry of one line-transform launch.
struct LineGeom {
    int n = 0;        // samples per line
    int levels = 0;   // floor(log2(n))
    int lds_row = 0;  // padded LDS row length
    int tile = 0;     // lines per workgroup
    int64_t lines = 0;          // contiguous: total lines; strided: lines per outer index (ny)
    int outer = 1;              // strided: number of outer indices
    int64_t sample_stride = 1;  // strided: distance between consecutive samples of a line
    int64_t outer_stride = 0;   // strided: distance between outer indices
};

// Launch shape of the line-transform kernels.
struct TileConfig {
    int contig_tile = 4;    // lines per workgroup, contiguous (y) lines
    int strided_tile = 16;  // consecutive y per workgroup, strided (x, z) lines
    int threads = 256;
    bool allow_line_parallel = true;  // use the line-parallel kernels where one exists
};

TileConfig default_tile(int n, size_t elem, bool strided);

// 0: line-parallel kernels where one exists, 1: windowed tile kernels only
extern std::atomic<int> g_line_kernel_mode;

// Error plumbing: set the thread's last-error message and return `status`.
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

int check_wavelet(int wt, char bc);
int check_line(int n, int levels, int wt);
int ilog2(int n);

// A batch of `nb` problems of equal extents whose arrays lie `stride` elements
// apart, transformed by one launch; `waves` wavefronts per workgroup of the
// line-parallel kernels (1, 2 or 4).
struct LineBatch {
    int nb = 1;
    int64_t stride = 0;
    int waves = 4;
};

template <class T>
int transform_lines(T* dst, const T* src, int nx, int ny, int nz, int levels, char axis, char bc,
                    int wt, bool fwd, hipStream_t stream, const TileConfig* cfg = nullptr,
                    const LineBatch* batch = nullptr);

// Line-parallel transform (dst != src, curl CDF3.7 'n' / divergence CDF4.6 'd',
// 2^5..2^8 samples); returns -1 when no such kernel applies.
int lp_transform(float* dst, const float* src, int nx, int ny, int nz, int levels, char axis,
                 char bc, int wt, bool fwd, hipStream_t stream, const LineBatch& b);

// Recovery driver over `nb` problems of equal extents stored back to back (the
// arrays of problem b start b * (array size) elements after the first);
// recover3d_set_tiles picks the launch shapes (1/2/4 wavefronts per line-pass
// workgroup, 1/2/4/8 z cells per projection thread).
int recover3d_create(int nx, int ny, int nz, double dx, double dy, double dz, char bc, int wc,
                     int wd, whhd_recover3d_s** out, int nb);
int recover3d_solve(whhd_recover3d_s* h, float* qx, float* qy, float* qz, float* xo, float* yo,
                    float* zo, const float* xv, const float* yv, const float* zv, hipStream_t s);
int recover3d_set_tiles(whhd_recover3d_s* h, int lp_waves, int proj_zb);

// Gram table (8 rows of n) of the synthesis basis of (wt, bc) into device memory.
int gram_table_device(double* out, int wt, char bc, int n, hipStream_t stream);

char proj_bc(char bc);
int proj_wavelet(int wt);
const char* wavelet_name(int wt);

// Name of the kernel the calling thread launched last (set by the dispatchers
// of the line transforms and the projection).
const char* last_kernel();
void set_last_kernel(const char* name);

// Per-stage timing of a driver's solves: HIP events recorded on the solve's
// stream after each stage, read back by trace_end.  Recording costs one event
// record per launch and no synchronization.
struct StageTrace {
    static constexpr int kMaxStages = 64;
    int cap = 0;      // solves that can be recorded (0: tracing off)
    int solves = 0;   // solves recorded so far
    int stages = 0;   // stages per solve, fixed by the first recorded solve
    int cur = -1;     // stage index within the solve being recorded, -1: none
    std::vector<hipEvent_t> ev;
    std::vector<std::string> labels;
    int begin(int max_solves);
    void solve_start(hipStream_t s);
    void mark(hipStream_t s, const char* stage);
    void solve_end();
    int end(float* stage_us, int max_stages, int* n_stages, int* n_solves);
    void release();
    ~StageTrace() { release(); }
};

}  // namespace whhd
