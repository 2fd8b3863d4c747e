/*
 * whhd.h -- C ABI of the MI355X-native wavelet potential-recovery library
 * (libwhhd.so, HIP kernels for gfx950).
 *
 * The reference (yours321dog/WaveletPotentials) exposes this path as C++
 * classes behind an empty CUWHHD_API macro (WaveletPotentials/cuWHHD_Export.h:
 * 11-20).  Each entry point below replaces one of those methods; the comment
 * on each names the reference interface (file:line).  Conventions:
 *
 *   - Every pointer argument is device memory owned by the caller, except
 *     handle out-parameters and host-side result structs.
 *   - Arrays are y-fastest: element (x, y, z) of an array with extents
 *     (nx, ny, nz) is at y + x*ny + z*nx*ny (WaveletPotentials/Utils.h:46-47).
 *     MAC staggering follows the reference: for nx*ny*nz cells,
 *     xv (nx+1, ny, nz), yv (nx, ny+1, nz), zv (nx, ny, nz+1),
 *     qx (nx, ny+1, nz+1), qy (nx+1, ny, nz+1), qz (nx+1, ny+1, nz).
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Calls enqueue work on it and return without synchronizing unless the
 *     comment says otherwise.
 *   - Functions return a whhd_status; they never abort the process.  The
 *     message of the last failure on the calling thread is whhd_last_error().
 *   - Bulk arithmetic is FP32 (float*); entry points suffixed _f64 take double*.
 */
#ifndef WHHD_H
#define WHHD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WHHD_API __attribute__((visibility("default")))

typedef enum whhd_status {
    WHHD_OK = 0,
    WHHD_INVALID_ARG = 1,
    WHHD_NOT_POW2 = 2,             /* wavelet path needs 2^L (+1 on staggered axes) samples */
    WHHD_UNSUPPORTED_WAVELET = 3,  /* (type, bc) has no kernel in the reference either */
    WHHD_HIP_ERROR = 4,
    WHHD_OOM = 5,
    WHHD_NONCONVERGED = 6
} whhd_status;

/* WaveletType, same enumerator order as WaveletPotentials/WaveletTypes.h:7-28 */
typedef enum whhd_wavelet {
    WHHD_WT_CDF3_9 = 0, WHHD_WT_CDF3_11, WHHD_WT_CDF3_5, WHHD_WT_CDF2_10, WHHD_WT_CDF2_12,
    WHHD_WT_CDF2_6, WHHD_WT_CDF4_4, WHHD_WT_CDF4_6, WHHD_WT_CDF4_8, WHHD_WT_CDF4_10,
    WHHD_WT_CDF2_8, WHHD_WT_CDF3_7, WHHD_WT_CDF26_44, WHHD_WT_CDF28_46, WHHD_WT_CDF44_26,
    WHHD_WT_CDF46_28, WHHD_WT_CDF2_0, WHHD_WT_CDF1_1, WHHD_WT_CDF1_7
} whhd_wavelet;

/* ---- library ---------------------------------------------------------- */
WHHD_API const char* whhd_status_string(int status);
WHHD_API const char* whhd_last_error(void);
WHHD_API const char* whhd_version(void);
/* proj_bc (WaveletPotentials/WaveletTypes.h:102-118) */
WHHD_API char whhd_proj_bc(char bc);
/* proj_WaveletType (WaveletPotentials/WaveletTypes.h:120-143); -1 if none */
WHHD_API int whhd_proj_wavelet(int wavelet);
/* 1 if (wavelet, bc) has a kernel (WaveletPotentials/cuWHHDForward.cu:744-1080) */
WHHD_API int whhd_wavelet_supported(int wavelet, char bc);
/* Process-wide choice of line-transform kernels (no reference counterpart; the
 * reference has a single kernel family): 0 = line-parallel register kernels
 * where one exists for the (wavelet, bc, length) and dst != src, windowed
 * LDS-tile kernels otherwise (default); 1 = windowed kernels only. */
WHHD_API int whhd_set_line_kernel(int mode);

/* ---- 1-D line transforms ---------------------------------------------- */
/* CuWHHDForward::Solve1D(dst, src, int3 dim, level, direction, bc, type, is_forward)
 * (WaveletPotentials/cuWHHDForward.cuh:20-21, cuWHHDForward.cu:1122-1253):
 * multi-level forward (forward=1) or inverse (0) transform of every line of
 * the (nx, ny, nz) array along axis 'x', 'y' or 'z'.  dst may equal src.
 * `level` must be floor(log2(line length)) and the length 2^level or
 * 2^level + 1. */
WHHD_API int whhd_transform1d(float* dst, const float* src, int nx, int ny, int nz, int level,
                              char axis, char bc, int wavelet, int forward, void* stream);
WHHD_API int whhd_transform1d_f64(double* dst, const double* src, int nx, int ny, int nz,
                                  int level, char axis, char bc, int wavelet, int forward,
                                  void* stream);

/* CuGetInnerProduct::GenerateIp(type, bc, n) (WaveletPotentials/cuGetInnerProduct.cu:
 * 92-126 with cuSelectIpPart :11-77): Gram table of the synthesis basis on n
 * samples, written as 8 rows of n doubles to device memory `out` (row 0 the
 * diagonal, rows 1-4 the +-1/+-2 off-diagonals, rows 5-7 the parent entries).
 * Computed in FP64.  Synchronizes `stream`. */
WHHD_API int whhd_gram_table(double* out, int wavelet, char bc, int n, void* stream);

/* ---- per-cell projection in wavelet space ------------------------------ */
typedef struct whhd_project3d_s* whhd_project3d_t;
/* CuDivLocalProject3D(int3 dim, int3 levels, DTYPE3 dx, WaveletType type, char bc)
 * (WaveletPotentials/cuDivLocalProject3D.cuh:11, .cu:43-72).  nx, ny, nz are
 * the cell counts (powers of two; the reference passes dim+1), wavelet/bc the
 * divergence ("base") type and its bc.  Only bc 'd' (velocity bc 'n') is
 * implemented. */
WHHD_API int whhd_project3d_create(int nx, int ny, int nz, double dx, double dy, double dz,
                                   int wavelet_div, char bc_div, whhd_project3d_t* out);
/* ProjectLocal_q_ccc(q..., v..., levels) and its 9-argument form
 * (WaveletPotentials/cuDivLocalProject3D.cuh:17-20, .cu:5273-5308, 5966-6004).
 * xv_out/yv_out/zv_out may be NULL; when given they receive s x q at the
 * projected coefficients and are left untouched elsewhere. */
WHHD_API int whhd_project3d_run(whhd_project3d_t h, float* qx, float* qy, float* qz,
                                float* xv_out, float* yv_out, float* zv_out, const float* xv,
                                const float* yv, const float* zv, void* stream);
WHHD_API void whhd_project3d_destroy(whhd_project3d_t h);

typedef struct whhd_project2d_s* whhd_project2d_t;
/* CuDivLocalProject(int2 dim, int2 levels, DTYPE2 dx, WaveletType, char bc)
 * (WaveletPotentials/cuDivLocalProject.cuh:10, .cu:26-42); nx, ny cell counts. */
WHHD_API int whhd_project2d_create(int nx, int ny, double dx, double dy, int wavelet_div,
                                   char bc_div, whhd_project2d_t* out);
/* ProjectSingle(q, xv, yv) (WaveletPotentials/cuDivLocalProject.cu:2162-2198). */
WHHD_API int whhd_project2d_run(whhd_project2d_t h, float* qz, const float* xv,
                                const float* yv, void* stream);
WHHD_API void whhd_project2d_destroy(whhd_project2d_t h);

/* ---- potential recovery ------------------------------------------------ */
typedef struct whhd_recover3d_s* whhd_recover3d_t;
/* CuWaveletPotentialRecover3D(int3 dim, DTYPE3 dx, char bc, WaveletType curl,
 * WaveletType div) (WaveletPotentials/cuWaveletPotentialRecover3D.cuh:12,
 * .cu:6-12).  Allocates the handle's scratch: one velocity-sized and one
 * potential-sized array per component (passes run out of place). */
WHHD_API int whhd_recover3d_create(int nx, int ny, int nz, double dx, double dy, double dz,
                                   char bc, int wavelet_curl, int wavelet_div,
                                   whhd_recover3d_t* out);
/* Device scratch, in bytes, that whhd_recover3d_create allocates for these
 * cell counts (the reference draws its scratch "tmp", "tmp1", "tmp2" from the
 * CuMemoryManager singleton, WaveletPotentials/cuWaveletPotentialRecover3D.cu:
 * 18-22): per component one velocity-sized array and one array of the larger
 * of the velocity and potential sizes.  Host only; needs no device. */
WHHD_API int whhd_recover3d_scratch_bytes(int nx, int ny, int nz, unsigned long long* bytes);
/* Solve(qx, qy, qz, xv, yv, zv) (cuWaveletPotentialRecover3D.cu:16-61). */
WHHD_API int whhd_recover3d_solve(whhd_recover3d_t h, float* qx, float* qy, float* qz,
                                  const float* xv, const float* yv, const float* zv,
                                  void* stream);
/* Solve(qx, qy, qz, xv_out, yv_out, zv_out, xv, yv, zv) (cuWaveletPotentialRecover3D.cu:
 * 63-121): also returns the divergence-free projected velocity. */
WHHD_API int whhd_recover3d_solve_proj(whhd_recover3d_t h, float* qx, float* qy, float* qz,
                                       float* xv_out, float* yv_out, float* zv_out,
                                       const float* xv, const float* yv, const float* zv,
                                       void* stream);
/* CutCoefficients(q_out..., q..., level) (cuWaveletPotentialRecover3D.cu:123-239). */
WHHD_API int whhd_recover3d_cut(whhd_recover3d_t h, float* qx_out, float* qy_out,
                                float* qz_out, const float* qx, const float* qy,
                                const float* qz, int level, void* stream);
/* Per-stage timing (no reference counterpart; the reference times whole
 * solves on the host, WaveletPotentials/TestLocalProjection.cpp:757-775).
 * After trace_begin, the next max_solves calls of solve / solve_proj / cut
 * record a HIP event on their stream after every launch (no synchronization).
 * trace_end waits for the last recorded event, writes the average duration of
 * each stage in microseconds to stage_us[0 .. min(n_stages, max_stages)),
 * the stage and solve counts to *n_stages / *n_solves (either may be NULL),
 * and keeps the labels ("<fwd|inv> <array> <axis> | <kernel>") readable by
 * trace_label until the next trace_begin or destroy.  Only solves with the
 * stage list of the first recorded one are averaged. */
WHHD_API int whhd_recover3d_trace_begin(whhd_recover3d_t h, int max_solves);
WHHD_API int whhd_recover3d_trace_end(whhd_recover3d_t h, float* stage_us, int max_stages,
                                      int* n_stages, int* n_solves);
WHHD_API const char* whhd_recover3d_trace_label(whhd_recover3d_t h, int stage);
WHHD_API void whhd_recover3d_destroy(whhd_recover3d_t h);

/* ---- batched recovery ---------------------------------------------------- */
typedef struct whhd_batch3d_s* whhd_batch3d_t;
/* Batched 3D recovery of n independent grids (no reference counterpart: the
 * reference recovers one grid per CuWaveletPotentialRecover3D handle,
 * WaveletPotentials/cuWaveletPotentialRecover3D.cu:6-61; each problem gets that
 * solve).  dims holds (nx, ny, nz) and spacing (dx, dy, dz) per problem; bc and
 * wavelets are shared.  Problems with equal extents and spacing form a size
 * class recovered by one launch per stage; the classes run as parallel
 * branches.  The handle owns its scratch, not the problems' arrays. */
WHHD_API int whhd_batch3d_create(int n, const int* dims, const double* spacing, char bc,
                                 int wavelet_curl, int wavelet_div, whhd_batch3d_t* out);
/* Floats in the caller's arena and, when offsets is not NULL, offsets[6*i + k]
 * (k = xv, yv, zv, qx, qy, qz) of problem i's arrays from the arena start. */
WHHD_API int whhd_batch3d_layout(whhd_batch3d_t h, long long* arena_floats, long long* offsets);
/* Recover every problem: q arrays from the velocity arrays of the arena. */
WHHD_API int whhd_batch3d_solve(whhd_batch3d_t h, float* arena, void* stream);
/* Capture the whole batch on `arena` as one HIP graph (replaces any earlier). */
WHHD_API int whhd_batch3d_capture(whhd_batch3d_t h, float* arena);
/* Launch the captured graph on `stream`. */
WHHD_API int whhd_batch3d_replay(whhd_batch3d_t h, void* stream);
/* Time each size class's launch shapes (1/2/4 wavefronts per line-pass
 * workgroup, then 1/2/4/8 z cells per projection thread) over `reps` solves of
 * the class on `arena` and keep the fastest; recaptures a captured graph.
 * Synchronizes `stream`; overwrites the arena's q arrays. */
WHHD_API int whhd_batch3d_autotune(whhd_batch3d_t h, float* arena, int reps, void* stream);
WHHD_API int whhd_batch3d_num_classes(whhd_batch3d_t h);
/* info[6] = nx, ny, nz, problems, line-pass wavefronts, projection z block. */
WHHD_API int whhd_batch3d_class_info(whhd_batch3d_t h, int cls, int* info);
/* Set a class's launch shapes by hand; drops a captured graph. */
WHHD_API int whhd_batch3d_set_tiles(whhd_batch3d_t h, int cls, int lp_waves, int proj_zb);
WHHD_API void whhd_batch3d_destroy(whhd_batch3d_t h);

typedef struct whhd_recover2d_s* whhd_recover2d_t;
/* CuWaveletPotentialRecover2D(int2 dim, DTYPE2 dx, char bc, WaveletType curl,
 * WaveletType div) (WaveletPotentials/cuWaveletPotentialRecover2D.cuh:12, .cu:6-12). */
WHHD_API int whhd_recover2d_create(int nx, int ny, double dx, double dy, char bc,
                                   int wavelet_curl, int wavelet_div, whhd_recover2d_t* out);
/* Solve(qz, xv, yv) (WaveletPotentials/cuWaveletPotentialRecover2D.cu:16-37). */
WHHD_API int whhd_recover2d_solve(whhd_recover2d_t h, float* qz, const float* xv,
                                  const float* yv, void* stream);
WHHD_API void whhd_recover2d_destroy(whhd_recover2d_t h);

/* ---- solid geometry ---------------------------------------------------- */
/* CuSolidLevelSet3D::InitLsSphere (WaveletPotentials/cuSolidLevelSet3D.cu:11-26,
 * 376-383) for a union of `count` (1..64) spheres: ls (nx, ny, nz) receives
 * min_i |p - c_i| - r_i at the cell centres p = ((x+.5)dx, (y+.5)dy, (z+.5)dz);
 * negative inside a solid.  centers (3*count) and radii (count) are host arrays. */
WHHD_API int whhd_levelset_spheres(float* ls, int nx, int ny, int nz, double dx, double dy,
                                   double dz, const double* centers, const double* radii,
                                   int count, void* stream);
/* CuSolidLevelSet3D::GetFrac(ax, ay, az) (WaveletPotentials/cuSolidLevelSet3D.cu:348-353 with
 * cuComputeFrac :28-34 and cuGetAx/Ay/Az :36-90): wet face fractions
 * ax (nx+1, ny, nz), ay (nx, ny+1, nz), az (nx, ny, nz+1) of the cell level set ls. */
WHHD_API int whhd_frac_from_ls(float* ax, float* ay, float* az, const float* ls, int nx, int ny,
                               int nz, double dx, double dy, double dz, void* stream);
/* The static CuSolidLevelSet3D::GetFrac(ax, ay, az, ls, dim, dx)
 * (WaveletPotentials/cuSolidLevelSet3D.cu:359-372 with cuGetAx/Ay/Az_divide
 * :164-317 and the samplers of Interpolation.cuh:16-29, 80-157): a face is 1
 * where the quadratic B-spline interpolant of ls at its centre exceeds the
 * spacing along its normal, 0 where it is <= minus that spacing, and otherwise
 * the share of its 128 x 128 midpoint sub-samples where the trilinear
 * interpolant of ls is positive (a multiple of 1/16384).  Extents as
 * whhd_frac_from_ls; needs at least 2 cells per axis. */
WHHD_API int whhd_frac_divide(float* ax, float* ay, float* az, const float* ls, int nx, int ny,
                              int nz, double dx, double dy, double dz, void* stream);
/* CuFastSweeping::Solve(ls, dim, dx, n_iters) (WaveletPotentials/cuFastSweeping.cu:
 * 215-257): n_iters relaxations of the eikonal update on cells with ls > -1e-12
 * (solid cells stay fixed).  Each relaxation reads the previous iterate
 * (the reference updates in place, so its result depends on thread order);
 * `scratch` is a caller-owned array of nx*ny*nz floats, distinct from ls. */
WHHD_API int whhd_eikonal_sweep(float* ls, float* scratch, int nx, int ny, int nz, double h,
                                int n_iters, void* stream);

/* ---- fractional-area multigrid ----------------------------------------- */
typedef struct whhd_mgfrac3d_s* whhd_mgfrac3d_t;
/* CuMultigridFrac3D(Ax, Ay, Az, int3 dim, DTYPE3 dx, char bc)
 * (WaveletPotentials/cuMultigridFrac3D.cuh:15, .cu:9-134): copies the face
 * fractions ax (nx+1, ny, nz), ay (nx, ny+1, nz), az (nx, ny, nz+1) and builds
 * the level hierarchy (extents halved rounding up until a side reaches 1;
 * fractions restricted, spacings summed).  bc 'n' closes the walls; 'd' puts
 * a zero Dirichlet value on the wall faces (the reference's 'z' with zero
 * exterior spacing); 'z' uses the cell spacing as exterior spacing.
 * Synchronizes the null stream. */
WHHD_API int whhd_mgfrac3d_create(const float* ax, const float* ay, const float* az, int nx,
                                  int ny, int nz, double dx, double dy, double dz, char bc,
                                  whhd_mgfrac3d_t* out);
/* Extents of the levels (3 ints each, finest first) into dims[0 .. 3*max_levels);
 * returns the number of levels (-1 for a null handle). */
WHHD_API int whhd_mgfrac3d_levels(whhd_mgfrac3d_t h, int* dims, int max_levels);
/* Solve(dst, pre_v, f, n_iters, nu1, nu2, weight) (WaveletPotentials/cuMultigridFrac3D.cu:
 * 249-327): n_iters V-cycles with nu1 / nu2 weighted-Jacobi sweeps before /
 * after the coarse correction, 50 ('n') or 20 sweeps on the coarsest level
 * (in FP64), starting from pre_v (NULL: zero).  dst, pre_v, f are (nx, ny, nz) cell arrays;
 * dst may equal pre_v. */
WHHD_API int whhd_mgfrac3d_solve(whhd_mgfrac3d_t h, float* dst, const float* pre_v, const float* f,
                                 int n_iters, int nu1, int nu2, float weight, void* stream);
/* Solve to a relative residual (no reference counterpart: the reference calls
 * Solve for a fixed cycle count, WaveletPotentials/cuMultigridFrac3D.cu:249-265,
 * in its build precision).  Mixed-precision refinement: the iterate is the
 * unevaluated sum dst + dst_lo of two FP32 arrays (an FP32 iterate alone has a
 * relative-residual floor near 1e-6 on fine grids).  Each cycle forms
 * r = f - A (dst + dst_lo) in FP64 with its FP64 norm, stops when
 * ||r|| <= tol ||f|| or after max_cycles V-cycles, otherwise runs 1-4
 * V-cycles (as Solve; as many as the observed convergence rate says reach tol,
 * two before a rate is known, never past max_cycles) on A e = r from e = 0 and
 * adds e to the pair; those V-cycles read the finest level's fractions rounded
 * to FP16 (the FP64 residual, which decides convergence, uses the FP32 ones).
 * pre_v (NULL: zero; may equal dst) starts dst; dst_lo starts at zero.
 * Writes the V-cycles run and the last relative residual; returns
 * WHHD_NONCONVERGED if tol was not reached.  Synchronizes `stream` once per
 * FP64 residual (16-byte read-back). */
WHHD_API int whhd_mgfrac3d_solve_tol(whhd_mgfrac3d_t h, float* dst, float* dst_lo,
                                     const float* pre_v, const float* f, int max_cycles,
                                     double tol, int nu1, int nu2, float weight, int* cycles,
                                     double* rel, void* stream);
/* Solve to a relative residual by conjugate gradients preconditioned by one
 * V-cycle (no reference counterpart; the V-cycle of Solve, nu1 / nu2
 * sweeps, on the residual from zero, reading the finest level's fractions
 * rounded to FP16 and running 10 coarsest-level sweeps instead of 50 for
 * bc 'n').  The iterate is the pair dst + dst_lo as in
 * whhd_mgfrac3d_solve_tol.  r = f - A (dst + dst_lo) is formed in FP64 at the
 * start; each iteration applies the V-cycle z = M r, builds the direction
 * p = z + beta p and A p (FP32 fractions, FP64 sums), and in one pass adds
 * alpha p to the pair and updates r <- r - alpha A p (beta, alpha stay on the
 * device).  Once that recurrence reaches max(1e-5, 10 tol) relative, every
 * further iteration forms the FP64 residual of the pair afresh, and that
 * decides convergence.  Needs fewer V-cycles than the refinement for the same
 * tolerance (a symmetric V-cycle, nu1 = nu2, is what conjugate gradients
 * assume).  pre_v (NULL: zero; may equal dst) starts dst.  Writes the
 * iterations (one V-cycle each) and the last FP64 relative residual; returns
 * WHHD_NONCONVERGED if tol was not reached.  Waits once per iteration for a
 * read-back of 8 or 16 bytes.  tol < 0: exactly max_iters iterations (the
 * FP64 residual formed every iteration) and nothing read
 * back or waited for (iters = max_iters, rel = -1), so the call can be
 * captured in a HIP graph. */
WHHD_API int whhd_mgfrac3d_solve_pcg(whhd_mgfrac3d_t h, float* dst, float* dst_lo,
                                     const float* pre_v, const float* f, int max_iters,
                                     double tol, int nu1, int nu2, float weight, int* iters,
                                     double* rel, void* stream);
/* ||f - A v||_2 / ||f||_2 with the sums in FP64 (CuWeightedJacobi3D::GetFracRhs,
 * WaveletPotentials/cuWeightedJacobi3D.cu:1384-1407).  Synchronizes `stream`. */
WHHD_API int whhd_mgfrac3d_residual(whhd_mgfrac3d_t h, const float* v, const float* f, double* rel,
                                    void* stream);
WHHD_API void whhd_mgfrac3d_destroy(whhd_mgfrac3d_t h);

/* ---- learned initial guess ---------------------------------------------- */
typedef struct whhd_unet_s* whhd_unet_t;
/* 3-level 3D U-Net (16/32/64 channels, 3x3x3 convolutions on the BF16 matrix
 * cores with FP32 accumulation, ReLU, 2x2x2 average pooling, nearest
 * upsampling, concatenated skips) predicting the fractional multigrid's
 * solution from (f, solid level set); its output is the pre_v of
 * whhd_mgfrac3d_solve / _solve_tol.  No reference counterpart: the reference's
 * multigrid takes any initial guess as pre_v (WaveletPotentials/
 * cuMultigridFrac3D.cu:249-258).  nx, ny, nz are the solve's cells; the
 * network runs on the grid coarser by `factor` (inputs averaged over factor^3
 * blocks, the prediction interpolated back trilinearly), so the extents must
 * be multiples of 4 * factor.  Weights are He-normal from `seed`. */
WHHD_API int whhd_unet_create(int nx, int ny, int nz, int factor, unsigned long long seed,
                              whhd_unet_t* out);
WHHD_API long long whhd_unet_num_params(whhd_unet_t h);
/* Copy the FP32 master weights (which = 0) or the last gradients (1) to host dst. */
WHHD_API int whhd_unet_get_params(whhd_unet_t h, float* dst, int which);
/* Load FP32 master weights from host memory; resets the Adam moments. */
WHHD_API int whhd_unet_set_params(whhd_unet_t h, const float* src);
/* p0 = prediction for the cell arrays f (right-hand side) and ls (level set),
 * all on the solve's grid; `length` is the domain's x extent (inputs are
 * scaled by rms(f) and length). */
WHHD_API int whhd_unet_predict(whhd_unet_t h, float* p0, const float* f, const float* ls,
                               double length, void* stream);
/* Forward and backward for the target p: loss = mean((e - mean e)^2) with
 * e = prediction - p in the network's scaled units; gradients kept in the
 * handle.  Writes the loss when `loss` is not NULL (synchronizes `stream`). */
WHHD_API int whhd_unet_loss_grad(whhd_unet_t h, const float* f, const float* ls, const float* p,
                                 double length, double* loss, void* stream);
/* Units of the network output: p = output * rms(f) * length^2 * scale
 * (default 1; training sets it from its data so the output is O(1)). */
WHHD_API int whhd_unet_set_output_scale(whhd_unet_t h, double scale);
WHHD_API double whhd_unet_output_scale(whhd_unet_t h);
/* One Adam step (beta 0.9 / 0.999) with the kept gradients. */
WHHD_API int whhd_unet_adam(whhd_unet_t h, float lr, void* stream);
WHHD_API void whhd_unet_destroy(whhd_unet_t h);

/* ---- field operators and reductions ------------------------------------ */
/* CuGradient::Gradient3D_Frac_P(xv, yv, zv, p, ax, ay, az, dim, dx, dy, dz, dx_e, bc)
 * (WaveletPotentials/cuGradient.cu:1841-1951, 2514-2545) for uniform spacing:
 * face gradients of the cell field p, zero on faces with fraction <= 1e-12.
 * bc 'n': wall faces 0; 'd': zero wall value on the face (distance h/2); 'z':
 * zero value one cell outside (distance h).  The far-wall face tests its own
 * fraction (the reference's spacing-array 'z' kernel tests the neighbouring
 * face's, cuGradient.cu:1939-1950). */
WHHD_API int whhd_gradient3d_frac(float* xv, float* yv, float* zv, const float* p, const float* ax,
                                  const float* ay, const float* az, int nx, int ny, int nz,
                                  double dx, double dy, double dz, char bc, void* stream);
/* u <- a * (u - grad p) on every face in one pass: the gradient above followed
 * by CudaMatAdd(u, grad, -1) and CudaMatMul(u, u, a)
 * (WaveletPotentials/TestLocalProjection.cpp:690-701).  No reference counterpart
 * as a single call. */
WHHD_API int whhd_subtract_gradient_frac(float* xv, float* yv, float* zv, const float* p,
                                         const float* ax, const float* ay, const float* az, int nx,
                                         int ny, int nz, double dx, double dy, double dz, char bc,
                                         void* stream);
/* The same step for a potential held as the sum of two floats (the
 * multigrid's p_hi + p_lo from whhd_mgfrac3d_solve_tol): u <- a (u - grad p_hi
 * - grad p_lo), the two gradients taken separately, in one pass
 * (WaveletPotentials/TestLocalProjection.cpp:690-701 with p = p_hi + p_lo).
 * No reference counterpart as a single call. */
WHHD_API int whhd_subtract_gradient_frac2(float* xv, float* yv, float* zv, const float* p_hi,
                                          const float* p_lo, const float* ax, const float* ay,
                                          const float* az, int nx, int ny, int nz, double dx,
                                          double dy, double dz, char bc, void* stream);
/* CuConvergence::GetB_3d_frac_vel(b, xv, yv, zv, ax, ay, az, dim, dx, bc)
 * (WaveletPotentials/cuConvergence.cu:40-124, 435-447): b = -div(a u) per cell,
 * over all six faces (both of the reference's kernels compute this value). */
WHHD_API int whhd_divergence3d_frac(float* b, const float* xv, const float* yv, const float* zv,
                                    const float* ax, const float* ay, const float* az, int nx,
                                    int ny, int nz, double dx, double dy, double dz, void* stream);
/* CudaMatAdd(dst, src, n, scale): dst += scale * src; CudaMatMul(dst, l, r, n):
 * dst = l * r (WaveletPotentials/cudaMath.cu:110-168). */
WHHD_API int whhd_axpy(float* dst, const float* src, long long n, float scale, void* stream);
WHHD_API int whhd_mul(float* dst, const float* l, const float* r, long long n, void* stream);
/* Reductions of CudaArraySum / CudaArrayNormSum / CudaArrayAbsSum / CudaFindMaxValue
 * (WaveletPotentials/cudaMath.cu:35-108, 1230-1280), accumulated in FP64 on the
 * device: `scratch` is caller-owned device memory of whhd_reduce_scratch_doubles()
 * doubles; the result lands in the host double *out.  Synchronizes `stream`. */
enum { WHHD_REDUCE_SUM = 0, WHHD_REDUCE_SUM_SQ = 1, WHHD_REDUCE_SUM_ABS = 2, WHHD_REDUCE_MAX_ABS = 3 };
WHHD_API int whhd_reduce(const float* x, long long n, int op, double* scratch, double* out,
                         void* stream);
WHHD_API int whhd_reduce_scratch_doubles(void);
/* whhd_reduce without the workspace argument: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under stream capture,
 * serving loops, harnesses that forbid hipMalloc).  It allocates nothing: the
 * partials live in a static device array of whhd_reduce_scratch_doubles()
 * doubles (8 KiB + 8 B per GPU, held until the process ends), which every n
 * fits.  Runs on the calling thread's current device; calls that overlap on
 * one GPU share that array and race.  Same kernels and results as whhd_reduce.
 * Synchronizes `stream`. */
WHHD_API int whhd_reduce_noalloc(const float* x, long long n, int op, double* out, void* stream);

/* ---- sweep + DST comparator ---------------------------------------------- */
typedef struct whhd_dst3d_s* whhd_dst3d_t;
/* CuDst3D(int3 dim, DTYPE3 dx) (WaveletPotentials/cuDst3D.cu:183-252): a DST-I
 * Poisson solver on an (n1, n2, n3) node grid.  Holds the sine matrices of the
 * three axes and one scratch array. */
WHHD_API int whhd_dst3d_create(int n1, int n2, int n3, double dx, double dy, double dz,
                               whhd_dst3d_t* out);
/* CuDst3D::Solve(dst, f) (WaveletPotentials/cuDst3D.cu:265-341): phi with
 * (-L_h) phi = b for the 7-point Laplacian, zero one node outside the grid; the
 * eigenvalues use each axis's own spacing (the reference scales every axis by
 * dx.x*dx.y, the same for cubic cells).  phi may equal b. */
WHHD_API int whhd_dst3d_solve(whhd_dst3d_t h, float* phi, const float* b, void* stream);
WHHD_API void whhd_dst3d_destroy(whhd_dst3d_t h);
/* FP64 forms of the DST solver (the reference built with USE_DOUBLE): a handle
 * made by a *_f64 create is solved only by whhd_dst3d_solve_f64, and the reverse
 * (WHHD_INVALID_ARG otherwise). */
WHHD_API int whhd_dst3d_create_f64(int n1, int n2, int n3, double dx, double dy, double dz,
                                   whhd_dst3d_t* out);
WHHD_API int whhd_dst2d_create_f64(int n1, int n2, double dx, double dy, whhd_dst3d_t* out);
WHHD_API int whhd_dst3d_solve_f64(whhd_dst3d_t h, double* phi, const double* b, void* stream);
/* CuDst(int2 dim, DTYPE2 dx) (WaveletPotentials/cuDst.cu:4-36): the 2D solver on an
 * (n1, n2) node grid; solve with whhd_dst3d_solve (CuDst::Solve,
 * WaveletPotentials/cuDst.cu:207-311: (-L_h) phi = b for the 5-point Laplacian,
 * each axis with its own spacing), release with whhd_dst3d_destroy. */
WHHD_API int whhd_dst2d_create(int n1, int n2, double dx, double dy, whhd_dst3d_t* out);

typedef struct whhd_sweep3d_s* whhd_sweep3d_t;
/* CuParallelSweep3D(int3 dim, DTYPE3 dx, char bc) (WaveletPotentials/cuParallelSweep3D.cu:
 * 876-895) for nx*ny*nz cells, velocity bc 'n', 'd' or 'z' ('z' takes the 'd'
 * branches, as in the reference).  Owns a DST solver on the (nx-1, ny-1, nz-1)
 * interior nodes ('n') or on all (nx+1, ny+1, nz+1) nodes ('d'). */
WHHD_API int whhd_sweep3d_create(int nx, int ny, int nz, double dx, double dy, double dz, char bc,
                                 whhd_sweep3d_t* out);
/* Solve(qx, qy, qz, xv, yv, zv) (WaveletPotentials/cuParallelSweep3D.cu:926-965): qx
 * and qy integrate yv and -xv along z from their bottom planes, qz = 0.  bc 'n':
 * the bottom planes are zero, then the top planes of qx, qy are cleared and
 * qz's top plane takes the x-integral of the cleared qx; zv is not read.
 * bc 'd': qx's bottom plane is zero and qy's integrates zv along x
 * (SweepingQy_x); no top step. */
WHHD_API int whhd_sweep3d_solve(whhd_sweep3d_t h, float* qx, float* qy, float* qz,
                                const float* xv, const float* yv, const float* zv, void* stream);
/* Project(qx, qy, qz) (WaveletPotentials/cuParallelSweep3D.cu:1316-1326).  bc 'n':
 * ProjectZero (:1054-1064), q -= grad phi with (-L_h) phi = -div q at the
 * interior nodes, phi = 0 on the boundary nodes.  bc 'd': ProjectD (:1066-1079),
 * phi on every node from the wall-clamped divergence (cuGetB_q_d_3d), zero one
 * node outside, and q -= grad phi over every edge. */
WHHD_API int whhd_sweep3d_project(whhd_sweep3d_t h, float* qx, float* qy, float* qz, void* stream);
WHHD_API void whhd_sweep3d_destroy(whhd_sweep3d_t h);
/* CuParallelSweep3D(int2 dim, DTYPE2 dx, char bc) (WaveletPotentials/cuParallelSweep3D.cu:
 * 897-916) on nx*ny cells; release with whhd_sweep3d_destroy.  Only
 * whhd_sweep2d_solve applies to it. */
WHHD_API int whhd_sweep2d_create(int nx, int ny, double dx, double dy, char bc,
                                 whhd_sweep3d_t* out);
/* Solve(qz, xv, yv) (WaveletPotentials/cuParallelSweep3D.cu:982-995): qz (nx+1, ny+1)
 * integrates -yv along x from its x = 0 column: zero for bc 'n', the integral of
 * xv along y for 'd' and 'z' (from qz(0, 0) = 0, which the reference leaves as
 * the array held it). */
WHHD_API int whhd_sweep2d_solve(whhd_sweep3d_t h, float* qz, const float* xv, const float* yv,
                                void* stream);
/* FP64 forms of the sweep comparator (USE_DOUBLE): same semantics on double
 * arrays; a handle made by a *_f64 create takes only the *_f64 solves. */
WHHD_API int whhd_sweep3d_create_f64(int nx, int ny, int nz, double dx, double dy, double dz,
                                     char bc, whhd_sweep3d_t* out);
WHHD_API int whhd_sweep3d_solve_f64(whhd_sweep3d_t h, double* qx, double* qy, double* qz,
                                    const double* xv, const double* yv, const double* zv,
                                    void* stream);
WHHD_API int whhd_sweep3d_project_f64(whhd_sweep3d_t h, double* qx, double* qy, double* qz,
                                      void* stream);
WHHD_API int whhd_sweep2d_create_f64(int nx, int ny, double dx, double dy, char bc,
                                     whhd_sweep3d_t* out);
WHHD_API int whhd_sweep2d_solve_f64(whhd_sweep3d_t h, double* qz, const double* xv,
                                    const double* yv, void* stream);

/* ---- verification helpers on host arrays ------------------------------- */
/* SetSliceZero(val, dim, axis) (WaveletPotentials/StreamInitial.cpp:5-43): zero the
 * first and last slices of an (nx, ny, nz) array normal to axis 'x', 'y' or 'z'. */
WHHD_API int whhd_set_slice_zero(float* val, int nx, int ny, int nz, char axis);
/* StreamInitial::GradientStream 3D (WaveletPotentials/StreamInitial.cpp:87-143): face
 * velocities = discrete curl of the edge potential on nx*ny*nz cells. */
WHHD_API int whhd_gradient_stream3d(float* xv, float* yv, float* zv, const float* qx,
                                    const float* qy, const float* qz, int nx, int ny, int nz,
                                    double dx, double dy, double dz);
/* StreamInitial::GradientStream 2D (WaveletPotentials/StreamInitial.cpp:145-175):
 * xv = dqz/dy, yv = -dqz/dx. */
WHHD_API int whhd_gradient_stream2d(float* xv, float* yv, const float* qz, int nx, int ny,
                                    double dx, double dy);
/* CalculateErrorL2 (WaveletPotentials/Utils.cpp:184-207, 296-331) over `count`
 * (base, check, n) arrays, FP64 sums; eps 1e-12 in the denominator when count > 1. */
WHHD_API int whhd_error_l2(const float* const* base, const float* const* check,
                           const long long* n, int count, double* out);

/* ---- field I/O on host arrays ------------------------------------------ */
/* The reference's dumps (WaveletPotentials/Utils.cpp:812-990); arrays are host
 * memory, y fastest.  Text: one row of ny values per (x, z) in memory order,
 * `digits` after the point, fixed or scientific (WriteDataTxt, WriteDataETxt);
 * read back with whhd_read_txt (ReadsDataTxt). */
WHHD_API int whhd_write_txt(const char* path, const float* data, int nx, int ny, int nz,
                            int digits, int scientific);
WHHD_API int whhd_read_txt(const char* path, float* data, int nx, int ny, int nz);
/* "x,y,z,value" lines for |value| >= threshold (WriteDataTxt_sparse) and
 * "x,y,z,vx,vy,vz" for |v| >= threshold (WriteVectorTxt_sparse). */
WHHD_API int whhd_write_txt_sparse(const char* path, const float* data, int nx, int ny, int nz,
                                   double threshold);
WHHD_API int whhd_write_vector_txt_sparse(const char* path, const float* vx, const float* vy,
                                          const float* vz, int nx, int ny, int nz,
                                          double threshold);
/* Raw FP32 values (WriteDataBinary) and their reader. */
WHHD_API int whhd_write_binary(const char* path, const float* data, long long n);
WHHD_API int whhd_read_binary(const char* path, float* data, long long n);
/* PLY point cloud (no reference counterpart; the reference includes a PLY
 * library, WaveletPotentials/happly.h, but writes none): a vertex with
 * property "value" at each cell centre where |value| >= threshold; ASCII or
 * binary little endian; the vertex count to *written (may be NULL). */
WHHD_API int whhd_write_ply(const char* path, const float* data, int nx, int ny, int nz, double dx,
                            double dy, double dz, double threshold, int binary,
                            long long* written);

#ifdef __cplusplus
}
#endif
#endif /* WHHD_H */
