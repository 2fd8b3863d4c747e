/*
 * ORACLE -- test infrastructure only.
 *
 * CPU FP64 restatement of the reference's wavelet potential-recovery path
 * (yours321dog/WaveletPotentials).  It is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load or call it.  Nothing in the shipped library depends on it.
 *
 * Layout convention (WaveletPotentials/Utils.h:46-47): a 3D array of extents
 * (nx, ny, nz) is stored y-fastest: index(i=y, j=x, k=z) = i + j*ny + k*nx*ny.
 *
 * Every function returns 0 on success, a negative code on invalid input, and
 * ORC_UNDEFINED when the reference's algorithm would read outside the line
 * for that (wavelet type, boundary condition, length) -- the reference gives
 * no defined result there.
 */
#ifndef WHHD_ORACLE_H
#define WHHD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_INVALID (-1)
#define ORC_UNSUPPORTED (-2)
#define ORC_UNDEFINED (-3)

/* WaveletType enumerators in the order of WaveletPotentials/WaveletTypes.h:7-28 */
enum orc_wavelet {
    ORC_CDF3_9 = 0, ORC_CDF3_11, ORC_CDF3_5, ORC_CDF2_10, ORC_CDF2_12, ORC_CDF2_6,
    ORC_CDF4_4, ORC_CDF4_6, ORC_CDF4_8, ORC_CDF4_10, ORC_CDF2_8, ORC_CDF3_7,
    ORC_CDF26_44, ORC_CDF28_46, ORC_CDF44_26, ORC_CDF46_28, ORC_CDF2_0, ORC_CDF1_1, ORC_CDF1_7
};

int orc_wavelet_supported(int wt, char bc);
int orc_wavelet_halo(int wt);
char orc_proj_bc(char bc);
int orc_proj_wavelet(int wt);

/* One multi-level transform of every line along `axis` ('x','y','z') of a
 * (nx,ny,nz) array.  dst may equal src.  levels = floor(log2(line length)). */
int orc_transform(double* dst, const double* src, int nx, int ny, int nz, int levels,
                  char axis, char bc, int wt, int forward);

/* Literal emulation of the reference's 32-lane y-axis kernels (window, halo,
 * shuffle and shared-memory ping-pong semantics), for a single line. */
int orc_emulate_line(double* dst, const double* src, int n, int levels, char bc, int wt,
                     int forward);

/* Gram table of the synthesis basis of (wt, bc) on n samples: 8 rows of n,
 * rows as WaveletPotentials/cuGetInnerProduct.cu:11-77 (row 5 is 0 where the
 * reference leaves it unwritten). */
int orc_gram_table(double* out8n, int wt, char bc, int n);

/* 3D per-cell projection: bc_div 'd' (velocity BC 'n') cuProjectLocal_q_3d_dq_ccc,
 * bc_div 'n' (velocity BC 'd') cuProjectLocal_q_3d_nq_ccc.  N = 2^L cells per axis
 * (possibly different per axis). ip_* are Gram diagonals. */
int orc_project3d(double* qx, double* qy, double* qz, double* xv_out, double* yv_out,
                  double* zv_out, const double* xv, const double* yv, const double* zv,
                  int nx, int ny, int nz, double dx, double dy, double dz, int wt_div,
                  char bc_div);

/* 2D node projection: bc_div 'd' cuProjectLocal_q_single, bc_div 'n'
 * cuProjectLocal_q_single_n. */
int orc_project2d(double* qz, const double* xv, const double* yv, int nx, int ny,
                  double dx, double dy, int wt_div, char bc_div);

/* CuWaveletPotentialRecover3D::Solve (6- and 9-argument forms; v_out may be NULL). */
int orc_recover3d(double* qx, double* qy, double* qz, double* xv_out, double* yv_out,
                  double* zv_out, const double* xv, const double* yv, const double* zv,
                  int nx, int ny, int nz, double dx, double dy, double dz, char bc,
                  int wt_curl, int wt_div);

/* GetDivL0Matrix: the 16-entry corner table for level counts (lx, ly). */
void orc_div_l0_matrix(double* dst, int lx, int ly);

/* CuWaveletPotentialRecover2D::Solve */
int orc_recover2d(double* qz, const double* xv, const double* yv, int nx, int ny,
                  double dx, double dy, char bc, int wt_curl, int wt_div);

/* CuWaveletPotentialRecover3D::CutCoefficients */
int orc_cut3d(double* qx_out, double* qy_out, double* qz_out, const double* qx,
              const double* qy, const double* qz, int nx, int ny, int nz, char bc,
              int wt_curl, int wt_div, int level);

/* Discrete curl of an edge potential to face velocities (StreamInitial.cpp:87-175). */
void orc_curl3d(double* xv, double* yv, double* zv, const double* qx, const double* qy,
                const double* qz, int nx, int ny, int nz, double dx, double dy, double dz);
void orc_curl2d(double* xv, double* yv, const double* qz, int nx, int ny, double dx,
                double dy);
/* Cell divergence of face velocities (nx*ny*nz cells). */
void orc_div3d(double* div, const double* xv, const double* yv, const double* zv, int nx,
               int ny, int nz, double dx, double dy, double dz);

/* ---- solids (solids_oracle.c) ---- */
/* cuInitLsSphere for a union of `count` spheres: min_i |p - c_i| - r_i at cell centres. */
int orc_levelset_spheres(double* ls, int nx, int ny, int nz, double dx, double dy, double dz,
                         const double* centers, const double* radii, int count);
/* cuGetAx/Ay/Az with cuComputeFrac: face fractions ax (nx+1,ny,nz), ay (nx,ny+1,nz), az (nx,ny,nz+1). */
int orc_frac_from_ls(double* ax, double* ay, double* az, const double* ls, int nx, int ny, int nz,
                     double dx, double dy, double dz);
/* solve_eikonal_nomin for neighbour minima along x, y, z and spacings dx, dy, dz. */
double orc_eikonal_update(double mx, double my, double mz, double dx, double dy, double dz);
/* n_iters Jacobi relaxations of fast_sweep_kernel_pos (in place on ls). */
int orc_eikonal_sweep(double* ls, int nx, int ny, int nz, double h, int n_iters);
/* cuGetAx/Ay/Az_divide (the static GetFrac): quadratic band test, then the
 * share of 128 x 128 sub-face samples of the linearly interpolated level set
 * that are > 0.  Extents as orc_frac_from_ls; needs nx, ny, nz >= 2. */
int orc_frac_divide(double* ax, double* ay, double* az, const double* ls, int nx, int ny, int nz,
                    double dx, double dy, double dz);
/* InterpolateLinear3D / InterpolateQuadratic3D at cell-centred grid
 * coordinates (x, y, z) of a (nx, ny, nz) y-fastest array. */
double orc_interp_linear3d(const double* v, double x, double y, double z, int nx, int ny, int nz);
double orc_interp_quadratic3d(const double* v, double x, double y, double z, int nx, int ny,
                              int nz);

/* ---- fractional multigrid (mg_oracle.c) ---- */
/* Level extents of CuMultigridFrac3D's hierarchy (3 ints per level); returns the count. */
int orc_mgfrac3d_levels(int nx, int ny, int nz, int* dims_out);
/* CuMultigridFrac3D(ax, ay, az, dim, dx, bc).Solve(dst, pre_v, f, n_iters, nu1, nu2, weight). */
int orc_mgfrac3d_solve(double* dst, const double* pre_v, const double* f, const double* ax,
                       const double* ay, const double* az, int nx, int ny, int nz, double dx,
                       double dy, double dz, char bc, int n_iters, int nu1, int nu2,
                       double weight);
/* One weighted Jacobi sweep dst <- src, or (residual != 0) dst = f - A src, uniform spacing. */
int orc_wj_frac3d(double* dst, const double* src, const double* f, const double* ax,
                  const double* ay, const double* az, int nx, int ny, int nz, double dx, double dy,
                  double dz, char bc, double weight, int residual);
/* ||f - A v||_2 / ||f||_2 (CuWeightedJacobi3D::GetFracRhs). */
double orc_mgfrac3d_rel_residual(const double* v, const double* f, const double* ax,
                                 const double* ay, const double* az, int nx, int ny, int nz,
                                 double dx, double dy, double dz, char bc);
/* The transfer operators of the V-cycle on their own: spacing-weighted cell
 * restriction (extents of the fine field and its per-axis spacings),
 * prolongation added into a fine field, face-fraction restriction (extents of
 * the fine face array and its normal axis 0/1/2), and the fractions and
 * spacings of one level of the hierarchy. */
int orc_mg_restrict(double* dst, const double* src, int nx, int ny, int nz, const double* hx,
                    const double* hy, const double* hz);
int orc_mg_prolong_add(double* fine, const double* coarse, int nx, int ny, int nz);
int orc_mg_restrict_faces(double* dst, const double* src, int fx, int fy, int fz, int normal);
int orc_mg_level(const double* ax, const double* ay, const double* az, int nx, int ny, int nz,
                 double dx, double dy, double dz, int level, double* ax_out, double* ay_out,
                 double* az_out, double* hx_out, double* hy_out, double* hz_out);
/* ---- field operators (fields_oracle.c) ---- */
/* Gradient3D_Frac_P for uniform spacing: face gradients xv (nx+1,ny,nz), yv, zv of p. */
int orc_gradient3d_frac(double* xv, double* yv, double* zv, const double* p, const double* ax,
                        const double* ay, const double* az, int nx, int ny, int nz, double dx,
                        double dy, double dz, char bc);
/* GetB_3d_frac_vel: b = -div(a u) per cell. */
int orc_divergence3d_frac(double* b, const double* xv, const double* yv, const double* zv,
                          const double* ax, const double* ay, const double* az, int nx, int ny,
                          int nz, double dx, double dy, double dz);
/* ---- sweep + DST comparator (sweep_oracle.c) ---- */
/* CuParallelSweep3D::Solve for velocity bc 'n': qx, qy by z sweeps, qz = 0, top gauge fix. */
int orc_sweep3d(double* qx, double* qy, double* qz, const double* xv, const double* yv,
                int nx, int ny, int nz, double dx, double dy, double dz);
/* (-L_h) phi = b on an (n1, n2, n3) node grid, zero one node outside (CuDst3D::Solve). */
int orc_dst3d_solve(double* phi, const double* b, int n1, int n2, int n3, double dx, double dy,
                    double dz);
/* CuDst::Solve: 2D DST-I Poisson solve on an (n1, n2) node grid, phi may equal b. */
int orc_dst2d_solve(double* phi, const double* b, int n1, int n2, double dx, double dy);
/* CuParallelSweep3D::Solve for velocity bc 'd' ('z' takes the same branches). */
int orc_sweep3d_d(double* qx, double* qy, double* qz, const double* xv, const double* yv,
                  const double* zv, int nx, int ny, int nz, double dx, double dy, double dz);
/* CuParallelSweep3D::Solve(qz, xv, yv): the 2D sweep, bc 'n' or 'd' ('z' as 'd'). */
int orc_sweep2d(double* qz, const double* xv, const double* yv, int nx, int ny, double dx,
                double dy, char bc);
/* CuParallelSweep3D::ProjectD: the projection on every node of the potential grid. */
int orc_project_d(double* qx, double* qy, double* qz, int nx, int ny, int nz, double dx,
                  double dy, double dz);
/* CuParallelSweep3D::ProjectZero: q -= grad phi with (-L_h) phi = -div q at interior nodes. */
int orc_project_zero(double* qx, double* qy, double* qz, int nx, int ny, int nz, double dx,
                     double dy, double dz);
/* ---- U-Net convolutions (unet_oracle.c) ---- */
/* 3x3x3 convolution, zero padding, channels-last; input = concat(a (ca ch, half
 * resolution when up_a), b (cb ch)); w[co][27][ca+cb]; bias may be NULL. */
int orc_conv3(double* y, const double* a, int ca, const double* b, int cb, int up_a, int nz,
              int nx, int ny, const double* w, const double* bias, int co);
/* Input gradient at full resolution (cin = ca + cb channels) from g[V][co]. */
int orc_conv3_bwd_data(double* dx, const double* g, int co, int nz, int nx, int ny,
                       const double* w, int cin);
/* Weight and bias gradients (db may be NULL). */
int orc_conv3_bwd_weight(double* dw, double* db, const double* g, int co, const double* a, int ca,
                         const double* b, int cb, int up_a, int nz, int nx, int ny);
#ifdef __cplusplus
}
#endif
#endif
