/*
 * ORACLE -- test infrastructure only (see whhd_oracle.h).
 *
 * FP64 restatement of the reference's solid-geometry kernels:
 *   sphere level set   WaveletPotentials/cuSolidLevelSet3D.cu:11-26
 *   face fractions     WaveletPotentials/cuSolidLevelSet3D.cu:28-90 (cuComputeFrac, cuGetAx/Ay/Az)
 *   supersampled face fractions
 *                      WaveletPotentials/cuSolidLevelSet3D.cu:164-317, 359-372
 *                      (cuGetAx/Ay/Az_divide, the static GetFrac), with the
 *                      samplers of WaveletPotentials/Interpolation.cuh:16-29,
 *                      80-157 (cuGetIxFx, InterpolateLinear3D,
 *                      InterpolateQuadratic3D)
 *   eikonal relaxation WaveletPotentials/cuFastSweeping.cu:105-146, 215-257
 * The reference's sweep updates its array in place from many threads at
 * once; this restatement relaxes all cells from the previous iterate (Jacobi),
 * which is the deterministic form of the same update.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "whhd_oracle.h"

#define IDX(y, x, z, NX, NY) ((size_t)(y) + (size_t)(x) * (NY) + (size_t)(z) * (NX) * (NY))

int orc_levelset_spheres(double* ls, int nx, int ny, int nz, double dx, double dy, double dz,
                         const double* centers, const double* radii, int count) {
    if (!ls || !centers || !radii || count < 1 || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                /* cuSolidLevelSet3D.cu:21-23: cell centre, distance minus radius */
                const double px = (x + 0.5) * dx, py = (y + 0.5) * dy, pz = (z + 0.5) * dz;
                double v = INFINITY;
                for (int i = 0; i < count; ++i) {
                    const double ex = px - centers[3 * i], ey = py - centers[3 * i + 1],
                                 ez = pz - centers[3 * i + 2];
                    const double d = sqrt(ex * ex + ey * ey + ez * ez) - radii[i];
                    if (d < v) v = d;
                }
                ls[IDX(y, x, z, nx, ny)] = v;
            }
    return ORC_OK;
}

/* cuComputeFrac, cuSolidLevelSet3D.cu:28-34 */
static double frac(double h, double l, double r) {
    const double dif = r - l;
    double d2 = h * h - dif * dif;
    if (d2 < 1e-14) d2 = 1e-14;
    double t = (1.0 - (r + l) / sqrt(d2)) * 0.5;
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
    return 1.0 - t;
}

int orc_frac_from_ls(double* ax, double* ay, double* az, const double* ls, int nx, int ny, int nz,
                     double dx, double dy, double dz) {
    if (!ax || !ay || !az || !ls || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    /* cuGetAx, cuSolidLevelSet3D.cu:36-52: ax has extents (nx+1, ny, nz) */
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x <= nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const int r = x < nx - 1 ? x : nx - 1, l = x - 1 > 0 ? x - 1 : 0;
                ax[IDX(y, x, z, nx + 1, ny)] =
                    frac(dx, ls[IDX(y, l, z, nx, ny)], ls[IDX(y, r, z, nx, ny)]);
            }
    /* cuGetAy, :54-71: ay has extents (nx, ny+1, nz) */
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y <= ny; ++y) {
                const int r = y < ny - 1 ? y : ny - 1, l = y - 1 > 0 ? y - 1 : 0;
                ay[IDX(y, x, z, nx, ny + 1)] =
                    frac(dy, ls[IDX(l, x, z, nx, ny)], ls[IDX(r, x, z, nx, ny)]);
            }
    /* cuGetAz, :73-90: az has extents (nx, ny, nz+1) */
    for (int z = 0; z <= nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const int r = z < nz - 1 ? z : nz - 1, l = z - 1 > 0 ? z - 1 : 0;
                az[IDX(y, x, z, nx, ny)] =
                    frac(dz, ls[IDX(y, x, l, nx, ny)], ls[IDX(y, x, r, nx, ny)]);
            }
    return ORC_OK;
}

/* solve_eikonal_nomin, cuFastSweeping.cu:105-146, with per-axis spacings */
double orc_eikonal_update(double mx, double my, double mz, double dx, double dy, double dz) {
    double m[3] = {mx, my, mz}, d[3] = {dx, dy, dz};
    for (int i = 1; i < 3; ++i)
        for (int j = 0; j < 3 - i; ++j)
            if (m[j] > m[j + 1]) {
                double t = m[j]; m[j] = m[j + 1]; m[j + 1] = t;
                t = d[j]; d[j] = d[j + 1]; d[j + 1] = t;
            }
    const double m0 = m[0], m1 = m[1], m2 = m[2], d0 = d[0], d1 = d[1], d2 = d[2];
    const double q0 = m0 * m0, q1 = m1 * m1, q2 = m2 * m2;
    const double e0 = d0 * d0, e1 = d1 * d1, e2 = d2 * d2;
    double u = m0 + d0;
    if (u > m1) {
        const double s = sqrt(-q0 + 2 * m0 * m1 - q1 + e0 + e1);
        u = (m1 * e0 + m0 * e1 + d0 * d1 * s) / (e0 + e1);
        if (u > m2) {
            const double a = sqrt(-q0 * e1 - q0 * e2 + 2 * m0 * m1 * e2 - q1 * e0 - q1 * e2 +
                                  2 * m0 * m2 * e1 - q2 * e0 - q2 * e1 + 2 * m1 * m2 * e0 +
                                  e0 * e1 + e0 * e2 + e1 * e2);
            u = (m2 * e0 * e1 + m1 * e0 * e2 + m0 * e1 * e2 + d0 * d1 * d2 * a) /
                (e0 * e1 + e0 * e2 + e1 * e2);
        }
    }
    return u;
}

int orc_eikonal_sweep(double* ls, int nx, int ny, int nz, double h, int n_iters) {
    if (!ls || nx < 1 || ny < 1 || nz < 1 || n_iters < 0) return ORC_INVALID;
    const size_t n = (size_t)nx * ny * nz;
    double* prev = (double*)malloc(n * sizeof(double));
    if (!prev) return ORC_INVALID;
    const double far = 1e8; /* cuFastSweeping.cu:232 */
    for (int it = 0; it < n_iters; ++it) {
        memcpy(prev, ls, n * sizeof(double));
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x)
                for (int y = 0; y < ny; ++y) {
                    const size_t i = IDX(y, x, z, nx, ny);
                    if (!(prev[i] > -1e-12)) continue; /* cuFastSweeping.cu:218 */
                    const double l = x > 0 ? prev[i - ny] : far;
                    const double r = x < nx - 1 ? prev[i + ny] : far;
                    const double u = y < ny - 1 ? prev[i + 1] : far;
                    const double d = y > 0 ? prev[i - 1] : far;
                    const double f = z > 0 ? prev[i - (size_t)nx * ny] : far;
                    const double b = z < nz - 1 ? prev[i + (size_t)nx * ny] : far;
                    ls[i] = orc_eikonal_update(l < r ? l : r, u < d ? u : d, f < b ? f : b, h, h, h);
                }
    }
    free(prev);
    return ORC_OK;
}

/* ---- supersampled face fractions ------------------------------------------
 * Grid coordinates are cell-centred: coordinate g of a point at world position
 * p along an axis of spacing h is p / h - 1/2 (Interpolation callers,
 * cuSolidLevelSet3D.cu:177, 189 with c_min = 0, dx_c = dx).  CLAMP_L is the
 * reference's FP32-build value 1e-5 (Interpolation.cuh:11-13); the library
 * computes in FP32. */
#define ORC_CLAMP_L 1e-5

/* cuGetIxFx, Interpolation.cuh:16-29 */
static void ixfx(int* ix, double* fx, double x, int n) {
    if (x < ORC_CLAMP_L) x = ORC_CLAMP_L;
    *ix = (int)x;
    if (*ix >= n - 1) {
        *ix = n - 2;
        *fx = 1.0 - ORC_CLAMP_L;
    } else {
        *fx = x - *ix;
    }
}

/* InterpolateLinear3D, Interpolation.cuh:80-117 (same summation order) */
double orc_interp_linear3d(const double* v, double x, double y, double z, int nx, int ny, int nz) {
    int ix, iy, iz;
    double fx, fy, fz;
    ixfx(&ix, &fx, x, nx);
    ixfx(&iy, &fy, y, ny);
    ixfx(&iz, &fz, z, nz);
    const double wx[2] = {1 - fx, fx}, wy[2] = {1 - fy, fy}, wz[2] = {1 - fz, fz};
    const size_t slice = (size_t)nx * ny;
    size_t i = (size_t)iy + (size_t)ix * ny + (size_t)iz * slice;
    double r = v[i] * wx[0] * wy[0] * wz[0];
    r += v[i + 1] * wx[0] * wy[1] * wz[0];
    r += v[i + ny] * wx[1] * wy[0] * wz[0];
    r += v[i + ny + 1] * wx[1] * wy[1] * wz[0];
    i += slice;
    r += v[i] * wx[0] * wy[0] * wz[1];
    r += v[i + 1] * wx[0] * wy[1] * wz[1];
    r += v[i + ny] * wx[1] * wy[0] * wz[1];
    r += v[i + ny + 1] * wx[1] * wy[1] * wz[1];
    return r;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* InterpolateQuadratic3D, Interpolation.cuh:119-157: quadratic B-spline
 * weights on the three nodes from floor(g - 1/2), indices clamped to the grid */
double orc_interp_quadratic3d(const double* v, double x, double y, double z, int nx, int ny,
                              int nz) {
    const int ix = (int)floor(x - 0.5), iy = (int)floor(y - 0.5), iz = (int)floor(z - 0.5);
    const double fx = x - ix, fy = y - iy, fz = z - iz;
#define SQ(a) ((a) * (a))
    const double wx[3] = {0.5 * SQ(1.5 - fx), 0.75 - SQ(fx - 1.0), 0.5 * SQ(fx - 0.5)};
    const double wy[3] = {0.5 * SQ(1.5 - fy), 0.75 - SQ(fy - 1.0), 0.5 * SQ(fy - 0.5)};
    const double wz[3] = {0.5 * SQ(1.5 - fz), 0.75 - SQ(fz - 1.0), 0.5 * SQ(fz - 0.5)};
#undef SQ
    double r = 0.0;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                const int X = clampi(ix + j, 0, nx - 1), Y = clampi(iy + i, 0, ny - 1),
                          Z = clampi(iz + k, 0, nz - 1);
                r += wx[j] * wy[i] * wz[k] * v[IDX(Y, X, Z, nx, ny)];
            }
    return r;
}

/* One face of cuGetA{x,y,z}_divide, cuSolidLevelSet3D.cu:164-317.  The face is
 * normal to `axis`; g0[] holds the grid coordinates of its lower corner in the
 * two in-plane axes and of its plane along `axis`.  Quadratic band test at
 * the face centre against +-h (h: the normal axis's spacing), then the share
 * of 128 x 128 midpoint sub-samples where the linear interpolant is > 0. */
#define ORC_SUB 128
static double divide_face(const double* ls, int nx, int ny, int nz, int axis, const double g0[3],
                          double h) {
    double c[3] = {g0[0], g0[1], g0[2]};
    for (int a = 0; a < 3; ++a)
        if (a != axis) c[a] += 0.5;
    const double lc = orc_interp_quadratic3d(ls, c[0], c[1], c[2], nx, ny, nz);
    if (lc > h) return 1.0;
    if (!(lc > -h)) return 0.0;
    const int p = axis == 0 ? 1 : 0, q = axis == 2 ? 1 : 2; /* in-plane axes */
    long pos = 0;
    for (int k = 0; k < ORC_SUB; ++k)
        for (int i = 0; i < ORC_SUB; ++i) {
            double g[3] = {g0[0], g0[1], g0[2]};
            g[p] += (i + 0.5) / ORC_SUB;
            g[q] += (k + 0.5) / ORC_SUB;
            if (orc_interp_linear3d(ls, g[0], g[1], g[2], nx, ny, nz) > 0.0) ++pos;
        }
    double r = (double)pos / (ORC_SUB * ORC_SUB);
    return r < 0.0 ? 0.0 : (r > 1.0 ? 1.0 : r);
}

int orc_frac_divide(double* ax, double* ay, double* az, const double* ls, int nx, int ny, int nz,
                    double dx, double dy, double dz) {
    if (!ax || !ay || !az || !ls || nx < 2 || ny < 2 || nz < 2) return ORC_INVALID;
    /* face (x, y, z) of ax lies in the plane at world x = x*dx; its lower
     * corner has grid coordinates (x - 1/2, y - 1/2, z - 1/2) */
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x <= nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double g0[3] = {x - 0.5, y - 0.5, z - 0.5};
                ax[IDX(y, x, z, nx + 1, ny)] = divide_face(ls, nx, ny, nz, 0, g0, dx);
            }
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y <= ny; ++y) {
                const double g0[3] = {x - 0.5, y - 0.5, z - 0.5};
                ay[IDX(y, x, z, nx, ny + 1)] = divide_face(ls, nx, ny, nz, 1, g0, dy);
            }
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z <= nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double g0[3] = {x - 0.5, y - 0.5, z - 0.5};
                az[IDX(y, x, z, nx, ny)] = divide_face(ls, nx, ny, nz, 2, g0, dz);
            }
    return ORC_OK;
}
