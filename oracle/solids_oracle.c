/*
 * ORACLE -- test infrastructure only (see whhd_oracle.h).
 *
 * FP64 restatement of the reference's solid-geometry kernels:
 *   sphere level set   WaveletPotentials/cuSolidLevelSet3D.cu:11-26
 *   face fractions     WaveletPotentials/cuSolidLevelSet3D.cu:28-90 (cuComputeFrac, cuGetAx/Ay/Az)
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
