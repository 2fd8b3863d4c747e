/*
 * ORACLE -- test infrastructure only (see whhd_oracle.h).
 *
 * FP64 restatement of the face/cell operators around the pressure solve:
 *   gradient with fractions  WaveletPotentials/cuGradient.cu:1841-1951 (uniform spacing)
 *   divergence of a*u        WaveletPotentials/cuConvergence.cu:40-124
 * bc 'n': wall faces get no gradient; 'd': the zero wall value sits on the
 * face (exterior spacing 0); 'z': one cell outside (exterior spacing h).
 */
#include <stddef.h>

#include "whhd_oracle.h"

#define CELL(Y_, X_, Z_) ((size_t)(Y_) + (size_t)(X_) * ny + (size_t)(Z_) * nx * ny)
#define XF(Y_, X_, Z_) ((size_t)(Y_) + (size_t)(X_) * ny + (size_t)(Z_) * (nx + 1) * ny)
#define YF(Y_, X_, Z_) ((size_t)(Y_) + (size_t)(X_) * (ny + 1) + (size_t)(Z_) * nx * (ny + 1))
#define ZF(Y_, X_, Z_) ((size_t)(Y_) + (size_t)(X_) * ny + (size_t)(Z_) * nx * ny)

/* one face between cell lo (index k-1) and hi (index k) of an axis with c cells */
static double grad1(double a, int k, int c, double p_lo, double p_hi, double h, char bc) {
    if (!(a > 1e-12)) return 0.0;
    if (k > 0 && k < c) return (p_hi - p_lo) / (h + h) * 2.0;
    if (bc == 'n') return 0.0;
    const double he = bc == 'd' ? 0.0 : h; /* exterior spacing */
    return k == 0 ? (p_hi - 0.0) / (h + he) * 2.0 : -p_lo / (h + he) * 2.0;
}

int orc_gradient3d_frac(double* xv, double* yv, double* zv, const double* p, const double* ax,
                        const double* ay, const double* az, int nx, int ny, int nz, double dx,
                        double dy, double dz, char bc) {
    if (!xv || !yv || !zv || !p || !ax || !ay || !az || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    if (bc != 'n' && bc != 'd' && bc != 'z') return ORC_INVALID;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x <= nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double lo = x > 0 ? p[CELL(y, x - 1, z)] : 0.0;
                const double hi = x < nx ? p[CELL(y, x, z)] : 0.0;
                xv[XF(y, x, z)] = grad1(ax[XF(y, x, z)], x, nx, lo, hi, dx, bc);
            }
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y <= ny; ++y) {
                const double lo = y > 0 ? p[CELL(y - 1, x, z)] : 0.0;
                const double hi = y < ny ? p[CELL(y, x, z)] : 0.0;
                yv[YF(y, x, z)] = grad1(ay[YF(y, x, z)], y, ny, lo, hi, dy, bc);
            }
    for (int z = 0; z <= nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double lo = z > 0 ? p[CELL(y, x, z - 1)] : 0.0;
                const double hi = z < nz ? p[CELL(y, x, z)] : 0.0;
                zv[ZF(y, x, z)] = grad1(az[ZF(y, x, z)], z, nz, lo, hi, dz, bc);
            }
    return ORC_OK;
}

int orc_divergence3d_frac(double* b, const double* xv, const double* yv, const double* zv,
                          const double* ax, const double* ay, const double* az, int nx, int ny,
                          int nz, double dx, double dy, double dz) {
    if (!b || !xv || !yv || !zv || !ax || !ay || !az || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const double v =
                    (xv[XF(y, x + 1, z)] * ax[XF(y, x + 1, z)] - xv[XF(y, x, z)] * ax[XF(y, x, z)]) / dx +
                    (yv[YF(y + 1, x, z)] * ay[YF(y + 1, x, z)] - yv[YF(y, x, z)] * ay[YF(y, x, z)]) / dy +
                    (zv[ZF(y, x, z + 1)] * az[ZF(y, x, z + 1)] - zv[ZF(y, x, z)] * az[ZF(y, x, z)]) / dz;
                b[CELL(y, x, z)] = -v;
            }
    return ORC_OK;
}
