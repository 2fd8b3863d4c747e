/*
 * ORACLE -- test infrastructure only (see whhd_oracle.h).
 *
 * FP64 building blocks of the learned initial guess's 3D U-Net (3x3x3
 * convolutions with zero padding, channels-last activations): the forward
 * convolution, its input gradient and its weight/bias gradients.  The network
 * has no reference counterpart; these are the textbook definitions
 *     y[v][co] = b[co] + sum_tap sum_ci x[v + off(tap)][ci] w[co][tap][ci],
 * off(tap) = (tap / 9 - 1, (tap / 3) % 3 - 1, tap % 3 - 1) in (z, x, y),
 * and their adjoints; oracle.py composes them into the network and the loss,
 * and its finite-difference test pins the composition.
 *
 * Voxels are numbered ((z * nx) + x) * ny + y.  The convolution input x is
 * the concatenation of `a` (ca channels; at half resolution, read by nearest
 * upsampling, when up_a) and `b` (cb channels, full resolution).
 */
#include <stddef.h>
#include <string.h>

#include "whhd_oracle.h"

static double fused(const double* a, int ca, const double* b, int cb, int up_a, int nz, int nx,
                    int ny, int z, int x, int y, int ci) {
    if (z < 0 || z >= nz || x < 0 || x >= nx || y < 0 || y >= ny) return 0.0;
    if (ci < ca) {
        if (up_a) {
            const size_t v = ((size_t)(z / 2) * (nx / 2) + x / 2) * (ny / 2) + y / 2;
            return a[v * ca + ci];
        }
        return a[(((size_t)z * nx + x) * ny + y) * ca + ci];
    }
    return b[(((size_t)z * nx + x) * ny + y) * cb + (ci - ca)];
}

int orc_conv3(double* y, const double* a, int ca, const double* b, int cb, int up_a, int nz,
              int nx, int ny, const double* w, const double* bias, int co) {
    if (!y || !a || (cb > 0 && !b) || !w || ca < 1 || cb < 0 || co < 1 || nz < 1 || nx < 1 || ny < 1)
        return ORC_INVALID;
    if (up_a && (nz % 2 || nx % 2 || ny % 2)) return ORC_INVALID;
    const int cin = ca + cb;
    const long V = (long)nz * nx * ny;
#pragma omp parallel for schedule(static)
    for (long v = 0; v < V; ++v) {
        const int yy = (int)(v % ny), xx = (int)((v / ny) % nx), zz = (int)(v / ((long)ny * nx));
        for (int o = 0; o < co; ++o) {
            double s = bias ? bias[o] : 0.0;
            for (int t = 0; t < 27; ++t) {
                const int dz = t / 9 - 1, dx = (t / 3) % 3 - 1, dy = t % 3 - 1;
                for (int ci = 0; ci < cin; ++ci)
                    s += fused(a, ca, b, cb, up_a, nz, nx, ny, zz + dz, xx + dx, yy + dy, ci) *
                         w[((size_t)o * 27 + t) * cin + ci];
            }
            y[v * co + o] = s;
        }
    }
    return ORC_OK;
}

int orc_conv3_bwd_data(double* dx, const double* g, int co, int nz, int nx, int ny,
                       const double* w, int cin) {
    if (!dx || !g || !w || co < 1 || cin < 1 || nz < 1 || nx < 1 || ny < 1) return ORC_INVALID;
    const long V = (long)nz * nx * ny;
#pragma omp parallel for schedule(static)
    for (long u = 0; u < V; ++u) {
        const int yy = (int)(u % ny), xx = (int)((u / ny) % nx), zz = (int)(u / ((long)ny * nx));
        for (int ci = 0; ci < cin; ++ci) {
            double s = 0.0;
            for (int t = 0; t < 27; ++t) {
                /* the output voxel v with v + off(t) = u */
                const int z = zz - (t / 9 - 1), x = xx - ((t / 3) % 3 - 1), y = yy - (t % 3 - 1);
                if (z < 0 || z >= nz || x < 0 || x >= nx || y < 0 || y >= ny) continue;
                const size_t v = ((size_t)z * nx + x) * ny + y;
                for (int o = 0; o < co; ++o) s += g[v * co + o] * w[((size_t)o * 27 + t) * cin + ci];
            }
            dx[u * cin + ci] = s;
        }
    }
    return ORC_OK;
}

int orc_conv3_bwd_weight(double* dw, double* db, const double* g, int co, const double* a, int ca,
                         const double* b, int cb, int up_a, int nz, int nx, int ny) {
    if (!dw || !g || !a || (cb > 0 && !b) || co < 1 || ca < 1 || cb < 0 || nz < 1 || nx < 1 || ny < 1)
        return ORC_INVALID;
    const int cin = ca + cb;
    const long V = (long)nz * nx * ny;
#pragma omp parallel for schedule(static) collapse(2)
    for (int o = 0; o < co; ++o)
        for (int t = 0; t < 27; ++t) {
            const int dz = t / 9 - 1, dx = (t / 3) % 3 - 1, dy = t % 3 - 1;
            for (int ci = 0; ci < cin; ++ci) {
                double s = 0.0;
                for (long v = 0; v < V; ++v) {
                    const int yy = (int)(v % ny), xx = (int)((v / ny) % nx),
                              zz = (int)(v / ((long)ny * nx));
                    s += g[v * co + o] *
                         fused(a, ca, b, cb, up_a, nz, nx, ny, zz + dz, xx + dx, yy + dy, ci);
                }
                dw[((size_t)o * 27 + t) * cin + ci] = s;
            }
        }
    if (db)
        for (int o = 0; o < co; ++o) {
            double s = 0.0;
            for (long v = 0; v < V; ++v) s += g[v * co + o];
            db[o] = s;
        }
    return ORC_OK;
}
