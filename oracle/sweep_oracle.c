/*
 * ORACLE -- test infrastructure only (see whhd_oracle.h).
 *
 * FP64 restatement of the sweep + DST comparator (velocity bc 'n'):
 *   Solve        WaveletPotentials/cuParallelSweep3D.cu:9-66 (SweepingQx/Qy, cuSweepingQ_x),
 *                939-965 (Solve: qz = 0, qx/qy bottom planes 0, z sweeps, top gauge fix)
 *   ProjectZero  WaveletPotentials/cuParallelSweep3D.cu:1054-1064 with
 *                GetB_3d_q_zeros (WaveletPotentials/cuConvergence.cu:126-160, 449-455),
 *                CuDst3D::Solve (WaveletPotentials/cuDst3D.cu:265-341),
 *                GradientMinus3D_zero_Q (WaveletPotentials/cuGradient.cu:2353-2401)
 * The DST solve returns phi with (-L_h) phi = b for the 7-point Laplacian with
 * zero values one node outside the grid, per-axis spacing in the eigenvalues
 * (the reference scales all axes by dx.x*dx.y; equal for cubic cells).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "whhd_oracle.h"

#define I3(Y_, X_, Z_, NX, NY) ((size_t)(Y_) + (size_t)(X_) * (NY) + (size_t)(Z_) * (NX) * (NY))

int orc_sweep3d(double* qx, double* qy, double* qz, const double* xv, const double* yv,
                int nx, int ny, int nz, double dx, double dy, double dz) {
    if (!qx || !qy || !qz || !xv || !yv || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    (void)dy;
    /* qz = 0 (nx+1, ny+1, nz) */
    memset(qz, 0, sizeof(double) * (size_t)(nx + 1) * (ny + 1) * nz);
    /* qx (nx, ny+1, nz+1): bottom plane 0, then qx(z) = qx(z-1) + yv(z-1) dz (SweepingQx) */
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y <= ny; ++y) {
            qx[I3(y, x, 0, nx, ny + 1)] = 0.0;
            for (int z = 1; z <= nz; ++z)
                qx[I3(y, x, z, nx, ny + 1)] =
                    qx[I3(y, x, z - 1, nx, ny + 1)] + yv[I3(y, x, z - 1, nx, ny + 1)] * dz;
        }
    /* qy (nx+1, ny, nz+1): qy(z) = qy(z-1) - xv(z-1) dz (SweepingQy) */
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y < ny; ++y) {
            qy[I3(y, x, 0, nx + 1, ny)] = 0.0;
            for (int z = 1; z <= nz; ++z)
                qy[I3(y, x, z, nx + 1, ny)] =
                    qy[I3(y, x, z - 1, nx + 1, ny)] - xv[I3(y, x, z - 1, nx + 1, ny)] * dz;
        }
    /* bc 'n' top: qp(x, y) = sum_{x' < x} qx(x', y, top) dx (cuSweepingQ_x); top planes of
     * qx and qy become 0 and the top plane of qz takes -qp / dz */
    double* qp = (double*)calloc((size_t)(nx + 1) * (ny + 1), sizeof(double));
    if (!qp) return ORC_INVALID;
    for (int y = 0; y <= ny; ++y)
        for (int x = 1; x <= nx; ++x)
            qp[(size_t)y + (size_t)x * (ny + 1)] =
                qp[(size_t)y + (size_t)(x - 1) * (ny + 1)] + qx[I3(y, x - 1, nz, nx, ny + 1)] * dx;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y <= ny; ++y) qx[I3(y, x, nz, nx, ny + 1)] = 0.0;
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y < ny; ++y) qy[I3(y, x, nz, nx + 1, ny)] = 0.0;
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y <= ny; ++y)
            qz[I3(y, x, nz - 1, nx + 1, ny + 1)] -= qp[(size_t)y + (size_t)x * (ny + 1)] / dz;
    free(qp);
    return ORC_OK;
}

/* orthonormal DST-I along one axis of an (n1, n2, n3) y-fastest array, in place */
static void dst_axis(double* a, int n1, int n2, int n3, int axis, double* tmp) {
    const int n = axis == 0 ? n1 : axis == 1 ? n2 : n3;
    const double s = sqrt(2.0 / (n + 1));
    const size_t st = axis == 0 ? (size_t)n2 : axis == 1 ? 1 : (size_t)n1 * n2;
    for (int z = 0; z < (axis == 2 ? 1 : n3); ++z)
        for (int x = 0; x < (axis == 0 ? 1 : n1); ++x)
            for (int y = 0; y < (axis == 1 ? 1 : n2); ++y) {
                const size_t b = (size_t)y + (size_t)x * n2 + (size_t)z * n1 * n2;
                for (int k = 0; k < n; ++k) {
                    double acc = 0.0;
                    for (int j = 0; j < n; ++j) acc += a[b + j * st] * sin(M_PI * (k + 1) * (j + 1) / (n + 1));
                    tmp[k] = s * acc;
                }
                for (int k = 0; k < n; ++k) a[b + k * st] = tmp[k];
            }
}

int orc_dst3d_solve(double* phi, const double* b, int n1, int n2, int n3, double dx, double dy,
                    double dz) {
    if (!phi || !b || n1 < 1 || n2 < 1 || n3 < 1) return ORC_INVALID;
    const size_t N = (size_t)n1 * n2 * n3;
    int m = n1 > n2 ? n1 : n2;
    if (n3 > m) m = n3;
    double* tmp = (double*)malloc(sizeof(double) * m);
    if (!tmp) return ORC_INVALID;
    memmove(phi, b, N * sizeof(double));
    for (int a = 0; a < 3; ++a) dst_axis(phi, n1, n2, n3, a, tmp);
    for (int z = 0; z < n3; ++z)
        for (int x = 0; x < n1; ++x)
            for (int y = 0; y < n2; ++y) {
                const double lx = (2.0 - 2.0 * cos(M_PI * (x + 1) / (n1 + 1))) / (dx * dx);
                const double ly = (2.0 - 2.0 * cos(M_PI * (y + 1) / (n2 + 1))) / (dy * dy);
                const double lz = (2.0 - 2.0 * cos(M_PI * (z + 1) / (n3 + 1))) / (dz * dz);
                phi[I3(y, x, z, n1, n2)] /= (lx + ly + lz);
            }
    for (int a = 2; a >= 0; --a) dst_axis(phi, n1, n2, n3, a, tmp);
    free(tmp);
    return ORC_OK;
}

/* CuParallelSweep3D::Solve for velocity bc 'd' (and 'z', which takes the same
 * branches): WaveletPotentials/cuParallelSweep3D.cu:926-965.  qz = 0, the
 * bottom plane of qx is 0 and that of qy integrates zv along x
 * (SweepingQy_x, :68-85); then the z sweeps from those planes; no top gauge. */
int orc_sweep3d_d(double* qx, double* qy, double* qz, const double* xv, const double* yv,
                  const double* zv, int nx, int ny, int nz, double dx, double dy, double dz) {
    if (!qx || !qy || !qz || !xv || !yv || !zv || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    (void)dy;
    memset(qz, 0, sizeof(double) * (size_t)(nx + 1) * (ny + 1) * nz);
    for (int y = 0; y < ny; ++y) {
        qy[I3(y, 0, 0, nx + 1, ny)] = 0.0;
        for (int x = 1; x <= nx; ++x)
            qy[I3(y, x, 0, nx + 1, ny)] = qy[I3(y, x - 1, 0, nx + 1, ny)] + zv[I3(y, x - 1, 0, nx, ny)] * dx;
    }
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y <= ny; ++y) {
            qx[I3(y, x, 0, nx, ny + 1)] = 0.0;
            for (int z = 1; z <= nz; ++z)
                qx[I3(y, x, z, nx, ny + 1)] =
                    qx[I3(y, x, z - 1, nx, ny + 1)] + yv[I3(y, x, z - 1, nx, ny + 1)] * dz;
        }
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 1; z <= nz; ++z)
                qy[I3(y, x, z, nx + 1, ny)] =
                    qy[I3(y, x, z - 1, nx + 1, ny)] - xv[I3(y, x, z - 1, nx + 1, ny)] * dz;
    return ORC_OK;
}

/* CuParallelSweep3D::ProjectD (WaveletPotentials/cuParallelSweep3D.cu:1066-1079):
 * b = -div q at every node of the potential grid with the edge index clamped at
 * the walls (cuGetB_q_d_3d, cuConvergence.cu:161-200), a DST-I solve on those
 * (nx+1, ny+1, nz+1) nodes, and q -= grad phi over every edge
 * (cuGradientMinus3D_d_Q, cuGradient.cu:2408-2446). */
int orc_project_d(double* qx, double* qy, double* qz, int nx, int ny, int nz, double dx,
                  double dy, double dz) {
    if (!qx || !qy || !qz || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    const int px = nx + 1, py = ny + 1, pz = nz + 1;
    double* phi = (double*)malloc(sizeof(double) * (size_t)px * py * pz);
    if (!phi) return ORC_INVALID;
    for (int k = 0; k < pz; ++k)
        for (int j = 0; j < px; ++j)
            for (int i = 0; i < py; ++i) {
                const int jl = j > 0 ? j - 1 : 0, jr = j < nx - 1 ? j : nx - 1;
                const int ib = i > 0 ? i - 1 : 0, it = i < ny - 1 ? i : ny - 1;
                const int kf = k > 0 ? k - 1 : 0, kb = k < nz - 1 ? k : nz - 1;
                const double d = (qx[I3(i, jr, k, nx, ny + 1)] - qx[I3(i, jl, k, nx, ny + 1)]) / dx +
                                 (qy[I3(it, j, k, nx + 1, ny)] - qy[I3(ib, j, k, nx + 1, ny)]) / dy +
                                 (qz[I3(i, j, kb, nx + 1, ny + 1)] - qz[I3(i, j, kf, nx + 1, ny + 1)]) / dz;
                phi[I3(i, j, k, px, py)] = -d;
            }
    orc_dst3d_solve(phi, phi, px, py, pz, dx, dy, dz);
    for (int z = 0; z < pz; ++z)
        for (int x = 0; x < px; ++x)
            for (int y = 0; y < py; ++y) {
                const double p = phi[I3(y, x, z, px, py)];
                if (x < nx) qx[I3(y, x, z, nx, ny + 1)] -= (phi[I3(y, x + 1, z, px, py)] - p) / dx;
                if (y < ny) qy[I3(y, x, z, nx + 1, ny)] -= (phi[I3(y + 1, x, z, px, py)] - p) / dy;
                if (z < nz) qz[I3(y, x, z, nx + 1, ny + 1)] -= (phi[I3(y, x, z + 1, px, py)] - p) / dz;
            }
    free(phi);
    return ORC_OK;
}

/* CuParallelSweep3D::Solve(qz, xv, yv), the 2D sweep
 * (WaveletPotentials/cuParallelSweep3D.cu:982-995, kernels :87-142): qz (nx+1,
 * ny+1) integrates -yv along x from its x = 0 column, which is 0 for bc 'n'
 * (SweepingQz_x_0) and the integral of xv along y for 'd' / 'z' (SweepingQz_y,
 * starting from qz(0, 0) = 0). */
int orc_sweep2d(double* qz, const double* xv, const double* yv, int nx, int ny, double dx,
                double dy, char bc) {
    if (!qz || !xv || !yv || nx < 1 || ny < 1) return ORC_INVALID;
    qz[0] = 0.0;
    for (int y = 1; y <= ny; ++y) qz[y] = bc == 'n' ? 0.0 : qz[y - 1] + xv[y - 1] * dy;
    for (int y = 0; y <= ny; ++y)
        for (int x = 1; x <= nx; ++x)
            qz[y + (size_t)x * (ny + 1)] = qz[y + (size_t)(x - 1) * (ny + 1)] - yv[y + (size_t)(x - 1) * (ny + 1)] * dx;
    return ORC_OK;
}

/* CuDst::Solve (WaveletPotentials/cuDst.cu:207-311): the 2D DST-I Poisson solve,
 * (-L_h) phi = b for the 5-point Laplacian on an (n1, n2) node grid with zero
 * values one node outside; each axis uses its own spacing (the reference
 * scales both by dx.x*dx.y, the same for square cells). */
int orc_dst2d_solve(double* phi, const double* b, int n1, int n2, double dx, double dy) {
    if (!phi || !b || n1 < 1 || n2 < 1) return ORC_INVALID;
    const size_t N = (size_t)n1 * n2;
    double* tmp = (double*)malloc(sizeof(double) * (n1 > n2 ? n1 : n2));
    if (!tmp) return ORC_INVALID;
    memmove(phi, b, N * sizeof(double));
    dst_axis(phi, n1, n2, 1, 0, tmp);
    dst_axis(phi, n1, n2, 1, 1, tmp);
    for (int x = 0; x < n1; ++x)
        for (int y = 0; y < n2; ++y) {
            const double lx = (2.0 - 2.0 * cos(M_PI * (x + 1) / (n1 + 1))) / (dx * dx);
            const double ly = (2.0 - 2.0 * cos(M_PI * (y + 1) / (n2 + 1))) / (dy * dy);
            phi[y + (size_t)x * n2] /= (lx + ly);
        }
    dst_axis(phi, n1, n2, 1, 1, tmp);
    dst_axis(phi, n1, n2, 1, 0, tmp);
    free(tmp);
    return ORC_OK;
}

/* b = -div q at the interior nodes (1..nx-1, 1..ny-1, 1..nz-1) of the potential grid */
static void node_div(double* b, const double* qx, const double* qy, const double* qz, int nx,
                     int ny, int nz, double dx, double dy, double dz) {
    const int px = nx - 1, py = ny - 1, pz = nz - 1;
    for (int k = 0; k < pz; ++k)
        for (int j = 0; j < px; ++j)
            for (int i = 0; i < py; ++i) {
                /* node (x=j+1, y=i+1, z=k+1) */
                const double dqx = qx[I3(i + 1, j + 1, k + 1, nx, ny + 1)] - qx[I3(i + 1, j, k + 1, nx, ny + 1)];
                const double dqy = qy[I3(i + 1, j + 1, k + 1, nx + 1, ny)] - qy[I3(i, j + 1, k + 1, nx + 1, ny)];
                const double dqz = qz[I3(i + 1, j + 1, k + 1, nx + 1, ny + 1)] - qz[I3(i + 1, j + 1, k, nx + 1, ny + 1)];
                b[I3(i, j, k, px, py)] = -(dqx / dx + dqy / dy + dqz / dz);
            }
}

int orc_project_zero(double* qx, double* qy, double* qz, int nx, int ny, int nz, double dx,
                     double dy, double dz) {
    if (!qx || !qy || !qz || nx < 2 || ny < 2 || nz < 2) return ORC_INVALID;
    const int px = nx - 1, py = ny - 1, pz = nz - 1;
    const size_t N = (size_t)px * py * pz;
    double* b = (double*)malloc(N * sizeof(double));
    double* phi = (double*)malloc(N * sizeof(double));
    if (!b || !phi) {
        free(b);
        free(phi);
        return ORC_INVALID;
    }
    node_div(b, qx, qy, qz, nx, ny, nz, dx, dy, dz);
    orc_dst3d_solve(phi, b, px, py, pz, dx, dy, dz);
    /* q -= grad phi with phi = 0 on the boundary nodes (GradientMinus3D_zero_Q) */
#define PHI(X_, Y_, Z_) (((X_) < 1 || (X_) > px || (Y_) < 1 || (Y_) > py || (Z_) < 1 || (Z_) > pz) \
                             ? 0.0 : phi[I3((Y_) - 1, (X_) - 1, (Z_) - 1, px, py)])
    for (int z = 0; z <= nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y <= ny; ++y)
                qx[I3(y, x, z, nx, ny + 1)] -= (PHI(x + 1, y, z) - PHI(x, y, z)) / dx;
    for (int z = 0; z <= nz; ++z)
        for (int x = 0; x <= nx; ++x)
            for (int y = 0; y < ny; ++y)
                qy[I3(y, x, z, nx + 1, ny)] -= (PHI(x, y + 1, z) - PHI(x, y, z)) / dy;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x <= nx; ++x)
            for (int y = 0; y <= ny; ++y)
                qz[I3(y, x, z, nx + 1, ny + 1)] -= (PHI(x, y, z + 1) - PHI(x, y, z)) / dz;
#undef PHI
    free(b);
    free(phi);
    return ORC_OK;
}
