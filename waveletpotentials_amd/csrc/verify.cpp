// Host verification helpers of the reference's tests, on host arrays (y
// fastest; element (x, y, z) of an (nx, ny, nz) array at y + x*ny + z*nx*ny):
// the discrete curl of an edge potential to face velocities
// (StreamInitial::GradientStream, WaveletPotentials/StreamInitial.cpp:87-175),
// the zeroing of an array's two end slices normal to one axis (SetSliceZero,
// StreamInitial.cpp:5-43), and the relative L2 error (CalculateErrorL2,
// WaveletPotentials/Utils.cpp:184-207, 296-331), accumulated in FP64.
#include <cmath>
#include <cstdint>

#include "whhd_internal.h"

namespace {

using whhd::fail;

inline int64_t at(int x, int y, int z, int nx, int ny) {
    return y + (int64_t)x * ny + (int64_t)z * nx * ny;
}

}  // namespace

extern "C" {

WHHD_API int whhd_set_slice_zero(float* val, int nx, int ny, int nz, char axis) {
    if (!val || nx <= 0 || ny <= 0 || nz <= 0) return fail(WHHD_INVALID_ARG, "set_slice_zero: bad argument");
    switch (axis) {
    case 'x':
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) val[at(0, y, z, nx, ny)] = val[at(nx - 1, y, z, nx, ny)] = 0.f;
        return WHHD_OK;
    case 'y':
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x) val[at(x, 0, z, nx, ny)] = val[at(x, ny - 1, z, nx, ny)] = 0.f;
        return WHHD_OK;
    case 'z':
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) val[at(x, y, 0, nx, ny)] = val[at(x, y, nz - 1, nx, ny)] = 0.f;
        return WHHD_OK;
    default:
        return fail(WHHD_INVALID_ARG, "set_slice_zero: axis must be 'x', 'y' or 'z'");
    }
}

// xv = dqz/dy - dqy/dz, yv = dqx/dz - dqz/dx, zv = dqy/dx - dqx/dy on the MAC grid
// of nx*ny*nz cells: qx (nx, ny+1, nz+1), qy (nx+1, ny, nz+1), qz (nx+1, ny+1, nz).
WHHD_API int whhd_gradient_stream3d(float* xv, float* yv, float* zv, const float* qx,
                                    const float* qy, const float* qz, int nx, int ny, int nz,
                                    double dx, double dy, double dz) {
    if (!xv || !yv || !zv || !qx || !qy || !qz || nx <= 0 || ny <= 0 || nz <= 0 ||
        !(dx > 0 && dy > 0 && dz > 0))
        return fail(WHHD_INVALID_ARG, "gradient_stream3d: bad argument");
    const float ix = (float)(1.0 / dx), iy = (float)(1.0 / dy), iz = (float)(1.0 / dz);
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x <= nx; ++x)
            for (int y = 0; y < ny; ++y)
                xv[at(x, y, z, nx + 1, ny)] =
                    (qz[at(x, y + 1, z, nx + 1, ny + 1)] - qz[at(x, y, z, nx + 1, ny + 1)]) * iy -
                    (qy[at(x, y, z + 1, nx + 1, ny)] - qy[at(x, y, z, nx + 1, ny)]) * iz;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y <= ny; ++y)
                yv[at(x, y, z, nx, ny + 1)] =
                    (qx[at(x, y, z + 1, nx, ny + 1)] - qx[at(x, y, z, nx, ny + 1)]) * iz -
                    (qz[at(x + 1, y, z, nx + 1, ny + 1)] - qz[at(x, y, z, nx + 1, ny + 1)]) * ix;
    for (int z = 0; z <= nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y)
                zv[at(x, y, z, nx, ny)] =
                    (qy[at(x + 1, y, z, nx + 1, ny)] - qy[at(x, y, z, nx + 1, ny)]) * ix -
                    (qx[at(x, y + 1, z, nx, ny + 1)] - qx[at(x, y, z, nx, ny + 1)]) * iy;
    return WHHD_OK;
}

// 2D: xv = dqz/dy, yv = -dqz/dx; qz (nx+1, ny+1), xv (nx+1, ny), yv (nx, ny+1).
WHHD_API int whhd_gradient_stream2d(float* xv, float* yv, const float* qz, int nx, int ny,
                                    double dx, double dy) {
    if (!xv || !yv || !qz || nx <= 0 || ny <= 0 || !(dx > 0 && dy > 0))
        return fail(WHHD_INVALID_ARG, "gradient_stream2d: bad argument");
    const float ix = (float)(1.0 / dx), iy = (float)(1.0 / dy);
    for (int x = 0; x <= nx; ++x)
        for (int y = 0; y < ny; ++y)
            xv[y + (int64_t)x * ny] = (qz[y + 1 + (int64_t)x * (ny + 1)] - qz[y + (int64_t)x * (ny + 1)]) * iy;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y <= ny; ++y)
            yv[y + (int64_t)x * (ny + 1)] =
                -(qz[y + (int64_t)(x + 1) * (ny + 1)] - qz[y + (int64_t)x * (ny + 1)]) * ix;
    return WHHD_OK;
}

// sqrt(sum (base - check)^2 / sum base^2) over `count` arrays given as
// (base, check, n) triples; the denominator gets eps = 1e-12 when count > 1, as
// the reference's three-array form does.
WHHD_API int whhd_error_l2(const float* const* base, const float* const* check,
                           const long long* n, int count, double* out) {
    if (!base || !check || !n || !out || count < 1) return fail(WHHD_INVALID_ARG, "error_l2: bad argument");
    double err = 0.0, tot = 0.0;
    for (int a = 0; a < count; ++a) {
        if (!base[a] || !check[a] || n[a] < 0) return fail(WHHD_INVALID_ARG, "error_l2: bad array");
        for (long long i = 0; i < n[a]; ++i) {
            const double d = (double)base[a][i] - check[a][i];
            err += d * d;
            tot += (double)base[a][i] * base[a][i];
        }
    }
    *out = std::sqrt(err / (tot + (count > 1 ? 1e-12 : 0.0)));
    return WHHD_OK;
}

}  // extern "C"
