// Per-cell local projection in wavelet space.
//
// 3D, divergence bc 'd' (velocity bc 'n'): cuProjectLocal_q_3d_dq_ccc
// (WaveletPotentials/cuDivLocalProject3D.cu:4013-4246; 9-argument form with
// v_out = s x q :5717-5965).  For cell (x, y, z) the kernel reads xv(x+1,y,z),
// yv(x,y+1,z), zv(x,y,z+1) and writes qx(x,y+1,z+1), qy(x+1,y,z+1),
// qz(x+1,y+1,z): interior cells get q = (s x v)/|s|^2, cells on one index-0
// plane get one Gram-weighted component, and the index-0 lines and corner
// write zeros, covering every potential coefficient exactly once.
//
// 3D, divergence bc 'n' (velocity bc 'd'): cuProjectLocal_q_3d_nq_ccc
// (cuDivLocalProject3D.cu:4474-4877).  Same interior and the same coverage; a
// cell on an index-0 plane also reads the index-0 sample of the in-plane
// component along the plane's normal (xv(0,y,z), yv(x,0,z) or zv(x,y,0)),
// whose two samples give two coefficients directly (divided by the in-plane
// scale, cuCalculateP0Val_single :156-180); the remaining component is
// corrected by their difference over the domain length and projected with the
// Gram weights.  Lines take the direct form twice, the corner integrates with
// the domain lengths.  Its v_out form (:5311-5715) leaves the transformed
// velocity unchanged, so v_out receives the velocity as read.
//
// 2D: cuProjectLocal_q_single (WaveletPotentials/cuDivLocalProject.cu:765-1045):
// node (X, Y) with X, Y > 1 gets
// (xv*sy*B1x*B0y - yv*sx*B0x*B1y) / (sx^2*B0x*B1y + sy^2*B1x*B0y + eps);
// the kernel there evaluates this in two tile orientations and adds halves.
// Divergence bc 'n' (velocity bc 'd'), cuProjectLocal_q_single_n (:1047-1478):
// the same interior; nodes X = 0, 1 (Y > 1) solve the 2x2 system that couples
// the two x scaling functions (Gram entries <b0,b0>, <b0,b1>, scale
// 1/length), nodes Y = 0, 1 (X > 1) the y analogue, and the four corner nodes
// apply a fixed table (by level-count difference) to xv(0,1), xv(1,1),
// yv(1,0), yv(1,1).  The table carries no length, so that path reproduces
// curl q = u exactly only on the unit square, as the reference does.
//
// Both are single streaming passes, one thread per cell, y fastest so that
// every wavefront touches contiguous rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "whhd_internal.h"

namespace whhd {

#ifndef WHHD_PROJ_ZB
#define WHHD_PROJ_ZB 4
#endif
#ifndef WHHD_PROJ_NT
#define WHHD_PROJ_NT 0
#endif

constexpr double kEps = 1e-12;  // eps<DTYPE>, WaveletPotentials/Utils.h:52-53

static int bitlen(int v) {
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

// Per-axis coefficient rows: s[c], B0[c] (projection-type Gram diagonal at c),
// B1[c] (divergence-type Gram diagonal at c+1).
struct AxisCoef {
    const float* s;
    const float* b0;
    const float* b1;
};

// Problems of a batch lie back to back: velocity component c of problem b at
// b * bv[c] elements from the first one, potential component c at b * bq[c].
struct Proj3Args {
    float *qx, *qy, *qz, *xo, *yo, *zo;
    const float *xv, *yv, *zv;
    AxisCoef cx, cy, cz;
    int nx, ny, nz;
    int zblocks;  // grid z rows per problem
    int64_t bv[3], bq[3];
    int nq;                       // divergence bc 'n' boundary cells
    float lix, liy, liz;          // 1 / domain length per axis
    float lenx, leny, lenz;       // domain length per axis
};

// One thread per (x, y) column walking ZB consecutive z cells, y fastest
// within the workgroup; offsets in index type I (32-bit whenever every array
// fits).  Tile row tz covers z block (tz % zblocks) of problem
// (tz / zblocks).
template <class I, int ZB>
__global__ void __launch_bounds__(256) project3d_kernel(Proj3Args p) {
    // workgroups go to the 8 XCDs round-robin in launch order; each XCD takes a
    // contiguous run of tiles, so rows that share a cache line share an L2
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned nwg = gx * gy * gridDim.z;
    const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned r8 = lin % 8u;
    const unsigned tile = r8 * (nwg / 8u) + (r8 < nwg % 8u ? r8 : nwg % 8u) + lin / 8u;
    const unsigned tx = tile % gx, ty = (tile / gx) % gy, tz = tile / (gx * gy);
    const int y = tx * blockDim.x + threadIdx.x;
    const int x = ty * blockDim.y + threadIdx.y;
    const int pb = tz / p.zblocks;
    const int z0 = (tz - pb * p.zblocks) * ZB;
    if (y >= p.ny || x >= p.nx) return;
    if (pb) {
        p.xv += pb * p.bv[0]; p.yv += pb * p.bv[1]; p.zv += pb * p.bv[2];
        p.qx += pb * p.bq[0]; p.qy += pb * p.bq[1]; p.qz += pb * p.bq[2];
        if (p.xo) { p.xo += pb * p.bv[0]; p.yo += pb * p.bv[1]; p.zo += pb * p.bv[2]; }
    }
    const I NY = p.ny, NX = p.nx;
    const I sxv_z = (NX + 1) * NY, syv_z = NX * (NY + 1), szv_z = NX * NY;
    const I sqx_x = NY + 1, sqx_z = NX * (NY + 1);
    const I sqy_x = NY, sqy_z = (NX + 1) * NY;
    const I sqz_x = NY + 1, sqz_z = (NX + 1) * (NY + 1);
    I ixv = y + (x + 1) * NY + z0 * sxv_z;
    I iyv = (y + 1) + x * (NY + 1) + z0 * syv_z;
    I izv = y + x * NY + (z0 + 1) * szv_z;
    I iqx = (y + 1) + x * sqx_x + (z0 + 1) * sqx_z;
    I iqy = y + (x + 1) * sqy_x + (z0 + 1) * sqy_z;
    I iqz = (y + 1) + (x + 1) * sqz_x + z0 * sqz_z;
    const float sx = p.cx.s[x], sy = p.cy.s[y];
    const float sxx_syy = sx * sx + sy * sy;
    const int zend = min(z0 + ZB, p.nz);
    // The velocities of all the thread's cells are loaded before any store:
    // as far as the compiler knows the arrays may alias, so a load inside the
    // loop would wait behind the previous cell's stores and every cell would
    // pay the whole memory latency.  Indices past the last cell are clamped to
    // it (always valid, never used).
    float pvx[ZB], pvy[ZB], pvz[ZB], psz[ZB];
#pragma unroll
    for (int k = 0; k < ZB; ++k) {
        const I kk = min(k, zend - z0 - 1);
#if WHHD_PROJ_NT
        pvx[k] = __builtin_nontemporal_load(p.xv + ixv + kk * sxv_z);
        pvy[k] = __builtin_nontemporal_load(p.yv + iyv + kk * syv_z);
        pvz[k] = __builtin_nontemporal_load(p.zv + izv + kk * szv_z);
#else
        pvx[k] = p.xv[ixv + kk * sxv_z];
        pvy[k] = p.yv[iyv + kk * syv_z];
        pvz[k] = p.zv[izv + kk * szv_z];
#endif
        psz[k] = p.cz.s[z0 + kk];
    }
#pragma unroll
    for (int k = 0; k < ZB; ++k) {
        const int z = z0 + k;
        if (z >= zend) break;
        const float sz = psz[k];
        const float vx = pvx[k], vy = pvy[k], vz = pvz[k];
        float lx = 0.f, ly = 0.f, lz = 0.f;
        if (x > 0 && y > 0 && z > 0) {
            const float inv = 1.f / (sxx_syy + sz * sz);
            lz = (sy * vx - sx * vy) * inv;
            lx = (sz * vy - sy * vz) * inv;
            ly = (sx * vz - sz * vx) * inv;
        } else {
            const float ip0x = p.cx.b0[x], ip1x = p.cx.b1[x];
            const float ip0y = p.cy.b0[y], ip1y = p.cy.b1[y];
            const float ip0z = p.cz.b0[z], ip1z = p.cz.b1[z];
            const float eps = (float)kEps;
            if (p.nq) {
                // index-0 samples of the in-plane components (read only where used)
                if (x > 0 && y > 0 && z == 0) {
                    const float z0 = p.zv[izv - szv_z];
                    const float sxBxBy = sx * ip0x * ip1y * ip0z;
                    const float syBxBy = sy * ip1x * ip0y * ip0z;
                    const float mz = 1.f / (sx * sxBxBy + sy * syBxBy + eps);
                    const float a0 = z0 / sx, a1 = vz / sx;
                    p.qx[iqx - sqx_z] = 0.f;
                    p.qy[iqy - sqy_z] = a0;
                    ly = a1;
                    const float vx2 = vx + (a1 - a0) * p.liz;
                    lz = (vx2 * syBxBy - vy * sxBxBy) * mz;
                } else if (x > 0 && z > 0 && y == 0) {
                    const float y0 = p.yv[iyv - 1];
                    const float sxBxBz = sx * ip0x * ip0y * ip1z;
                    const float szBxBz = sz * ip1x * ip0y * ip0z;
                    const float my = 1.f / (sx * sxBxBz + sz * szBxBz + eps);
                    const float a0 = y0 / sz, a1 = vy / sz;
                    p.qx[iqx - 1] = a0;
                    lx = a1;
                    const float vz2 = vz + (a1 - a0) * p.liy;
                    ly = (vz2 * sxBxBz - vx * szBxBz) * my;
                    p.qz[iqz - 1] = 0.f;
                } else if (y > 0 && z > 0 && x == 0) {
                    const float x0 = p.xv[ixv - NY];
                    const float syByBz = sy * ip0x * ip0y * ip1z;
                    const float szByBz = sz * ip0x * ip1y * ip0z;
                    const float mx = 1.f / (sy * syByBz + sz * szByBz + eps);
                    const float a0 = x0 / sy, a1 = vx / sy;
                    p.qz[iqz - sqz_x] = a0;
                    lz = a1;
                    const float vy2 = vy + (a1 - a0) * p.lix;
                    lx = (vy2 * szByBz - vz * syByBz) * mx;
                    p.qy[iqy - sqy_x] = 0.f;
                } else if (x > 0 && y == 0 && z == 0) {
                    const float y0 = p.yv[iyv - 1], z0 = p.zv[izv - szv_z];
                    p.qy[iqy - sqy_z] = z0 / sx;
                    ly = vz / sx;
                    p.qz[iqz - 1] = -y0 / sx;
                    lz = -vy / sx;
                    p.qx[iqx - 1] = 0.f;
                    p.qx[iqx - sqx_z - 1] = 0.f;
                    p.qx[iqx - sqx_z] = 0.f;
                } else if (y > 0 && x == 0 && z == 0) {
                    const float z0 = p.zv[izv - szv_z], x0 = p.xv[ixv - NY];
                    p.qx[iqx - sqx_z] = -z0 / sy;
                    lx = -vz / sy;
                    p.qz[iqz - sqz_x] = x0 / sy;
                    lz = vx / sy;
                    p.qy[iqy - sqy_x - sqy_z] = 0.f;
                    p.qy[iqy - sqy_z] = 0.f;
                    p.qy[iqy - sqy_x] = 0.f;
                } else if (z > 0 && x == 0 && y == 0) {
                    const float y0 = p.yv[iyv - 1], x0 = p.xv[ixv - NY];
                    p.qx[iqx - 1] = y0 / sz;
                    lx = vy / sz;
                    p.qy[iqy - sqy_x] = -x0 / sz;
                    ly = -vx / sz;
                    p.qz[iqz - sqz_x - 1] = 0.f;
                    p.qz[iqz - sqz_x] = 0.f;
                    p.qz[iqz - 1] = 0.f;
                } else {
                    const float y0 = p.yv[iyv - 1], x0 = p.xv[ixv - NY], z0 = p.zv[izv - szv_z];
                    p.qz[iqz - 1 - sqz_x] = 0.f;
                    p.qz[iqz - 1] = 0.f;
                    p.qz[iqz - sqz_x] = 0.f;
                    const float q010 = -z0 * p.leny;
                    p.qx[iqx - 1 - sqx_z] = 0.f;
                    p.qx[iqx - sqx_z] = q010;
                    p.qx[iqx - 1] = y0 * p.lenz;
                    lx = q010 + vy * p.lenz;
                    p.qy[iqy - sqy_x - sqy_z] = 0.f;
                    p.qy[iqy - sqy_z] = 0.f;
                    p.qy[iqy - sqy_x] = -x0 * p.lenz;
                    ly = -vx * p.lenz;
                }
            } else if (y > 0 && z > 0 && x == 0) {
                const float syByBz = sy * ip0x * ip0y * ip1z;
                const float szByBz = sz * ip0x * ip1y * ip0z;
                const float mx = 1.f / (sy * syByBz + sz * szByBz + eps);
                lx = (vy * szByBz - vz * syByBz) * mx;
                p.qz[iqz - sqz_x] = 0.f;
                p.qy[iqy - sqy_x] = 0.f;
            } else if (x > 0 && z > 0 && y == 0) {
                const float sxBxBz = sx * ip0x * ip0y * ip1z;
                const float szBxBz = sz * ip1x * ip0y * ip0z;
                const float my = 1.f / (sx * sxBxBz + sz * szBxBz + eps);
                ly = (vz * sxBxBz - vx * szBxBz) * my;
                p.qx[iqx - 1] = 0.f;
                p.qz[iqz - 1] = 0.f;
            } else if (x > 0 && y > 0 && z == 0) {
                const float sxBxBy = sx * ip0x * ip1y * ip0z;
                const float syBxBy = sy * ip1x * ip0y * ip0z;
                const float mz = 1.f / (sx * sxBxBy + sy * syBxBy + eps);
                lz = (vx * syBxBy - vy * sxBxBy) * mz;
                p.qx[iqx - sqx_z] = 0.f;
                p.qy[iqy - sqy_z] = 0.f;
            } else if (y == 0 && x > 0 && z == 0) {
                p.qy[iqy - sqy_z] = 0.f;
                p.qz[iqz - 1] = 0.f;
                p.qx[iqx - 1] = 0.f;
                p.qx[iqx - sqx_z - 1] = 0.f;
                p.qx[iqx - sqx_z] = 0.f;
            } else if (y > 0 && x == 0 && z == 0) {
                p.qx[iqx - sqx_z] = 0.f;
                p.qz[iqz - sqz_x] = 0.f;
                p.qy[iqy - sqy_x - sqy_z] = 0.f;
                p.qy[iqy - sqy_z] = 0.f;
                p.qy[iqy - sqy_x] = 0.f;
            } else if (y == 0 && x == 0 && z > 0) {
                p.qx[iqx - 1] = 0.f;
                p.qy[iqy - sqy_x] = 0.f;
                p.qz[iqz - sqz_x - 1] = 0.f;
                p.qz[iqz - sqz_x] = 0.f;
                p.qz[iqz - 1] = 0.f;
            } else {
                p.qz[iqz - 1 - sqz_x] = 0.f;
                p.qz[iqz - 1] = 0.f;
                p.qz[iqz - sqz_x] = 0.f;
                p.qx[iqx - 1 - sqx_z] = 0.f;
                p.qx[iqx - sqx_z] = 0.f;
                p.qx[iqx - 1] = 0.f;
                p.qy[iqy - sqy_x - sqy_z] = 0.f;
                p.qy[iqy - sqy_z] = 0.f;
                p.qy[iqy - sqy_x] = 0.f;
            }
        }
        p.qx[iqx] = lx;
        p.qy[iqy] = ly;
        p.qz[iqz] = lz;
        if (p.xo) {
            if (p.nq) {
                p.xo[ixv] = vx;
                p.yo[iyv] = vy;
                p.zo[izv] = vz;
            } else {
                p.xo[ixv] = sy * lz - sz * ly;
                p.yo[iyv] = sz * lx - sx * lz;
                p.zo[izv] = sx * ly - sy * lx;
            }
        }
        ixv += sxv_z; iyv += syv_z; izv += szv_z;
        iqx += sqx_z; iqy += sqy_z; iqz += sqz_z;
    }
}

struct Proj2Args {
    float* qz;
    const float *xv, *yv;
    AxisCoef cx, cy;  // indexed by node X (resp. Y): s from X-1, B0[X-1], B1[X]
    int nx, ny;
    int nq;                             // divergence bc 'n'
    float sxl, syl;                     // 1 / domain length
    float bx00, bx01, by00, by01;       // <b0,b0>, <b0,b1> of the divergence basis
    float p0x, p0y;                     // projection-basis Gram diagonal at 0
    float corner[16];                   // level-difference table
};

__global__ void __launch_bounds__(256) project2d_kernel(Proj2Args p) {
    const int Y = blockIdx.x * 64 + threadIdx.x;
    const int X = blockIdx.y * 4 + threadIdx.y;
    if (Y > p.ny || X > p.nx) return;
    const int64_t NY = p.ny;
    // node-indexed velocities: xv(X, Y) = xv[(Y-1) + X NY] (0 at Y = 0),
    // yv(X, Y) = yv[Y + (X-1)(NY+1)] (0 at X = 0)
    auto XV = [&](int x, int y) { return y >= 1 ? p.xv[(y - 1) + x * NY] : 0.f; };
    auto YV = [&](int x, int y) { return x >= 1 ? p.yv[y + (x - 1) * (NY + 1)] : 0.f; };
    float q = 0.f;
    if (X > 1 && Y > 1) {
        const float vx = XV(X, Y), vy = YV(X, Y);
        const float sx = p.cx.s[X], sy = p.cy.s[Y];
        const float ip0x = p.cx.b0[X], ip1x = p.cx.b1[X];
        const float ip0y = p.cy.b0[Y], ip1y = p.cy.b1[Y];
        const float mid = sx * sx * ip0x * ip1y + sy * sy * ip1x * ip0y + (float)kEps;
        q = (vx * sy * ip1x * ip0y - vy * sx * ip0x * ip1y) / mid;
    } else if (p.nq) {
        const float eps = (float)kEps;
        if (Y > 1) {  // X = 0, 1
            const float sy = p.cy.s[Y], ip0y = p.cy.b0[Y], ip1y = p.cy.b1[Y];
            const float xv0 = XV(0, Y), xv1 = XV(1, Y), v = YV(1, Y);
            const float ax = p.sxl * p.sxl * p.p0x * ip1y;
            const float p0 = ax + sy * sy * p.bx00 * ip0y;
            const float p1 = -ax + sy * sy * p.bx01 * ip0y;
            const float den = p0 * p0 - p1 * p1 + eps;
            const float m0 = p0 / den, m1 = -p1 / den;
            const float g = v * p.sxl * p.p0x * ip1y;
            const float f1 = g + xv0 * sy * p.bx00 * ip0y + xv1 * sy * p.bx01 * ip0y;
            const float f2 = -g + xv0 * sy * p.bx01 * ip0y + xv1 * sy * p.bx00 * ip0y;
            q = X == 0 ? m0 * f1 + m1 * f2 : m1 * f1 + m0 * f2;
        } else if (X > 1) {  // Y = 0, 1
            const float sx = p.cx.s[X], ip0x = p.cx.b0[X], ip1x = p.cx.b1[X];
            const float v = XV(X, 1), yv0 = YV(X, 0), yv1 = YV(X, 1);
            const float ay = p.syl * p.syl * ip1x * p.p0y;
            const float p0 = sx * sx * ip0x * p.by00 + ay;
            const float p1 = sx * sx * ip0x * p.by01 - ay;
            const float den = p0 * p0 - p1 * p1 + eps;
            const float m0 = p0 / den, m1 = -p1 / den;
            const float g = v * p.syl * ip1x * p.p0y;
            const float f1 = -g - yv0 * sx * ip0x * p.by00 - yv1 * sx * ip0x * p.by01;
            const float f2 = g - yv0 * sx * ip0x * p.by01 - yv1 * sx * ip0x * p.by00;
            q = Y == 0 ? m0 * f1 + m1 * f2 : m1 * f1 + m0 * f2;
        } else {  // corner
            const float* a = p.corner + (2 * Y + X) * 4;
            q = a[2] * XV(0, 1) + a[3] * XV(1, 1) - a[0] * YV(1, 0) - a[1] * YV(1, 1);
        }
    }
    p.qz[Y + X * (NY + 1)] = q;
}

// ---------------------------------------------------------------------------
// Host: coefficient tables
// ---------------------------------------------------------------------------
// The first `rows` rows of the 8-row Gram table (row 0: <b_i, b_i>, row 3:
// <b_i, b_i+1>), row r at r * n.
static int gram_rows_host(std::vector<double>& out, int wt, char bc, int n, int rows) {
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * 8 * (size_t)n) != hipSuccess)
        return fail(WHHD_OOM, "gram table allocation failed");
    int st = gram_table_device(d, wt, bc, n, nullptr);
    if (st == WHHD_OK) {
        out.resize((size_t)rows * n);
        hipError_t e = hipMemcpy(out.data(), d, sizeof(double) * rows * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = hip_fail(e, "gram table copy");
    }
    hipFree(d);
    return st;
}

static int gram_diag_host(std::vector<double>& diag, int wt, char bc, int n) {
    return gram_rows_host(diag, wt, bc, n, 1);
}

// GetDivL0Matrix (WaveletPotentials/DivL0Matrix.cpp:3-50): four numbers (a, b,
// c, e) by the difference d of the two level counts, swapped pairwise when the
// y axis has more levels, expanded to the 4x4 sign pattern.  For d = 1..8 the
// reference's table equals a, b = -2^(d-2) (1 +- 1/(4^d + 1)),
// c, e = -(1/4 +- 4^(d-1)/(4^d + 1)), computed here from that form; d = 0 or
// d > 8 take (-3/8, -1/8, -3/8, -1/8).  At d = 5 the reference initialises
// only the first entry, with the sum of the four (DivL0Matrix.cpp:27), and so
// does this table.
static void div_l0_table(float out[16], int lx, int ly) {
    const int m = lx < ly ? lx : ly;
    const int ex = lx - m, ey = ly - m;
    const int d = ex > ey ? ex : ey;
    double a = -0.375, b = -0.125, c = -0.375, e = -0.125;
    if (d >= 1 && d <= 8) {
        const double r = 1.0 / (std::ldexp(1.0, 2 * d) + 1.0);
        const double h = std::ldexp(1.0, d - 2), t = std::ldexp(1.0, 2 * d - 2) * r;
        a = -h * (1.0 + r); b = -h * (1.0 - r);
        c = -(0.25 + t); e = -(0.25 - t);
        if (d == 5) { a = a + b + c + e; b = c = e = 0.0; }
    }
    if (ex < ey) { std::swap(a, c); std::swap(b, e); }
    const double v[16] = {a, b, c, e, -a, -b, e, c, b, a, -c, -e, -b, -a, -e, -c};
    for (int i = 0; i < 16; ++i) out[i] = (float)v[i];
}

// rows: [s | B0 | B1], each of `len` floats.  mode3d: index c is the cell
// (s from bitlen(c), B0[c], B1[c+1]); otherwise index X is the node
// (s from bitlen(X-1), B0[X-1], B1[X]).
static int axis_coef(std::vector<float>& rows, int N, double h, int wt_div, char bc_div,
                     bool mode3d) {
    const int L = ilog2(N);
    std::vector<double> base, proj;
    int st = gram_diag_host(base, wt_div, bc_div, N + 1);
    if (st != WHHD_OK) return st;
    st = gram_diag_host(proj, proj_wavelet(wt_div), proj_bc(bc_div), N);
    if (st != WHHD_OK) return st;
    const int len = N + 1;
    rows.assign(3 * (size_t)len, 0.f);
    const double hinv = 1.0 / h;
    for (int c = 0; c < len; ++c) {
        int bl, i0, i1;
        if (mode3d) { bl = bitlen(c); i0 = c; i1 = c + 1; }
        else { bl = c >= 1 ? bitlen(c - 1) : 0; i0 = c - 1; i1 = c; }
        const int j = L - bl + 1;
        const double s = 4.0 * hinv / (std::ldexp(1.0, j) + kEps);
        rows[c] = (float)s;
        rows[len + c] = (i0 >= 0 && i0 < N) ? (float)proj[i0] : 0.f;
        rows[2 * len + c] = (i1 >= 0 && i1 <= N) ? (float)base[i1] : 0.f;
    }
    return WHHD_OK;
}

}  // namespace whhd

struct whhd_project3d_s {
    int nx, ny, nz;
    char bc;                // divergence bc
    double len[3];          // domain lengths nx dx, ny dy, nz dz
    float* coef = nullptr;  // 3 axes x 3 rows x (n+1)
    whhd::AxisCoef cx, cy, cz;
};

struct whhd_project2d_s {
    int nx, ny;
    float* coef = nullptr;
    whhd::AxisCoef cx, cy;
    bool nq = false;  // divergence bc 'n': boundary constants below
    float sxl = 0.f, syl = 0.f, bx00 = 0.f, bx01 = 0.f, by00 = 0.f, by01 = 0.f, p0x = 0.f,
          p0y = 0.f;
    float corner[16] = {};
};

namespace whhd {

static bool is_pow2(int n) { return n >= 4 && (n & (n - 1)) == 0; }

int project3d_create(int nx, int ny, int nz, double dx, double dy, double dz, int wt, char bc,
                     whhd_project3d_s** out) {
    if (!out) return fail(WHHD_INVALID_ARG, "project3d_create: null handle pointer");
    *out = nullptr;
    if (!is_pow2(nx) || !is_pow2(ny) || !is_pow2(nz))
        return fail(WHHD_NOT_POW2, "project3d: cell counts must be powers of two (>= 4)");
    if (!(dx > 0 && dy > 0 && dz > 0)) return fail(WHHD_INVALID_ARG, "project3d: spacing must be > 0");
    if (bc != 'd' && bc != 'n')
        return fail(WHHD_UNSUPPORTED_WAVELET,
                    "project3d: the divergence bc must be 'd' (velocity bc 'n') or 'n' (velocity bc 'd')");
    int st = check_wavelet(wt, bc);
    if (st != WHHD_OK) return st;
    if (proj_wavelet(wt) < 0) return fail(WHHD_UNSUPPORTED_WAVELET, "project3d: no projection type");
    std::vector<float> rx, ry, rz;
    if ((st = axis_coef(rx, nx, dx, wt, bc, true)) != WHHD_OK) return st;
    if ((st = axis_coef(ry, ny, dy, wt, bc, true)) != WHHD_OK) return st;
    if ((st = axis_coef(rz, nz, dz, wt, bc, true)) != WHHD_OK) return st;
    auto* h = new whhd_project3d_s;
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->bc = bc;
    h->len[0] = nx * dx; h->len[1] = ny * dy; h->len[2] = nz * dz;
    const size_t tot = rx.size() + ry.size() + rz.size();
    if (hipMalloc(&h->coef, tot * sizeof(float)) != hipSuccess) {
        delete h;
        return fail(WHHD_OOM, "project3d: coefficient allocation failed");
    }
    std::vector<float> all;
    all.insert(all.end(), rx.begin(), rx.end());
    all.insert(all.end(), ry.begin(), ry.end());
    all.insert(all.end(), rz.begin(), rz.end());
    hipError_t e = hipMemcpy(h->coef, all.data(), tot * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(h->coef);
        delete h;
        return hip_fail(e, "project3d: coefficient upload");
    }
    const float* b = h->coef;
    h->cx = {b, b + (nx + 1), b + 2 * (nx + 1)};
    b += 3 * (nx + 1);
    h->cy = {b, b + (ny + 1), b + 2 * (ny + 1)};
    b += 3 * (ny + 1);
    h->cz = {b, b + (nz + 1), b + 2 * (nz + 1)};
    *out = h;
    return WHHD_OK;
}

template <int ZB>
static void launch_project3d(const Proj3Args& a, int nb, bool wide, hipStream_t st) {
    // a workgroup spans up to 256 consecutive y of 256 / that many rows, so the
    // potentials' rows (ny + 1 floats, not 128-byte aligned) are written by one
    // workgroup each and few cache lines are shared between workgroups
    const int yb = a.ny >= 256 ? 256 : (a.ny >= 128 ? 128 : 64);
    dim3 block(yb, 256 / yb, 1);
    dim3 grid((a.ny + yb - 1) / yb, (a.nx + block.y - 1) / block.y, (unsigned)(a.zblocks * nb));
    if (!wide) hipLaunchKernelGGL((project3d_kernel<int, ZB>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((project3d_kernel<int64_t, ZB>), grid, block, 0, st, a);
}

int project3d_run(const whhd_project3d_s* h, float* qx, float* qy, float* qz, float* xo, float* yo,
                  float* zo, const float* xv, const float* yv, const float* zv, hipStream_t st,
                  int nb, int zb) {
    if (!h || !qx || !qy || !qz || !xv || !yv || !zv)
        return fail(WHHD_INVALID_ARG, "project3d_run: null argument");
    if ((xo || yo || zo) && !(xo && yo && zo))
        return fail(WHHD_INVALID_ARG, "project3d_run: give all three or none of the v_out arrays");
    if (zb <= 0) zb = WHHD_PROJ_ZB;
    if (zb != 1 && zb != 2 && zb != 4 && zb != 8)
        return fail(WHHD_INVALID_ARG, "project3d_run: z block must be 1, 2, 4 or 8");
    const int64_t X = h->nx, Y = h->ny, Z = h->nz;
    Proj3Args a{qx, qy, qz, xo, yo, zo, xv, yv, zv, h->cx, h->cy, h->cz, h->nx, h->ny, h->nz,
                (h->nz + zb - 1) / zb,
                {(X + 1) * Y * Z, X * (Y + 1) * Z, X * Y * (Z + 1)},
                {X * (Y + 1) * (Z + 1), (X + 1) * Y * (Z + 1), (X + 1) * (Y + 1) * Z},
                h->bc == 'n',
                (float)(1.0 / h->len[0]), (float)(1.0 / h->len[1]), (float)(1.0 / h->len[2]),
                (float)h->len[0], (float)h->len[1], (float)h->len[2]};
    if (nb < 1 || (int64_t)a.zblocks * nb > 65535)
        return fail(WHHD_INVALID_ARG, "project3d_run: batch too large for one launch");
    const bool wide = (X + 1) * (Y + 1) * (Z + 1) >= ((int64_t)1 << 31);
    switch (zb) {
    case 1: launch_project3d<1>(a, nb, wide, st); break;
    case 2: launch_project3d<2>(a, nb, wide, st); break;
    case 4: launch_project3d<4>(a, nb, wide, st); break;
    default: launch_project3d<8>(a, nb, wide, st); break;
    }
    set_last_kernel("project3d_kernel");
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "project3d kernel");
}

int project2d_create(int nx, int ny, double dx, double dy, int wt, char bc, whhd_project2d_s** out) {
    if (!out) return fail(WHHD_INVALID_ARG, "project2d_create: null handle pointer");
    *out = nullptr;
    if (!is_pow2(nx) || !is_pow2(ny))
        return fail(WHHD_NOT_POW2, "project2d: cell counts must be powers of two (>= 4)");
    if (!(dx > 0 && dy > 0)) return fail(WHHD_INVALID_ARG, "project2d: spacing must be > 0");
    if (bc != 'd' && bc != 'n')
        return fail(WHHD_UNSUPPORTED_WAVELET,
                    "project2d: the divergence bc must be 'd' (velocity bc 'n') or 'n' (velocity bc 'd')");
    int st = check_wavelet(wt, bc);
    if (st != WHHD_OK) return st;
    if (proj_wavelet(wt) < 0) return fail(WHHD_UNSUPPORTED_WAVELET, "project2d: no projection type");
    std::vector<float> rx, ry;
    if ((st = axis_coef(rx, nx, dx, wt, bc, false)) != WHHD_OK) return st;
    if ((st = axis_coef(ry, ny, dy, wt, bc, false)) != WHHD_OK) return st;
    std::vector<double> gx, gy, px, py;
    if (bc == 'n') {
        if ((st = gram_rows_host(gx, wt, bc, nx + 1, 4)) != WHHD_OK) return st;
        if ((st = gram_rows_host(gy, wt, bc, ny + 1, 4)) != WHHD_OK) return st;
        if ((st = gram_diag_host(px, proj_wavelet(wt), proj_bc(bc), nx)) != WHHD_OK) return st;
        if ((st = gram_diag_host(py, proj_wavelet(wt), proj_bc(bc), ny)) != WHHD_OK) return st;
    }
    auto* h = new whhd_project2d_s;
    h->nx = nx; h->ny = ny;
    if (bc == 'n') {
        h->nq = true;
        h->sxl = (float)(1.0 / (nx * dx));
        h->syl = (float)(1.0 / (ny * dy));
        h->bx00 = (float)gx[0]; h->bx01 = (float)gx[3 * (size_t)(nx + 1)];
        h->by00 = (float)gy[0]; h->by01 = (float)gy[3 * (size_t)(ny + 1)];
        h->p0x = (float)px[0]; h->p0y = (float)py[0];
        div_l0_table(h->corner, ilog2(nx), ilog2(ny));
    }
    const size_t tot = rx.size() + ry.size();
    if (hipMalloc(&h->coef, tot * sizeof(float)) != hipSuccess) {
        delete h;
        return fail(WHHD_OOM, "project2d: coefficient allocation failed");
    }
    std::vector<float> all(rx);
    all.insert(all.end(), ry.begin(), ry.end());
    hipError_t e = hipMemcpy(h->coef, all.data(), tot * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(h->coef);
        delete h;
        return hip_fail(e, "project2d: coefficient upload");
    }
    const float* b = h->coef;
    h->cx = {b, b + (nx + 1), b + 2 * (nx + 1)};
    b += 3 * (nx + 1);
    h->cy = {b, b + (ny + 1), b + 2 * (ny + 1)};
    *out = h;
    return WHHD_OK;
}

int project2d_run(const whhd_project2d_s* h, float* qz, const float* xv, const float* yv,
                  hipStream_t st) {
    if (!h || !qz || !xv || !yv) return fail(WHHD_INVALID_ARG, "project2d_run: null argument");
    Proj2Args a{qz, xv, yv, h->cx, h->cy, h->nx, h->ny, h->nq ? 1 : 0, h->sxl, h->syl,
                h->bx00, h->bx01, h->by00, h->by01, h->p0x, h->p0y, {}};
    for (int i = 0; i < 16; ++i) a.corner[i] = h->corner[i];
    dim3 block(64, 4, 1);
    dim3 grid((h->ny + 1 + 63) / 64, (h->nx + 1 + 3) / 4, 1);
    hipLaunchKernelGGL(project2d_kernel, grid, block, 0, st, a);
    set_last_kernel("project2d_kernel");
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WHHD_OK : hip_fail(e, "project2d kernel");
}

}  // namespace whhd

extern "C" {

WHHD_API int whhd_project3d_create(int nx, int ny, int nz, double dx, double dy, double dz,
                                   int wavelet_div, char bc_div, whhd_project3d_t* out) {
    return whhd::project3d_create(nx, ny, nz, dx, dy, dz, wavelet_div, bc_div, out);
}

WHHD_API int whhd_project3d_run(whhd_project3d_t h, float* qx, float* qy, float* qz,
                                float* xv_out, float* yv_out, float* zv_out, const float* xv,
                                const float* yv, const float* zv, void* stream) {
    return whhd::project3d_run(h, qx, qy, qz, xv_out, yv_out, zv_out, xv, yv, zv,
                               (hipStream_t)stream, 1, 0);
}

WHHD_API void whhd_project3d_destroy(whhd_project3d_t h) {
    if (!h) return;
    hipFree(h->coef);
    delete h;
}

WHHD_API int whhd_project2d_create(int nx, int ny, double dx, double dy, int wavelet_div,
                                   char bc_div, whhd_project2d_t* out) {
    return whhd::project2d_create(nx, ny, dx, dy, wavelet_div, bc_div, out);
}

WHHD_API int whhd_project2d_run(whhd_project2d_t h, float* qz, const float* xv, const float* yv,
                                void* stream) {
    return whhd::project2d_run(h, qz, xv, yv, (hipStream_t)stream);
}

WHHD_API void whhd_project2d_destroy(whhd_project2d_t h) {
    if (!h) return;
    hipFree(h->coef);
    delete h;
}

}  // extern "C"
