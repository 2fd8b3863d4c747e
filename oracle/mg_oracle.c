/*
 * ORACLE -- test infrastructure only (see whhd_oracle.h).
 *
 * FP64 restatement of the reference's fractional-area multigrid:
 *   hierarchy          WaveletPotentials/cuMultigridFrac3D.cu:9-134
 *   Solve / V-cycle    WaveletPotentials/cuMultigridFrac3D.cu:249-327
 *   dx restriction     WaveletPotentials/cuMultigridFrac2D.cu:4-26
 *   Jacobi sweeps      WaveletPotentials/cuWeightedJacobi3D.cu:620-695 ('n'), 859-936 ('z'),
 *                      1075-1105 (SolveFrac)
 *   residual           WaveletPotentials/cuWeightedJacobi3D.cu:1107-1181, 1384-1424
 *   coarsest level     WaveletPotentials/cuWeightedJacobi3D.cu:1426-1722 (50 sweeps 'n', 20 'z')
 *   restriction        WaveletPotentials/cuRestrict3D.cu:524-882 (p1/p2 kernels),
 *                      939-970 (SolveP1RN with dx), 972-1140 (SolveP1Ax/Ay/Az)
 *   prolongation       WaveletPotentials/cuProlongate3D.cu:353-423, 594-616
 * Arrays are y fastest; cells (nx, ny, nz); face fractions ax (nx+1, ny, nz),
 * ay (nx, ny+1, nz), az (nx, ny, nz+1).  bc 'd' runs as 'z' with zero
 * exterior spacing (cuMultigridFrac3D.cu:12-13): a Dirichlet wall at the face.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "whhd_oracle.h"

typedef struct { int x, y, z; } i3;

#define AT(a, Y_, X_, Z_, D) (a)[(size_t)(Y_) + (size_t)(X_) * (D).y + (size_t)(Z_) * (D).x * (D).y]

static int half_up(int n) { return (n & 1) ? (n >> 1) + 1 : n >> 1; }

/* One weighted Jacobi sweep dst <- src (cuWjFrac3D_n / cuWjFrac3D_z), or the
 * residual f - A src when res != 0 (cuWjGetFracRes_3d_n / _z). */
static void apply(double* dst, const double* src, const double* f, const double* ax,
                  const double* ay, const double* az, i3 d, const double* hx, const double* hy,
                  const double* hz, const double he[3], char bc, double w, int res) {
    const i3 dax = {d.x + 1, d.y, d.z}, day = {d.x, d.y + 1, d.z}, daz = {d.x, d.y, d.z + 1};
    const int z_bc = (bc == 'z');
    for (int z = 0; z < d.z; ++z)
        for (int x = 0; x < d.x; ++x)
            for (int y = 0; y < d.y; ++y) {
                double diag = 0.0, off = 0.0;
                /* (face fraction, neighbour spacing, own spacing, neighbour value, inside) */
                struct { double a, hn, h, v; int in; } fc[6] = {
                    {AT(ay, y, x, z, day), y > 0 ? hy[y - 1] : he[1], hy[y],
                     y > 0 ? AT(src, y - 1, x, z, d) : 0.0, y > 0},
                    {AT(ay, y + 1, x, z, day), y < d.y - 1 ? hy[y + 1] : he[1], hy[y],
                     y < d.y - 1 ? AT(src, y + 1, x, z, d) : 0.0, y < d.y - 1},
                    {AT(ax, y, x, z, dax), x > 0 ? hx[x - 1] : he[0], hx[x],
                     x > 0 ? AT(src, y, x - 1, z, d) : 0.0, x > 0},
                    {AT(ax, y, x + 1, z, dax), x < d.x - 1 ? hx[x + 1] : he[0], hx[x],
                     x < d.x - 1 ? AT(src, y, x + 1, z, d) : 0.0, x < d.x - 1},
                    {AT(az, y, x, z, daz), z > 0 ? hz[z - 1] : he[2], hz[z],
                     z > 0 ? AT(src, y, x, z - 1, d) : 0.0, z > 0},
                    {AT(az, y, x, z + 1, daz), z < d.z - 1 ? hz[z + 1] : he[2], hz[z],
                     z < d.z - 1 ? AT(src, y, x, z + 1, d) : 0.0, z < d.z - 1}};
                for (int k = 0; k < 6; ++k) {
                    if (!fc[k].in && !z_bc) continue;  /* 'n': no flux through the wall */
                    const double cf = fc[k].a / (fc[k].hn + fc[k].h) / fc[k].h * 2.0;
                    diag += cf;
                    if (fc[k].in) off -= cf * fc[k].v;
                }
                const double c = AT(src, y, x, z, d), fv = AT(f, y, x, z, d);
                double out = 0.0;
                if (diag > 1e-12)
                    out = res ? fv - (diag * c + off) : (1.0 - w) * c + w * (fv - off) / diag;
                AT(dst, y, x, z, d) = out;
            }
}

/* Coarsest level (cuWjFrac3D_kernel_n / _z): Jacobi on the volume-scaled
 * system; 50 sweeps for 'n', 20 for 'z', zero where no face is open. */
static void coarsest(double* v, const double* f, const double* ax, const double* ay,
                     const double* az, i3 d, const double* hx, const double* hy, const double* hz,
                     const double he[3], char bc, double w) {
    const size_t n = (size_t)d.x * d.y * d.z;
    const i3 dax = {d.x + 1, d.y, d.z}, day = {d.x, d.y + 1, d.z}, daz = {d.x, d.y, d.z + 1};
    double* p = (double*)malloc(n * sizeof(double));
    double* q = (double*)malloc(n * sizeof(double));
    double* c = (double*)malloc(n * 6 * sizeof(double));
    double* r = (double*)malloc(n * sizeof(double));
    const int z_bc = (bc == 'z');
    for (int z = 0; z < d.z; ++z)
        for (int x = 0; x < d.x; ++x)
            for (int y = 0; y < d.y; ++y) {
                const size_t i = (size_t)y + (size_t)x * d.y + (size_t)z * d.x * d.y;
                const double vol = hx[x] * hy[y] * hz[z];
                /* 'n' clamps the neighbour spacing index, 'z' uses the exterior spacing */
                const double hxl = x > 0 ? hx[x - 1] : (z_bc ? he[0] : hx[0]);
                const double hxr = x < d.x - 1 ? hx[x + 1] : (z_bc ? he[0] : hx[d.x - 1]);
                const double hyb = y > 0 ? hy[y - 1] : (z_bc ? he[1] : hy[0]);
                const double hyt = y < d.y - 1 ? hy[y + 1] : (z_bc ? he[1] : hy[d.y - 1]);
                const double hzf = z > 0 ? hz[z - 1] : (z_bc ? he[2] : hz[0]);
                const double hzb = z < d.z - 1 ? hz[z + 1] : (z_bc ? he[2] : hz[d.z - 1]);
                c[6 * i + 0] = AT(ax, y, x, z, dax) / (hxl + hx[x]) * hy[y] * hz[z] * 2.0;
                c[6 * i + 1] = AT(ax, y, x + 1, z, dax) / (hxr + hx[x]) * hy[y] * hz[z] * 2.0;
                c[6 * i + 2] = AT(ay, y, x, z, day) / (hyb + hy[y]) * hx[x] * hz[z] * 2.0;
                c[6 * i + 3] = AT(ay, y + 1, x, z, day) / (hyt + hy[y]) * hx[x] * hz[z] * 2.0;
                c[6 * i + 4] = AT(az, y, x, z, daz) / (hzf + hz[z]) * hx[x] * hy[y] * 2.0;
                c[6 * i + 5] = AT(az, y, x, z + 1, daz) / (hzb + hz[z]) * hx[x] * hy[y] * 2.0;
                r[i] = f[i] * vol;
                p[i] = v[i];
            }
    const int iters = z_bc ? 20 : 50;
    const size_t sx = (size_t)d.y, sz = (size_t)d.x * d.y;
    for (int it = 0; it < iters; ++it) {
        for (int z = 0; z < d.z; ++z)
            for (int x = 0; x < d.x; ++x)
                for (int y = 0; y < d.y; ++y) {
                    const size_t i = (size_t)y + x * sx + z * sz;
                    const int in[6] = {x > 0, x < d.x - 1, y > 0, y < d.y - 1, z > 0, z < d.z - 1};
                    const long long nb[6] = {-(long long)sx, (long long)sx, -1, 1, -(long long)sz,
                                             (long long)sz};
                    double diag = 0.0, off = 0.0;
                    for (int k = 0; k < 6; ++k) {
                        if (in[k]) {
                            diag += c[6 * i + k];
                            off -= c[6 * i + k] * p[(size_t)((long long)i + nb[k])];
                        } else if (z_bc) {
                            diag += c[6 * i + k];
                        }
                    }
                    q[i] = diag > 0.0 ? (1.0 - w) * p[i] + w * (r[i] - off) / diag : 0.0;
                }
        memcpy(p, q, n * sizeof(double));
    }
    memcpy(v, p, n * sizeof(double));
    free(p); free(q); free(c); free(r);
}

/* cuRestrict3D_p1_dx_x / _dy_y / _dz_z: spacing-weighted average of the one
 * or two children along one axis. */
static void restrict_axis_w(double* dst, const double* src, i3 ds, int axis, const double* h) {
    i3 dd = ds;
    if (axis == 0) dd.x = half_up(ds.x);
    if (axis == 1) dd.y = half_up(ds.y);
    if (axis == 2) dd.z = half_up(ds.z);
    for (int z = 0; z < dd.z; ++z)
        for (int x = 0; x < dd.x; ++x)
            for (int y = 0; y < dd.y; ++y) {
                int c0[3] = {x, y, z};
                const int nmax = axis == 0 ? ds.x : axis == 1 ? ds.y : ds.z;
                const int k0 = 2 * c0[axis];
                double val = 0.0, wsum = 0.0;
                for (int k = k0; k < k0 + 2 && k < nmax; ++k) {
                    int s[3] = {x, y, z};
                    s[axis] = k;
                    const double wt = h ? h[k] : 1.0;
                    val += wt * AT(src, s[1], s[0], s[2], ds);
                    wsum += wt;
                }
                AT(dst, y, x, z, dd) = val / wsum;
            }
}

/* Face-fraction restriction along one axis: `main` axes (the face normal)
 * inject the face at 2X (the last coarse face takes the last fine face,
 * cuRestrict3D_p1_main_* / p2_main_*), other axes average the children
 * (cuRestrict3D_p1_* / p2_lesser_*).  ds/dd are face-array extents. */
static void restrict_face_axis(double* dst, const double* src, i3 ds, int axis, int main_axis) {
    i3 dd = ds;
    const int ns = axis == 0 ? ds.x : axis == 1 ? ds.y : ds.z;
    const int nd = main_axis ? (ns >> 1) + 1 : half_up(ns);
    if (axis == 0) dd.x = nd;
    if (axis == 1) dd.y = nd;
    if (axis == 2) dd.z = nd;
    for (int z = 0; z < dd.z; ++z)
        for (int x = 0; x < dd.x; ++x)
            for (int y = 0; y < dd.y; ++y) {
                const int c0[3] = {x, y, z};
                double val;
                if (main_axis) {
                    int s[3] = {x, y, z};
                    s[axis] = 2 * c0[axis] < ns - 1 ? 2 * c0[axis] : ns - 1;
                    val = AT(src, s[1], s[0], s[2], ds);
                } else {
                    double sum = 0.0;
                    int cnt = 0;
                    for (int k = 2 * c0[axis]; k < 2 * c0[axis] + 2 && k < ns; ++k) {
                        int s[3] = {x, y, z};
                        s[axis] = k;
                        sum += AT(src, s[1], s[0], s[2], ds);
                        ++cnt;
                    }
                    val = sum / cnt;
                }
                AT(dst, y, x, z, dd) = val;
            }
}

/* SolveP1Ax/Ay/Az: `fine` extents are the face array's; restricts x, y, z in turn. */
static void restrict_faces(double* dst, const double* src, i3 fine, int normal) {
    const size_t n = (size_t)fine.x * fine.y * fine.z;
    double* t1 = (double*)malloc(n * sizeof(double));
    double* t2 = (double*)malloc(n * sizeof(double));
    i3 d = fine;
    restrict_face_axis(t1, src, d, 0, normal == 0);
    d.x = normal == 0 ? (fine.x >> 1) + 1 : half_up(fine.x);
    restrict_face_axis(t2, t1, d, 1, normal == 1);
    d.y = normal == 1 ? (fine.y >> 1) + 1 : half_up(fine.y);
    restrict_face_axis(dst, t2, d, 2, normal == 2);
    free(t1);
    free(t2);
}

/* SolveP1RN with spacings: x, then y, then z restriction of a cell field of
 * extents d (cuRestrict3D.cu:939-970). */
static void restrict_cells(double* dst, const double* src, i3 d, const double* hx,
                           const double* hy, const double* hz) {
    const size_t n = (size_t)d.x * d.y * d.z;
    double* t1 = (double*)malloc(n * sizeof(double));
    double* t2 = (double*)malloc(n * sizeof(double));
    i3 s = d;
    restrict_axis_w(t1, src, s, 0, hx);
    s.x = half_up(d.x);
    restrict_axis_w(t2, t1, s, 1, hy);
    s.y = half_up(d.y);
    restrict_axis_w(dst, t2, s, 2, hz);
    free(t1);
    free(t2);
}

/* SolveAddP1RN: each fine cell of extents d adds its parent's value
 * (cuProlongate3D.cu:353-423, 594-616). */
static void prolong_add(double* fine, const double* coarse, i3 d) {
    const i3 c = {half_up(d.x), half_up(d.y), half_up(d.z)};
    for (int z = 0; z < d.z; ++z)
        for (int x = 0; x < d.x; ++x)
            for (int y = 0; y < d.y; ++y)
                AT(fine, y, x, z, d) += AT(coarse, y / 2, x / 2, z / 2, c);
}

typedef struct {
    int levels;
    i3 dim[32];
    double *f[32], *v[32], *ax[32], *ay[32], *az[32], *hx[32], *hy[32], *hz[32];
} hierarchy;

static void free_hierarchy(hierarchy* H) {
    for (int l = 0; l < H->levels; ++l) {
        free(H->f[l]); free(H->v[l]); free(H->ax[l]); free(H->ay[l]); free(H->az[l]);
        free(H->hx[l]); free(H->hy[l]); free(H->hz[l]);
    }
    H->levels = 0;
}

static int build_hierarchy(hierarchy* H, const double* ax, const double* ay, const double* az,
                           i3 dim, double dx, double dy, double dz) {
    memset(H, 0, sizeof(*H));
    i3 d = dim, last = dim;
    int min_dim = dim.x < dim.y ? dim.x : dim.y;
    if (dim.z < min_dim) min_dim = dim.z;
    while (min_dim > 1 && H->levels < 32) {
        const int l = H->levels;
        const size_t n = (size_t)d.x * d.y * d.z;
        H->dim[l] = d;
        H->f[l] = (double*)calloc(n, sizeof(double));
        H->v[l] = (double*)calloc(n, sizeof(double));
        H->ax[l] = (double*)malloc((size_t)(d.x + 1) * d.y * d.z * sizeof(double));
        H->ay[l] = (double*)malloc((size_t)d.x * (d.y + 1) * d.z * sizeof(double));
        H->az[l] = (double*)malloc((size_t)d.x * d.y * (d.z + 1) * sizeof(double));
        H->hx[l] = (double*)malloc(d.x * sizeof(double));
        H->hy[l] = (double*)malloc(d.y * sizeof(double));
        H->hz[l] = (double*)malloc(d.z * sizeof(double));
        if (l == 0) {
            memcpy(H->ax[0], ax, (size_t)(d.x + 1) * d.y * d.z * sizeof(double));
            memcpy(H->ay[0], ay, (size_t)d.x * (d.y + 1) * d.z * sizeof(double));
            memcpy(H->az[0], az, (size_t)d.x * d.y * (d.z + 1) * sizeof(double));
            for (int i = 0; i < d.x; ++i) H->hx[0][i] = dx;
            for (int i = 0; i < d.y; ++i) H->hy[0][i] = dy;
            for (int i = 0; i < d.z; ++i) H->hz[0][i] = dz;
        } else {
            restrict_faces(H->ax[l], H->ax[l - 1], (i3){last.x + 1, last.y, last.z}, 0);
            restrict_faces(H->ay[l], H->ay[l - 1], (i3){last.x, last.y + 1, last.z}, 1);
            restrict_faces(H->az[l], H->az[l - 1], (i3){last.x, last.y, last.z + 1}, 2);
            /* cuRestrictDx: coarse spacing = sum of the child spacings */
            for (int i = 0; i < d.x; ++i)
                H->hx[l][i] = H->hx[l - 1][2 * i] + (2 * i + 1 < last.x ? H->hx[l - 1][2 * i + 1] : 0.0);
            for (int i = 0; i < d.y; ++i)
                H->hy[l][i] = H->hy[l - 1][2 * i] + (2 * i + 1 < last.y ? H->hy[l - 1][2 * i + 1] : 0.0);
            for (int i = 0; i < d.z; ++i)
                H->hz[l][i] = H->hz[l - 1][2 * i] + (2 * i + 1 < last.z ? H->hz[l - 1][2 * i + 1] : 0.0);
        }
        H->levels++;
        last = d;
        d = (i3){half_up(d.x), half_up(d.y), half_up(d.z)};
        min_dim = d.x < d.y ? d.x : d.y;
        if (d.z < min_dim) min_dim = d.z;
    }
    return H->levels;
}

static void smooth(hierarchy* H, int l, int n, char bc, const double he[3], double w) {
    const i3 d = H->dim[l];
    const size_t sz = (size_t)d.x * d.y * d.z;
    double* tmp = (double*)malloc(sz * sizeof(double));
    for (int it = 0; it < n; ++it) {
        apply(tmp, H->v[l], H->f[l], H->ax[l], H->ay[l], H->az[l], d, H->hx[l], H->hy[l],
              H->hz[l], he, bc, w, 0);
        memcpy(H->v[l], tmp, sz * sizeof(double));
    }
    free(tmp);
}

static void vcycle(hierarchy* H, int l, int nu1, int nu2, char bc, const double he[3], double w) {
    const i3 d = H->dim[l];
    if (l < H->levels - 1) {
        smooth(H, l, nu1, bc, he, w);
        const size_t n = (size_t)d.x * d.y * d.z;
        double* r = (double*)malloc(n * sizeof(double));
        apply(r, H->v[l], H->f[l], H->ax[l], H->ay[l], H->az[l], d, H->hx[l], H->hy[l], H->hz[l],
              he, bc, w, 1);
        restrict_cells(H->f[l + 1], r, d, H->hx[l], H->hy[l], H->hz[l]);
        free(r);
        const i3 c = H->dim[l + 1];
        memset(H->v[l + 1], 0, (size_t)c.x * c.y * c.z * sizeof(double));
        vcycle(H, l + 1, nu1, nu2, bc, he, w);
        prolong_add(H->v[l], H->v[l + 1], d);
        smooth(H, l, nu2, bc, he, w);
    } else {
        coarsest(H->v[l], H->f[l], H->ax[l], H->ay[l], H->az[l], d, H->hx[l], H->hy[l], H->hz[l],
                 he, bc, w);
    }
}

static int mg_bc(char bc, double dx, double dy, double dz, char* out, double he[3]) {
    if (bc == 'd') {
        *out = 'z';
        he[0] = he[1] = he[2] = 0.0;
    } else if (bc == 'n' || bc == 'z') {
        *out = bc;
        he[0] = dx; he[1] = dy; he[2] = dz;
    } else {
        return ORC_INVALID;
    }
    return ORC_OK;
}

int orc_mgfrac3d_levels(int nx, int ny, int nz, int* dims_out) {
    int n = 0;
    i3 d = {nx, ny, nz};
    int m = nx < ny ? nx : ny;
    if (nz < m) m = nz;
    while (m > 1 && n < 32) {
        if (dims_out) { dims_out[3 * n] = d.x; dims_out[3 * n + 1] = d.y; dims_out[3 * n + 2] = d.z; }
        ++n;
        d = (i3){half_up(d.x), half_up(d.y), half_up(d.z)};
        m = d.x < d.y ? d.x : d.y;
        if (d.z < m) m = d.z;
    }
    return n;
}

int orc_mgfrac3d_solve(double* dst, const double* pre_v, const double* f, const double* ax,
                       const double* ay, const double* az, int nx, int ny, int nz, double dx,
                       double dy, double dz, char bc, int n_iters, int nu1, int nu2,
                       double weight) {
    if (!dst || !f || !ax || !ay || !az || nx < 2 || ny < 2 || nz < 2 || n_iters < 0)
        return ORC_INVALID;
    char b;
    double he[3];
    if (mg_bc(bc, dx, dy, dz, &b, he) != ORC_OK) return ORC_INVALID;
    hierarchy H;
    build_hierarchy(&H, ax, ay, az, (i3){nx, ny, nz}, dx, dy, dz);
    const size_t n = (size_t)nx * ny * nz;
    if (pre_v) memcpy(H.v[0], pre_v, n * sizeof(double));
    memcpy(H.f[0], f, n * sizeof(double));
    for (int i = 0; i < n_iters; ++i) vcycle(&H, 0, nu1, nu2, b, he, weight);
    memcpy(dst, H.v[0], n * sizeof(double));
    free_hierarchy(&H);
    return ORC_OK;
}

int orc_wj_frac3d(double* dst, const double* src, const double* f, const double* ax,
                  const double* ay, const double* az, int nx, int ny, int nz, double dx, double dy,
                  double dz, char bc, double weight, int residual) {
    char b;
    double he[3];
    if (!dst || !src || !f || mg_bc(bc, dx, dy, dz, &b, he) != ORC_OK) return ORC_INVALID;
    double* hx = (double*)malloc(nx * sizeof(double));
    double* hy = (double*)malloc(ny * sizeof(double));
    double* hz = (double*)malloc(nz * sizeof(double));
    for (int i = 0; i < nx; ++i) hx[i] = dx;
    for (int i = 0; i < ny; ++i) hy[i] = dy;
    for (int i = 0; i < nz; ++i) hz[i] = dz;
    apply(dst, src, f, ax, ay, az, (i3){nx, ny, nz}, hx, hy, hz, he, b, weight, residual);
    free(hx); free(hy); free(hz);
    return ORC_OK;
}

double orc_mgfrac3d_rel_residual(const double* v, const double* f, const double* ax,
                                 const double* ay, const double* az, int nx, int ny, int nz,
                                 double dx, double dy, double dz, char bc) {
    const size_t n = (size_t)nx * ny * nz;
    double* r = (double*)malloc(n * sizeof(double));
    if (orc_wj_frac3d(r, v, f, ax, ay, az, nx, ny, nz, dx, dy, dz, bc, 0.0, 1) != ORC_OK) {
        free(r);
        return -1.0;
    }
    double sr = 0.0, sf = 0.0;
    for (size_t i = 0; i < n; ++i) { sr += r[i] * r[i]; sf += f[i] * f[i]; }
    free(r);
    return sqrt(sr) / sqrt(sf);
}

/* ---- test hooks: the transfer operators and the level data on their own ---- */

int orc_mg_restrict(double* dst, const double* src, int nx, int ny, int nz, const double* hx,
                    const double* hy, const double* hz) {
    if (!dst || !src || !hx || !hy || !hz || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    restrict_cells(dst, src, (i3){nx, ny, nz}, hx, hy, hz);
    return ORC_OK;
}

int orc_mg_prolong_add(double* fine, const double* coarse, int nx, int ny, int nz) {
    if (!fine || !coarse || nx < 1 || ny < 1 || nz < 1) return ORC_INVALID;
    prolong_add(fine, coarse, (i3){nx, ny, nz});
    return ORC_OK;
}

int orc_mg_restrict_faces(double* dst, const double* src, int fx, int fy, int fz, int normal) {
    if (!dst || !src || fx < 1 || fy < 1 || fz < 1 || normal < 0 || normal > 2) return ORC_INVALID;
    restrict_faces(dst, src, (i3){fx, fy, fz}, normal);
    return ORC_OK;
}

int orc_mg_level(const double* ax, const double* ay, const double* az, int nx, int ny, int nz,
                 double dx, double dy, double dz, int level, double* ax_out, double* ay_out,
                 double* az_out, double* hx_out, double* hy_out, double* hz_out) {
    if (!ax || !ay || !az || nx < 2 || ny < 2 || nz < 2 || level < 0) return ORC_INVALID;
    hierarchy H;
    build_hierarchy(&H, ax, ay, az, (i3){nx, ny, nz}, dx, dy, dz);
    if (level >= H.levels) {
        free_hierarchy(&H);
        return ORC_INVALID;
    }
    const i3 d = H.dim[level];
    const size_t sx = (size_t)(d.x + 1) * d.y * d.z, sy = (size_t)d.x * (d.y + 1) * d.z,
                 sz = (size_t)d.x * d.y * (d.z + 1);
    if (ax_out) memcpy(ax_out, H.ax[level], sx * sizeof(double));
    if (ay_out) memcpy(ay_out, H.ay[level], sy * sizeof(double));
    if (az_out) memcpy(az_out, H.az[level], sz * sizeof(double));
    if (hx_out) memcpy(hx_out, H.hx[level], d.x * sizeof(double));
    if (hy_out) memcpy(hy_out, H.hy[level], d.y * sizeof(double));
    if (hz_out) memcpy(hz_out, H.hz[level], d.z * sizeof(double));
    free_hierarchy(&H);
    return ORC_OK;
}
