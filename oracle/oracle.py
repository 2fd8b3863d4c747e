"""numpy wrappers over liboracle.so (test infrastructure only; see __init__.py).

Arrays follow the reference layout: a field of extents (nx, ny, nz) is a numpy
array of shape (nz, nx, ny) (y fastest), WaveletPotentials/Utils.h:46-47.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OK, INVALID, UNSUPPORTED, UNDEFINED = 0, -1, -2, -3

# WaveletType order of WaveletPotentials/WaveletTypes.h:7-28
WAVELETS = ["cdf3.9", "cdf3.11", "cdf3.5", "cdf2.10", "cdf2.12", "cdf2.6", "cdf4.4", "cdf4.6",
            "cdf4.8", "cdf4.10", "cdf2.8", "cdf3.7", "cdf26.44", "cdf28.46", "cdf44.26",
            "cdf46.28", "cdf2.0", "cdf1.1", "cdf1.7"]
WT = {name: i for i, name in enumerate(WAVELETS)}


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


_D = ctypes.POINTER(ctypes.c_double)


def _setup(L):
    i, c, d = ctypes.c_int, ctypes.c_char, ctypes.c_double
    L.orc_transform.argtypes = [_D, _D, i, i, i, i, c, c, i, i]
    L.orc_emulate_line.argtypes = [_D, _D, i, i, c, i, i]
    L.orc_gram_table.argtypes = [_D, i, c, i]
    L.orc_project3d.argtypes = [_D] * 9 + [i, i, i, d, d, d, i, c]
    L.orc_project2d.argtypes = [_D, _D, _D, i, i, d, d, i, c]
    L.orc_recover3d.argtypes = [_D] * 9 + [i, i, i, d, d, d, c, i, i]
    L.orc_recover2d.argtypes = [_D, _D, _D, i, i, d, d, c, i, i]
    L.orc_cut3d.argtypes = [_D] * 6 + [i, i, i, c, i, i, i]
    L.orc_curl3d.argtypes = [_D] * 6 + [i, i, i, d, d, d]
    L.orc_curl2d.argtypes = [_D] * 3 + [i, i, d, d]
    L.orc_div3d.argtypes = [_D] * 4 + [i, i, i, d, d, d]
    L.orc_levelset_spheres.argtypes = [_D, i, i, i, d, d, d, _D, _D, i]
    L.orc_frac_from_ls.argtypes = [_D] * 4 + [i, i, i, d, d, d]
    L.orc_eikonal_update.argtypes = [d] * 6
    L.orc_eikonal_update.restype = d
    L.orc_eikonal_sweep.argtypes = [_D, i, i, i, d, i]
    L.orc_mgfrac3d_levels.argtypes = [i, i, i, ctypes.POINTER(ctypes.c_int)]
    L.orc_mgfrac3d_solve.argtypes = [_D] * 6 + [i, i, i, d, d, d, c, i, i, i, d]
    L.orc_wj_frac3d.argtypes = [_D] * 6 + [i, i, i, d, d, d, c, d, i]
    L.orc_mgfrac3d_rel_residual.argtypes = [_D] * 5 + [i, i, i, d, d, d, c]
    L.orc_mgfrac3d_rel_residual.restype = d
    L.orc_gradient3d_frac.argtypes = [_D] * 7 + [i, i, i, d, d, d, c]
    L.orc_divergence3d_frac.argtypes = [_D] * 7 + [i, i, i, d, d, d]
    L.orc_sweep3d.argtypes = [_D] * 5 + [i, i, i, d, d, d]
    L.orc_dst3d_solve.argtypes = [_D, _D, i, i, i, d, d, d]
    L.orc_project_zero.argtypes = [_D] * 3 + [i, i, i, d, d, d]
    L.orc_curl3d.restype = None
    L.orc_curl2d.restype = None
    L.orc_div3d.restype = None
    L.orc_wavelet_supported.argtypes = [i, c]
    L.orc_wavelet_halo.argtypes = [i]
    L.orc_proj_wavelet.argtypes = [i]


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def _wt(w):
    return WT[w] if isinstance(w, str) else int(w)


def _b(ch):
    return ch.encode() if isinstance(ch, str) else ch


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc != OK:
        raise OracleError(f"{what}: oracle status {rc}")


def field(nx, ny, nz=1):
    return np.zeros((nz, nx, ny), dtype=np.float64)


def transform(src, levels, axis, bc, wt, forward=True, allow_undefined=False):
    """Solve1D on a (nz, nx, ny) float64 array; returns a new array."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    nz, nx, ny = src.shape
    dst = np.empty_like(src)
    rc = lib().orc_transform(_p(dst), _p(src), nx, ny, nz, levels, _b(axis), _b(bc), _wt(wt),
                             int(forward))
    if not (allow_undefined and rc == UNDEFINED):
        _check(rc, "transform")
    return dst, rc


def emulate_line(src, levels, bc, wt, forward=True):
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.empty_like(src)
    rc = lib().orc_emulate_line(_p(dst), _p(src), src.size, levels, _b(bc), _wt(wt),
                                int(forward))
    return dst, rc


def gram_table(wt, bc, n):
    out = np.zeros(8 * n, dtype=np.float64)
    _check(lib().orc_gram_table(_p(out), _wt(wt), _b(bc), n), "gram_table")
    return out.reshape(8, n)


def project3d(xv, yv, zv, n, dx, wt_div="cdf4.6", bc_div="d", with_vout=False):
    nx, ny, nz = n
    qx, qy, qz = field(nx, ny + 1, nz + 1), field(nx + 1, ny, nz + 1), field(nx + 1, ny + 1, nz)
    xo = yo = zo = None
    if with_vout:
        xo, yo, zo = xv.copy(), yv.copy(), zv.copy()
    _check(lib().orc_project3d(_p(qx), _p(qy), _p(qz), _p(xo), _p(yo), _p(zo), _p(xv), _p(yv),
                               _p(zv), nx, ny, nz, dx[0], dx[1], dx[2], _wt(wt_div), _b(bc_div)),
           "project3d")
    return (qx, qy, qz, xo, yo, zo) if with_vout else (qx, qy, qz)


def project2d(xv, yv, n, dx, wt_div="cdf4.6", bc_div="d"):
    nx, ny = n
    qz = field(nx + 1, ny + 1)
    _check(lib().orc_project2d(_p(qz), _p(xv), _p(yv), nx, ny, dx[0], dx[1], _wt(wt_div),
                               _b(bc_div)), "project2d")
    return qz


def recover3d(xv, yv, zv, n, dx, bc="n", wt_curl="cdf3.7", wt_div="cdf4.6", with_vout=False):
    nx, ny, nz = n
    qx, qy, qz = field(nx, ny + 1, nz + 1), field(nx + 1, ny, nz + 1), field(nx + 1, ny + 1, nz)
    xo = yo = zo = None
    if with_vout:
        xo, yo, zo = np.empty_like(xv), np.empty_like(yv), np.empty_like(zv)
    _check(lib().orc_recover3d(_p(qx), _p(qy), _p(qz), _p(xo), _p(yo), _p(zo), _p(xv), _p(yv),
                               _p(zv), nx, ny, nz, dx[0], dx[1], dx[2], _b(bc), _wt(wt_curl),
                               _wt(wt_div)), "recover3d")
    return (qx, qy, qz, xo, yo, zo) if with_vout else (qx, qy, qz)


def recover2d(xv, yv, n, dx, bc="n", wt_curl="cdf3.7", wt_div="cdf4.6"):
    nx, ny = n
    qz = field(nx + 1, ny + 1)
    _check(lib().orc_recover2d(_p(qz), _p(xv), _p(yv), nx, ny, dx[0], dx[1], _b(bc),
                               _wt(wt_curl), _wt(wt_div)), "recover2d")
    return qz


def cut3d(qx, qy, qz, n, level, bc="n", wt_curl="cdf3.7", wt_div="cdf4.6"):
    nx, ny, nz = n
    ox, oy, oz = np.empty_like(qx), np.empty_like(qy), np.empty_like(qz)
    _check(lib().orc_cut3d(_p(ox), _p(oy), _p(oz), _p(qx), _p(qy), _p(qz), nx, ny, nz, _b(bc),
                           _wt(wt_curl), _wt(wt_div), level), "cut3d")
    return ox, oy, oz


def curl3d(qx, qy, qz, n, dx):
    nx, ny, nz = n
    xv, yv, zv = field(nx + 1, ny, nz), field(nx, ny + 1, nz), field(nx, ny, nz + 1)
    lib().orc_curl3d(_p(xv), _p(yv), _p(zv), _p(np.ascontiguousarray(qx)),
                     _p(np.ascontiguousarray(qy)), _p(np.ascontiguousarray(qz)), nx, ny, nz,
                     dx[0], dx[1], dx[2])
    return xv, yv, zv


def curl2d(qz, n, dx):
    nx, ny = n
    xv, yv = field(nx + 1, ny), field(nx, ny + 1)
    lib().orc_curl2d(_p(xv), _p(yv), _p(np.ascontiguousarray(qz)), nx, ny, dx[0], dx[1])
    return xv, yv


def div3d(xv, yv, zv, n, dx):
    nx, ny, nz = n
    d = field(nx, ny, nz)
    lib().orc_div3d(_p(d), _p(xv), _p(yv), _p(zv), nx, ny, nz, dx[0], dx[1], dx[2])
    return d


def wavelet_supported(wt, bc):
    return bool(lib().orc_wavelet_supported(_wt(wt), _b(bc)))


def halo(wt):
    return lib().orc_wavelet_halo(_wt(wt))


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in (a if isinstance(a, (list, tuple)) else [a])])
    b = np.concatenate([np.ravel(x) for x in (b if isinstance(b, (list, tuple)) else [b])])
    return float(np.sqrt(np.sum((a - b) ** 2) / (np.sum(a ** 2) + 1e-300)))


# ---- solids ----------------------------------------------------------------
def levelset_spheres(n, dx, centers, radii):
    """Union-of-spheres signed distance at cell centres; returns (nz, nx, ny)."""
    nx, ny, nz = n
    c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
    ls = field(nx, ny, nz)
    _check(lib().orc_levelset_spheres(_p(ls), nx, ny, nz, *map(float, dx), _p(c), _p(r), len(r)),
           "levelset_spheres")
    return ls


def frac_from_ls(ls, dx):
    """Face fractions (ax, ay, az) from a cell level set of shape (nz, nx, ny)."""
    ls = np.ascontiguousarray(ls, dtype=np.float64)
    nz, nx, ny = ls.shape
    ax, ay, az = field(nx + 1, ny, nz), field(nx, ny + 1, nz), field(nx, ny, nz + 1)
    _check(lib().orc_frac_from_ls(_p(ax), _p(ay), _p(az), _p(ls), nx, ny, nz, *map(float, dx)),
           "frac_from_ls")
    return ax, ay, az


def eikonal_update(mx, my, mz, dx, dy, dz):
    return lib().orc_eikonal_update(mx, my, mz, dx, dy, dz)


def eikonal_sweep(ls, h, n_iters):
    """n_iters Jacobi relaxations of the positive region; returns a new array."""
    out = np.array(ls, dtype=np.float64, order="C", copy=True)
    nz, nx, ny = out.shape
    _check(lib().orc_eikonal_sweep(_p(out), nx, ny, nz, float(h), int(n_iters)), "eikonal_sweep")
    return out


# ---- fractional multigrid ---------------------------------------------------
def mg_levels(n):
    nx, ny, nz = n
    out = (ctypes.c_int * 96)()
    k = lib().orc_mgfrac3d_levels(nx, ny, nz, out)
    return [tuple(out[3 * i:3 * i + 3]) for i in range(k)]


def _fr(ax, ay, az):
    return [np.ascontiguousarray(a, dtype=np.float64) for a in (ax, ay, az)]


def mg_solve(f, ax, ay, az, dx, bc="n", n_iters=10, nu1=3, nu2=3, weight=0.8, pre_v=None):
    """CuMultigridFrac3D(ax, ay, az, dim, dx, bc).Solve(dst, pre_v, f, ...); f is (nz, nx, ny)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    nz, nx, ny = f.shape
    ax, ay, az = _fr(ax, ay, az)
    dst = np.empty_like(f)
    pv = None if pre_v is None else np.ascontiguousarray(pre_v, dtype=np.float64)
    _check(lib().orc_mgfrac3d_solve(_p(dst), _p(pv) if pv is not None else None, _p(f), _p(ax),
                                    _p(ay), _p(az), nx, ny, nz, *map(float, dx), _b(bc),
                                    int(n_iters), int(nu1), int(nu2), float(weight)), "mg_solve")
    return dst


def wj_sweep(v, f, ax, ay, az, dx, bc="n", weight=0.8, residual=False):
    v = np.ascontiguousarray(v, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    nz, nx, ny = v.shape
    ax, ay, az = _fr(ax, ay, az)
    out = np.empty_like(v)
    _check(lib().orc_wj_frac3d(_p(out), _p(v), _p(f), _p(ax), _p(ay), _p(az), nx, ny, nz,
                               *map(float, dx), _b(bc), float(weight), int(residual)), "wj_sweep")
    return out


def mg_rel_residual(v, f, ax, ay, az, dx, bc="n"):
    v = np.ascontiguousarray(v, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    nz, nx, ny = v.shape
    ax, ay, az = _fr(ax, ay, az)
    return lib().orc_mgfrac3d_rel_residual(_p(v), _p(f), _p(ax), _p(ay), _p(az), nx, ny, nz,
                                           *map(float, dx), _b(bc))


# ---- field operators --------------------------------------------------------
def gradient3d_frac(p, ax, ay, az, dx, bc="n"):
    p = np.ascontiguousarray(p, dtype=np.float64)
    nz, nx, ny = p.shape
    ax, ay, az = _fr(ax, ay, az)
    xv, yv, zv = field(nx + 1, ny, nz), field(nx, ny + 1, nz), field(nx, ny, nz + 1)
    _check(lib().orc_gradient3d_frac(_p(xv), _p(yv), _p(zv), _p(p), _p(ax), _p(ay), _p(az), nx, ny,
                                     nz, *map(float, dx), _b(bc)), "gradient3d_frac")
    return xv, yv, zv


def divergence3d_frac(xv, yv, zv, ax, ay, az, dx):
    xv, yv, zv = _fr(xv, yv, zv)
    ax, ay, az = _fr(ax, ay, az)
    nz, nx1, ny = xv.shape
    nx = nx1 - 1
    b = field(nx, ny, nz)
    _check(lib().orc_divergence3d_frac(_p(b), _p(xv), _p(yv), _p(zv), _p(ax), _p(ay), _p(az), nx,
                                       ny, nz, *map(float, dx)), "divergence3d_frac")
    return b


# ---- sweep + DST comparator ------------------------------------------------
def sweep3d(xv, yv, n, dx):
    """CuParallelSweep3D::Solve (bc 'n'): returns (qx, qy, qz)."""
    nx, ny, nz = n
    xv = np.ascontiguousarray(xv, dtype=np.float64)
    yv = np.ascontiguousarray(yv, dtype=np.float64)
    qx, qy, qz = field(nx, ny + 1, nz + 1), field(nx + 1, ny, nz + 1), field(nx + 1, ny + 1, nz)
    _check(lib().orc_sweep3d(_p(qx), _p(qy), _p(qz), _p(xv), _p(yv), nx, ny, nz, *map(float, dx)),
           "sweep3d")
    return qx, qy, qz


def dst3d_solve(b, dx):
    """(-L_h) phi = b with zero values one node outside; b is (n3, n1, n2)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    n3, n1, n2 = b.shape
    phi = np.empty_like(b)
    _check(lib().orc_dst3d_solve(_p(phi), _p(b), n1, n2, n3, *map(float, dx)), "dst3d_solve")
    return phi


def project_zero(qx, qy, qz, n, dx):
    """CuParallelSweep3D::ProjectZero; returns new arrays."""
    nx, ny, nz = n
    q = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (qx, qy, qz)]
    _check(lib().orc_project_zero(_p(q[0]), _p(q[1]), _p(q[2]), nx, ny, nz, *map(float, dx)),
           "project_zero")
    return tuple(q)
