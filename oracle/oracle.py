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
    L.orc_frac_divide.argtypes = [_D] * 4 + [i, i, i, d, d, d]
    L.orc_interp_linear3d.argtypes = [_D, d, d, d, i, i, i]
    L.orc_interp_linear3d.restype = d
    L.orc_interp_quadratic3d.argtypes = [_D, d, d, d, i, i, i]
    L.orc_interp_quadratic3d.restype = d
    L.orc_eikonal_update.argtypes = [d] * 6
    L.orc_eikonal_update.restype = d
    L.orc_eikonal_sweep.argtypes = [_D, i, i, i, d, i]
    L.orc_mgfrac3d_levels.argtypes = [i, i, i, ctypes.POINTER(ctypes.c_int)]
    L.orc_mgfrac3d_solve.argtypes = [_D] * 6 + [i, i, i, d, d, d, c, i, i, i, d]
    L.orc_wj_frac3d.argtypes = [_D] * 6 + [i, i, i, d, d, d, c, d, i]
    L.orc_mgfrac3d_rel_residual.argtypes = [_D] * 5 + [i, i, i, d, d, d, c]
    L.orc_mgfrac3d_rel_residual.restype = d
    L.orc_mg_restrict.argtypes = [_D, _D, i, i, i, _D, _D, _D]
    L.orc_mg_prolong_add.argtypes = [_D, _D, i, i, i]
    L.orc_mg_restrict_faces.argtypes = [_D, _D, i, i, i, i]
    L.orc_mg_level.argtypes = [_D, _D, _D, i, i, i, d, d, d, i] + [_D] * 6
    L.orc_gradient3d_frac.argtypes = [_D] * 7 + [i, i, i, d, d, d, c]
    L.orc_divergence3d_frac.argtypes = [_D] * 7 + [i, i, i, d, d, d]
    L.orc_sweep3d.argtypes = [_D] * 5 + [i, i, i, d, d, d]
    L.orc_dst3d_solve.argtypes = [_D, _D, i, i, i, d, d, d]
    L.orc_dst2d_solve.argtypes = [_D, _D, i, i, d, d]
    L.orc_sweep3d_d.argtypes = [_D] * 6 + [i, i, i, d, d, d]
    L.orc_project_d.argtypes = [_D] * 3 + [i, i, i, d, d, d]
    L.orc_sweep2d.argtypes = [_D] * 3 + [i, i, d, d, c]
    L.orc_project_zero.argtypes = [_D] * 3 + [i, i, i, d, d, d]
    L.orc_conv3.argtypes = [_D, _D, i, _D, i, i, i, i, i, _D, _D, i]
    L.orc_conv3_bwd_data.argtypes = [_D, _D, i, i, i, i, _D, i]
    L.orc_conv3_bwd_weight.argtypes = [_D, _D, _D, i, _D, i, _D, i, i, i, i, i]
    L.orc_curl3d.restype = None
    L.orc_curl2d.restype = None
    L.orc_div3d.restype = None
    L.orc_div_l0_matrix.argtypes = [_D, i, i]
    L.orc_div_l0_matrix.restype = None
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


def div_l0_matrix(lx, ly):
    """GetDivL0Matrix for level counts (lx, ly): 16 entries."""
    out = np.zeros(16)
    lib().orc_div_l0_matrix(_p(out), int(lx), int(ly))
    return out


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


_lib32 = None
_F = ctypes.POINTER(ctypes.c_float)


def lib_f32():
    """liboracle_f32.so: the FP32 build of the recovery path (orc_fp32.c), for timing."""
    global _lib32
    if _lib32 is None:
        path = os.path.join(_HERE, "liboracle_f32.so")
        if not os.path.exists(path):
            build()
        _lib32 = ctypes.CDLL(path)
        i, c, f = ctypes.c_int, ctypes.c_char, ctypes.c_float
        _lib32.orc_recover3d.argtypes = [_F] * 9 + [i, i, i, f, f, f, c, i, i]
    return _lib32


def recover3d_f32(xv, yv, zv, n, dx, bc="n", wt_curl="cdf3.7", wt_div="cdf4.6"):
    """recover3d computed in FP32 by liboracle_f32.so; float32 inputs and outputs."""
    nx, ny, nz = n

    def pf(a):
        assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
        return a.ctypes.data_as(_F)

    q = [np.zeros(s, dtype=np.float32) for s in ((nz + 1, nx, ny + 1), (nz + 1, nx + 1, ny),
                                                  (nz, nx + 1, ny + 1))]
    ins = [np.ascontiguousarray(a, dtype=np.float32) for a in (xv, yv, zv)]
    _check(lib_f32().orc_recover3d(pf(q[0]), pf(q[1]), pf(q[2]), None, None, None,
                                   *(pf(a) for a in ins), nx, ny, nz, dx[0], dx[1], dx[2],
                                   _b(bc), _wt(wt_curl), _wt(wt_div)), "recover3d (FP32)")
    return tuple(q)


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


def frac_divide(ls, dx):
    """Supersampled face fractions (the reference's static GetFrac) of a cell
    level set of shape (nz, nx, ny)."""
    ls = np.ascontiguousarray(ls, dtype=np.float64)
    nz, nx, ny = ls.shape
    ax, ay, az = field(nx + 1, ny, nz), field(nx, ny + 1, nz), field(nx, ny, nz + 1)
    _check(lib().orc_frac_divide(_p(ax), _p(ay), _p(az), _p(ls), nx, ny, nz, *map(float, dx)),
           "frac_divide")
    return ax, ay, az


def interp_linear3d(v, g):
    """InterpolateLinear3D of v (nz, nx, ny) at grid coordinates g = (x, y, z)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    nz, nx, ny = v.shape
    return lib().orc_interp_linear3d(_p(v), *map(float, g), nx, ny, nz)


def interp_quadratic3d(v, g):
    """InterpolateQuadratic3D of v (nz, nx, ny) at grid coordinates g = (x, y, z)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    nz, nx, ny = v.shape
    return lib().orc_interp_quadratic3d(_p(v), *map(float, g), nx, ny, nz)


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


def _half_up(n):
    return (n + 1) // 2


def mg_restrict(src, hx, hy, hz):
    """Spacing-weighted restriction of a cell field (nz, nx, ny) to ceil-halved
    extents (CuRestrict3D::SolveP1RN with dx, cuRestrict3D.cu:801-970)."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    nz, nx, ny = src.shape
    h = [np.ascontiguousarray(a, dtype=np.float64) for a in (hx, hy, hz)]
    assert [len(a) for a in h] == [nx, ny, nz]
    dst = np.empty((_half_up(nz), _half_up(nx), _half_up(ny)))
    _check(lib().orc_mg_restrict(_p(dst), _p(src), nx, ny, nz, *map(_p, h)), "mg_restrict")
    return dst


def mg_prolong_add(fine, coarse):
    """fine + (each fine cell's parent value), CuProlongate3D::SolveAddP1RN."""
    out = np.array(fine, dtype=np.float64, order="C")
    nz, nx, ny = out.shape
    coarse = np.ascontiguousarray(coarse, dtype=np.float64)
    assert coarse.shape == (_half_up(nz), _half_up(nx), _half_up(ny))
    _check(lib().orc_mg_prolong_add(_p(out), _p(coarse), nx, ny, nz), "mg_prolong_add")
    return out


def mg_restrict_faces(src, normal):
    """Face-fraction restriction (SolveP1Ax/Ay/Az, cuRestrict3D.cu:972-1140) of a
    face array (fz, fx, fy) whose normal axis is 0 (x), 1 (y) or 2 (z)."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    fz, fx, fy = src.shape
    ext = [fx, fy, fz]
    out = [(e >> 1) + 1 if a == normal else _half_up(e) for a, e in enumerate(ext)]
    dst = np.empty((out[2], out[0], out[1]))
    _check(lib().orc_mg_restrict_faces(_p(dst), _p(src), fx, fy, fz, int(normal)),
           "mg_restrict_faces")
    return dst


def mg_level(ax, ay, az, dx, level):
    """Face fractions and per-axis cell spacings of one level of the hierarchy the
    MG builds from fine fractions (CuMultigridFrac3D ctor, cuMultigridFrac3D.cu:9-134)."""
    ax, ay, az = _fr(ax, ay, az)
    nz, nx, ny = ax.shape[0], ax.shape[1] - 1, ax.shape[2]
    levels = mg_levels((nx, ny, nz))
    if not 0 <= level < len(levels):
        raise ValueError(f"level {level} outside 0..{len(levels) - 1}")
    cx, cy, cz = levels[level]
    out = [np.empty((cz, cx + 1, cy)), np.empty((cz, cx, cy + 1)), np.empty((cz + 1, cx, cy)),
           np.empty(cx), np.empty(cy), np.empty(cz)]
    _check(lib().orc_mg_level(_p(ax), _p(ay), _p(az), nx, ny, nz, *map(float, dx), int(level),
                              *map(_p, out)), "mg_level")
    return out


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


def sweep3d_d(xv, yv, zv, n, dx):
    """CuParallelSweep3D::Solve for velocity bc 'd' ('z'): returns (qx, qy, qz)."""
    nx, ny, nz = n
    v = [np.ascontiguousarray(a, dtype=np.float64) for a in (xv, yv, zv)]
    qx, qy, qz = field(nx, ny + 1, nz + 1), field(nx + 1, ny, nz + 1), field(nx + 1, ny + 1, nz)
    _check(lib().orc_sweep3d_d(_p(qx), _p(qy), _p(qz), *map(_p, v), nx, ny, nz, *map(float, dx)),
           "sweep3d_d")
    return qx, qy, qz


def sweep2d(xv, yv, n, dx, bc="n"):
    """CuParallelSweep3D::Solve(qz, xv, yv): the 2D sweep; xv (nx+1, ny), yv (nx, ny+1)."""
    nx, ny = n
    qz = np.zeros((nx + 1, ny + 1))
    xv = np.ascontiguousarray(xv, dtype=np.float64)
    yv = np.ascontiguousarray(yv, dtype=np.float64)
    _check(lib().orc_sweep2d(_p(qz), _p(xv), _p(yv), nx, ny, float(dx[0]), float(dx[1]), _b(bc)),
           "sweep2d")
    return qz


def project_d(qx, qy, qz, n, dx):
    """CuParallelSweep3D::ProjectD; returns new arrays."""
    nx, ny, nz = n
    q = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (qx, qy, qz)]
    _check(lib().orc_project_d(_p(q[0]), _p(q[1]), _p(q[2]), nx, ny, nz, *map(float, dx)),
           "project_d")
    return tuple(q)


def dst3d_solve(b, dx):
    """(-L_h) phi = b with zero values one node outside; b is (n3, n1, n2)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    n3, n1, n2 = b.shape
    phi = np.empty_like(b)
    _check(lib().orc_dst3d_solve(_p(phi), _p(b), n1, n2, n3, *map(float, dx)), "dst3d_solve")
    return phi


def dst2d_solve(b, dx):
    """(-L_h) phi = b on an (n1, n2) node grid, zero one node outside."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    n1, n2 = b.shape
    phi = np.empty_like(b)
    _check(lib().orc_dst2d_solve(_p(phi), _p(b), n1, n2, *map(float, dx)), "dst2d_solve")
    return phi


def project_zero(qx, qy, qz, n, dx):
    """CuParallelSweep3D::ProjectZero; returns new arrays."""
    nx, ny, nz = n
    q = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (qx, qy, qz)]
    _check(lib().orc_project_zero(_p(q[0]), _p(q[1]), _p(q[2]), nx, ny, nz, *map(float, dx)),
           "project_zero")
    return tuple(q)


# ---- learned initial guess: the 3-level U-Net (unet_oracle.c + composition) ----
# (channels from a, channels from b, a upsampled, output channels, level); the
# last layer has 16 output rows of which only row 0 is nonzero.
UNET_LAYERS = [(16, 0, False, 16, 0), (16, 0, False, 16, 0), (16, 0, False, 32, 1),
               (32, 0, False, 32, 1), (32, 0, False, 64, 2), (64, 0, False, 64, 2),
               (64, 32, True, 32, 1), (32, 0, False, 32, 1), (32, 16, True, 16, 0),
               (16, 0, False, 16, 0), (16, 0, False, 16, 0)]


def unet_layout():
    """[(weight offset, bias offset)] per layer in the flat parameter vector
    (weights [co][27][cin] then biases [co]) and its length."""
    off, out = 0, []
    for ca, cb, _, co, _ in UNET_LAYERS:
        wo = off
        off += co * 27 * (ca + cb)
        out.append((wo, off))
        off += co
    return out, off


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), via FP32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(a))


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def conv3(a, ca, b, cb, up, dims, w, bias, co):
    nz, nx, ny = dims
    y = np.empty((nz * nx * ny, co))
    a, b, w, bias = _c(a), _c(b), _c(w), _c(bias)
    _check(lib().orc_conv3(_p(y), _p(a), ca, _p(b), cb, int(up), nz, nx, ny, _p(w), _p(bias), co),
           "conv3")
    return y


def conv3_bwd_data(g, co, dims, w, cin):
    nz, nx, ny = dims
    dx = np.empty((nz * nx * ny, cin))
    g, w = _c(g), _c(w)
    _check(lib().orc_conv3_bwd_data(_p(dx), _p(g), co, nz, nx, ny, _p(w), cin), "conv3_bwd_data")
    return dx


def conv3_bwd_weight(g, co, a, ca, b, cb, up, dims):
    nz, nx, ny = dims
    dw = np.empty(co * 27 * (ca + cb))
    db = np.empty(co)
    g, a, b = _c(g), _c(a), _c(b)
    _check(lib().orc_conv3_bwd_weight(_p(dw), _p(db), _p(g), co, _p(a), ca, _p(b), cb, int(up),
                                      nz, nx, ny), "conv3_bwd_weight")
    return dw, db


def _pool(x, dims, C, mean=True):
    nz, nx, ny = dims
    t = x.reshape(nz // 2, 2, nx // 2, 2, ny // 2, 2, C).sum(axis=(1, 3, 5))
    return (t / 8.0 if mean else t).reshape(-1, C)


def _unpool(xc, dims, C):
    nz, nx, ny = dims
    t = xc.reshape(nz // 2, nx // 2, ny // 2, C)
    return np.repeat(np.repeat(np.repeat(t, 2, 0), 2, 1), 2, 2).reshape(-1, C)


def unet_forward(params, f, ls, length, bf16=False):
    """Network output (scaled units, one value per cell) and the activations.
    f, ls: cell arrays (nz, nx, ny).  With bf16, activations are rounded to
    BF16 where the GPU stores them."""
    R = bf16_round if bf16 else (lambda a: a)
    lay, _ = unet_layout()
    nz, nx, ny = f.shape
    d0, d1, d2 = (nz, nx, ny), (nz // 2, nx // 2, ny // 2), (nz // 4, nx // 4, ny // 4)
    ms = float(np.mean(np.asarray(f, dtype=np.float64) ** 2))
    x = np.zeros((nz * nx * ny, 16))
    x[:, 0] = np.ravel(f) / np.sqrt(ms)
    x[:, 1] = np.ravel(ls) / length
    A = {"x": R(x)}

    def layer(l, a, b, dims, relu=True):
        ca, cb, up, co, _ = UNET_LAYERS[l]
        wo, bo = lay[l]
        y = conv3(a, ca, b, cb, up, dims, params[wo:bo], params[bo:bo + co], co)
        return R(np.maximum(y, 0.0)) if relu else y

    A["e0a"] = layer(0, A["x"], None, d0)
    A["e0"] = layer(1, A["e0a"], None, d0)
    A["p0"] = R(_pool(A["e0"], d0, 16))
    A["e1a"] = layer(2, A["p0"], None, d1)
    A["e1"] = layer(3, A["e1a"], None, d1)
    A["p1"] = R(_pool(A["e1"], d1, 32))
    A["ba"] = layer(4, A["p1"], None, d2)
    A["b"] = layer(5, A["ba"], None, d2)
    A["d1a"] = layer(6, A["b"], A["e1"], d1)
    A["d1"] = layer(7, A["d1a"], None, d1)
    A["d0a"] = layer(8, A["d1"], A["e0"], d0)
    A["d0"] = layer(9, A["d0a"], None, d0)
    out = layer(10, A["d0"], None, d0, relu=False)[:, 0]
    return out, A, ms


def unet_down(a, F):
    """Means of the F^3 blocks of a (nz, nx, ny) cell array."""
    nz, nx, ny = a.shape
    return a.reshape(nz // F, F, nx // F, F, ny // F, F).mean(axis=(1, 3, 5))


def unet_up(c, F):
    """Trilinear interpolation of cell-centred values on the coarse grid at the
    fine cell centres (coarse coordinate (i + 1/2)/F - 1/2, clamped)."""
    def axis(nc):
        u = np.clip((np.arange(nc * F) + 0.5) / F - 0.5, 0, nc - 1)
        i0 = np.floor(u).astype(int)
        i1 = np.minimum(i0 + 1, nc - 1)
        return i0, i1, u - i0
    out = np.asarray(c, dtype=np.float64)
    for ax in range(3):
        i0, i1, w = axis(out.shape[ax])
        shape = [1, 1, 1]
        shape[ax] = -1
        out = (np.take(out, i0, axis=ax) * (1 - w).reshape(shape)
               + np.take(out, i1, axis=ax) * w.reshape(shape))
    return out


def unet_predict(params, f, ls, length, bf16=False, out_scale=1.0, factor=1):
    """The initial guess p0 = output * rms(f) * length^2 * out_scale, shaped like
    f; with factor > 1 the network runs on the block means of f and ls and its
    output is interpolated back."""
    if factor > 1:
        fc, lc = unet_down(f, factor), unet_down(ls, factor)
        out, _, ms = unet_forward(params, fc, lc, length, bf16)
        return unet_up(out.reshape(fc.shape), factor) * np.sqrt(ms) * length ** 2 * out_scale
    out, _, ms = unet_forward(params, f, ls, length, bf16)
    return (out * np.sqrt(ms) * length ** 2 * out_scale).reshape(f.shape)


def unet_loss_grad(params, f, ls, p, length, bf16=False, out_scale=1.0):
    """loss = mean((e - mean e)^2), e = output - p / (rms(f) length^2 out_scale),
    and its gradient with respect to the flat parameters."""
    R = bf16_round if bf16 else (lambda a: a)
    lay, n = unet_layout()
    out, A, ms = unet_forward(params, f, ls, length, bf16)
    nz, nx, ny = f.shape
    d0, d1, d2 = (nz, nx, ny), (nz // 2, nx // 2, ny // 2), (nz // 4, nx // 4, ny // 4)
    V0 = out.size
    e = out - np.ravel(p) / (np.sqrt(ms) * length ** 2 * out_scale)
    eb = float(e.mean())
    loss = float(np.mean(e * e) - eb * eb)
    g = np.zeros((V0, 16))
    g[:, 0] = 2.0 * (e - eb) / V0
    g = R(g)
    grads = np.zeros(n)

    def wgrad(l, G, a, b, dims):
        ca, cb, up, co, _ = UNET_LAYERS[l]
        wo, bo = lay[l]
        grads[wo:bo], grads[bo:bo + co] = conv3_bwd_weight(G, co, a, ca, b, cb, up, dims)

    def bdata(l, G, dims):
        ca, cb, _, co, _ = UNET_LAYERS[l]
        wo, bo = lay[l]
        return R(conv3_bwd_data(G, co, dims, params[wo:bo], ca + cb))

    def mask(G, Y):
        return G * (Y > 0)

    wgrad(10, g, A["d0"], None, d0)
    g0b = mask(bdata(10, g, d0), A["d0"])
    wgrad(9, g0b, A["d0a"], None, d0)
    g0a = mask(bdata(9, g0b, d0), A["d0a"])
    wgrad(8, g0a, A["d1"], A["e0"], d0)
    t = bdata(8, g0a, d0)
    u0, ge0 = t[:, :32], t[:, 32:]
    g1a = mask(R(_pool(u0, d0, 32, mean=False)), A["d1"])
    wgrad(7, g1a, A["d1a"], None, d1)
    g1b = mask(bdata(7, g1a, d1), A["d1a"])
    wgrad(6, g1b, A["b"], A["e1"], d1)
    t = bdata(6, g1b, d1)
    u1, ge1 = t[:, :64], t[:, 64:]
    g2a = mask(R(_pool(u1, d1, 64, mean=False)), A["b"])
    wgrad(5, g2a, A["ba"], None, d2)
    g2b = mask(bdata(5, g2a, d2), A["ba"])
    wgrad(4, g2b, A["p1"], None, d2)
    g2c = bdata(4, g2b, d2)
    ge1 = mask(R(ge1 + _unpool(g2c, d1, 32) / 8.0), A["e1"])
    wgrad(3, ge1, A["e1a"], None, d1)
    g1a = mask(bdata(3, ge1, d1), A["e1a"])
    wgrad(2, g1a, A["p0"], None, d1)
    g1c = bdata(2, g1a, d1)
    ge0 = mask(R(ge0 + _unpool(g1c, d0, 16) / 8.0), A["e0"])
    wgrad(1, ge0, A["e0a"], None, d0)
    g0b = mask(bdata(1, ge0, d0), A["e0a"])
    wgrad(0, g0b, A["x"], None, d0)
    return loss, grads, out
