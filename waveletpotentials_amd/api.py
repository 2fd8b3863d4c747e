"""Python mirror of the reference's recovery-path classes over libwhhd.so.

Names, argument order and meaning follow the reference's C++ interface so a
caller of ``CuWaveletPotentialRecover3D`` & co. finds the same surface:

* ``WaveletType`` / ``proj_bc`` / ``proj_WaveletType`` / ``type_from_string`` /
  ``string_from_type``  (WaveletPotentials/WaveletTypes.h:7-143, WaveletTypes.cpp:4-67)
* ``CuWHHDForward.GetInstance().Solve1D``  (WaveletPotentials/cuWHHDForward.cuh:17-25)
* ``CuGetInnerProduct.GetInstance().GenerateIp``  (WaveletPotentials/cuGetInnerProduct.cuh:21-22)
* ``CuDivLocalProject3D`` / ``CuDivLocalProject``  (WaveletPotentials/cuDivLocalProject3D.cuh:7-47,
  WaveletPotentials/cuDivLocalProject.cuh:7-33)
* ``CuWaveletPotentialRecover3D`` / ``CuWaveletPotentialRecover2D``
  (WaveletPotentials/cuWaveletPotentialRecover3D.cuh:8-31, cuWaveletPotentialRecover2D.cuh)

Fields are torch tensors on the GPU (float32; Solve1D also takes float64),
stored y-fastest: a field of extents (nx, ny, nz) is a contiguous tensor of
shape (nz, nx, ny) (2D: (nx, ny)).  Where the reference aborts the process on
an error, these raise :class:`WhhdError`.  Work is enqueued on torch's current
stream.
"""
import ctypes
import enum

import torch

from ._lib import WhhdError, check, lib

__all__ = ["WaveletType", "proj_bc", "proj_WaveletType", "type_from_string", "string_from_type",
           "CuSolidLevelSet3D", "CuFastSweeping", "CuMultigridFrac3D", "CuGradient",
           "CuConvergence", "CudaMatAdd", "CudaMatMul", "CudaArraySum", "CudaArrayNormSum",
           "CudaArrayAbsSum", "CudaFindMaxValue", "CuDst3D", "CuDst", "CuParallelSweep3D", "CuWHHDForward", "CuGetInnerProduct", "CuDivLocalProject3D", "CuDivLocalProject",
           "CuWaveletPotentialRecover3D", "CuWaveletPotentialRecover2D", "CuBatchRecover3D",
           "UNet3D", "WriteDataTxt", "WriteDataETxt", "ReadsDataTxt", "WriteDataTxt_sparse",
           "WriteVectorTxt_sparse", "WriteDataBinary", "ReadDataBinary", "WritePly",
           "SetSliceZero", "StreamInitial", "CalculateErrorL2",
           "WhhdError",
           "mac_shapes3d", "mac_shapes2d"]


class WaveletType(enum.IntEnum):
    """WaveletPotentials/WaveletTypes.h:7-28 (same enumerator order)."""
    WT_CDF3_9 = 0
    WT_CDF3_11 = 1
    WT_CDF3_5 = 2
    WT_CDF2_10 = 3
    WT_CDF2_12 = 4
    WT_CDF2_6 = 5
    WT_CDF4_4 = 6
    WT_CDF4_6 = 7
    WT_CDF4_8 = 8
    WT_CDF4_10 = 9
    WT_CDF2_8 = 10
    WT_CDF3_7 = 11
    WT_CDF26_44 = 12
    WT_CDF28_46 = 13
    WT_CDF44_26 = 14
    WT_CDF46_28 = 15
    WT_CDF2_0 = 16
    WT_CDF1_1 = 17
    WT_CDF1_7 = 18


_NAMES = {WaveletType.WT_CDF3_5: "cdf3.5", WaveletType.WT_CDF3_7: "cdf3.7",
          WaveletType.WT_CDF3_9: "cdf3.9", WaveletType.WT_CDF3_11: "cdf3.11",
          WaveletType.WT_CDF2_6: "cdf2.6", WaveletType.WT_CDF2_8: "cdf2.8",
          WaveletType.WT_CDF2_10: "cdf2.10", WaveletType.WT_CDF2_12: "cdf2.12",
          WaveletType.WT_CDF4_4: "cdf4.4", WaveletType.WT_CDF4_6: "cdf4.6",
          WaveletType.WT_CDF4_8: "cdf4.8", WaveletType.WT_CDF4_10: "cdf4.10",
          WaveletType.WT_CDF2_0: "cdf2.0", WaveletType.WT_CDF1_1: "cdf1.1",
          WaveletType.WT_CDF1_7: "cdf1.7"}


def proj_bc(bc):
    """WaveletPotentials/WaveletTypes.h:102-118: 'n' <-> 'd', others unchanged."""
    return {"n": "d", "d": "n"}.get(bc, bc)


def proj_WaveletType(t):
    """WaveletPotentials/WaveletTypes.h:120-143."""
    r = lib().whhd_proj_wavelet(int(t))
    return None if r < 0 else WaveletType(r)


def type_from_string(s):
    """WaveletPotentials/WaveletTypes.cpp:4-67: 'n_cdf3.7' -> WT_CDF3_7 (default WT_CDF3_5)."""
    tail = s[2:]
    for t, name in _NAMES.items():
        if name == tail:
            return t
    return WaveletType.WT_CDF3_5


def string_from_type(t, bc):
    """WaveletPotentials/WaveletTypes.h:67-100."""
    return f"{bc}_{_NAMES[WaveletType(t)]}"


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t, dtype=torch.float32, numel=None, name="tensor"):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name}: expected a CUDA/HIP torch tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name}: expected dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: expected {numel} elements, got {t.numel()}")
    return ctypes.c_void_p(t.data_ptr())


def _b(ch):
    return ch.encode() if isinstance(ch, str) else ch


def mac_shapes3d(nx, ny, nz):
    """Tensor shapes (nz, nx, ny)-ordered of xv, yv, zv, qx, qy, qz for nx*ny*nz cells."""
    return {"xv": (nz, nx + 1, ny), "yv": (nz, nx, ny + 1), "zv": (nz + 1, nx, ny),
            "qx": (nz + 1, nx, ny + 1), "qy": (nz + 1, nx + 1, ny), "qz": (nz, nx + 1, ny + 1)}


def mac_shapes2d(nx, ny):
    return {"xv": (nx + 1, ny), "yv": (nx, ny + 1), "qz": (nx + 1, ny + 1)}


class CuWHHDForward:
    """CuWHHDForward (WaveletPotentials/cuWHHDForward.cuh:11-37): line transforms."""
    _instance = None

    @classmethod
    def GetInstance(cls):
        if cls._instance is None:
            cls._instance = cls()
        return cls._instance

    def Solve1D(self, dst, src, dim, level, direction, bc="n", rsw_type=WaveletType.WT_CDF3_5,
                is_forward=True):
        """Multi-level transform of every line of the (dim.x, dim.y, dim.z) array
        along `direction`; src None means in place (cuWHHDForward.cu:1111-1120)."""
        if src is None:
            src = dst
        nx, ny, nz = dim
        n = nx * ny * nz
        if dst.dtype == torch.float64:
            f = lib().whhd_transform1d_f64
            dt = torch.float64
        else:
            f = lib().whhd_transform1d
            dt = torch.float32
        check(f(_ptr(dst, dt, n, "dst"), _ptr(src, dt, n, "src"), nx, ny, nz, int(level),
                _b(direction), _b(bc), int(rsw_type), int(bool(is_forward)), _stream()),
              "Solve1D")


class CuGetInnerProduct:
    """CuGetInnerProduct (WaveletPotentials/cuGetInnerProduct.cuh:8-28); tables are cached
    per (type, bc, n) like the reference's ips_map_."""
    _instance = None

    def __init__(self):
        self._cache = {}

    @classmethod
    def GetInstance(cls):
        if cls._instance is None:
            cls._instance = cls()
        return cls._instance

    def GenerateIp(self, type_or_str, bc_or_dim, dim=None):
        """GenerateIp(type, bc, n) or GenerateIp("d_cdf4.6", n): (8, n) float64 tensor."""
        if isinstance(type_or_str, str):
            t, bc, n = type_from_string(type_or_str), type_or_str[0], bc_or_dim
        else:
            t, bc, n = type_or_str, bc_or_dim, dim
        key = (int(t), bc, int(n))
        if key not in self._cache:
            out = torch.empty((8, n), dtype=torch.float64, device="cuda")
            check(lib().whhd_gram_table(_ptr(out, torch.float64), int(t), _b(bc), int(n),
                                        _stream()), "GenerateIp")
            self._cache[key] = out
        return self._cache[key]

    def GetIp(self, type_or_str, bc_or_dim, dim=None):
        if isinstance(type_or_str, str):
            key = (int(type_from_string(type_or_str)), type_or_str[0], int(bc_or_dim))
        else:
            key = (int(type_or_str), bc_or_dim, int(dim))
        return self._cache.get(key)


class _Handle:
    _destroy = None

    def __init__(self):
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if self._h and self._h.value and self._destroy:
                getattr(lib(), self._destroy)(self._h)
                self._h = ctypes.c_void_p()
        except Exception:
            pass


class CuDivLocalProject3D(_Handle):
    """CuDivLocalProject3D(int3 dim, int3 levels, DTYPE3 dx, WaveletType type, char bc)
    (WaveletPotentials/cuDivLocalProject3D.cuh:11).  As in the reference, `dim` is
    the node extent (cells + 1) and `type`/`bc` the divergence type and bc."""
    _destroy = "whhd_project3d_destroy"

    def __init__(self, dim, levels, dx, type, bc):
        super().__init__()
        self.levels = tuple(int(v) for v in levels)
        self.n = tuple(1 << v for v in self.levels)
        check(lib().whhd_project3d_create(*self.n, *[float(v) for v in dx], int(type), _b(bc),
                                          ctypes.byref(self._h)), "CuDivLocalProject3D")

    def ProjectLocal_q_ccc(self, q_lwt_x, q_lwt_y, q_lwt_z, *rest):
        """(q..., v..., levels) or (q..., v_out..., v..., levels)
        (WaveletPotentials/cuDivLocalProject3D.cuh:17-20)."""
        args = [a for a in rest if isinstance(a, torch.Tensor)]
        nx, ny, nz = self.n
        s = mac_shapes3d(nx, ny, nz)
        size = {k: int(torch.Size(v).numel()) for k, v in s.items()}
        q = [_ptr(q_lwt_x, numel=size["qx"], name="qx"), _ptr(q_lwt_y, numel=size["qy"], name="qy"),
             _ptr(q_lwt_z, numel=size["qz"], name="qz")]
        if len(args) == 6:
            vo = [_ptr(args[0], numel=size["xv"]), _ptr(args[1], numel=size["yv"]),
                  _ptr(args[2], numel=size["zv"])]
            v = args[3:]
        elif len(args) == 3:
            vo = [None, None, None]
            v = args
        else:
            raise TypeError("ProjectLocal_q_ccc expects 3 potential and 3 or 6 velocity tensors")
        vp = [_ptr(v[0], numel=size["xv"], name="xv"), _ptr(v[1], numel=size["yv"], name="yv"),
              _ptr(v[2], numel=size["zv"], name="zv")]
        check(lib().whhd_project3d_run(self._h, *q, *vo, *vp, _stream()), "ProjectLocal_q_ccc")


class CuDivLocalProject(_Handle):
    """CuDivLocalProject(int2 dim, int2 levels, DTYPE2 dx, WaveletType, char bc)
    (WaveletPotentials/cuDivLocalProject.cuh:10); `dim` is the node extent."""
    _destroy = "whhd_project2d_destroy"

    def __init__(self, dim, levels, dx, type, bc):
        super().__init__()
        self.levels = tuple(int(v) for v in levels)
        self.n = tuple(1 << v for v in self.levels)
        check(lib().whhd_project2d_create(*self.n, float(dx[0]), float(dx[1]), int(type), _b(bc),
                                          ctypes.byref(self._h)), "CuDivLocalProject")

    def ProjectSingle(self, p_lwt, xv_lwt, yv_lwt):
        nx, ny = self.n
        check(lib().whhd_project2d_run(self._h, _ptr(p_lwt, numel=(nx + 1) * (ny + 1)),
                                       _ptr(xv_lwt, numel=(nx + 1) * ny),
                                       _ptr(yv_lwt, numel=nx * (ny + 1)), _stream()),
              "ProjectSingle")


class CuWaveletPotentialRecover3D(_Handle):
    """CuWaveletPotentialRecover3D(int3 dim, DTYPE3 dx, char bc, WaveletType curl,
    WaveletType div) (WaveletPotentials/cuWaveletPotentialRecover3D.cuh:12); `dim`
    is the cell count per axis."""
    _destroy = "whhd_recover3d_destroy"

    def __init__(self, dim, dx, bc, type_w_curl, type_w_div):
        super().__init__()
        self.dim = tuple(int(v) for v in dim)
        self.dx = tuple(float(v) for v in dx)
        self.bc = bc
        check(lib().whhd_recover3d_create(*self.dim, *self.dx, _b(bc), int(type_w_curl),
                                          int(type_w_div), ctypes.byref(self._h)),
              "CuWaveletPotentialRecover3D")
        s = mac_shapes3d(*self.dim)
        self._size = {k: int(torch.Size(v).numel()) for k, v in s.items()}

    def _q(self, qx, qy, qz):
        s = self._size
        return [_ptr(qx, numel=s["qx"], name="qx"), _ptr(qy, numel=s["qy"], name="qy"),
                _ptr(qz, numel=s["qz"], name="qz")]

    def _v(self, xv, yv, zv):
        s = self._size
        return [_ptr(xv, numel=s["xv"], name="xv"), _ptr(yv, numel=s["yv"], name="yv"),
                _ptr(zv, numel=s["zv"], name="zv")]

    def Solve(self, qx, qy, qz, *v):
        """Solve(qx, qy, qz, xv, yv, zv) or Solve(qx, qy, qz, xv_out, yv_out, zv_out, xv, yv, zv)
        (WaveletPotentials/cuWaveletPotentialRecover3D.cuh:15-16)."""
        if len(v) == 3:
            check(lib().whhd_recover3d_solve(self._h, *self._q(qx, qy, qz), *self._v(*v),
                                             _stream()), "Solve")
        elif len(v) == 6:
            check(lib().whhd_recover3d_solve_proj(self._h, *self._q(qx, qy, qz), *self._v(*v[:3]),
                                                  *self._v(*v[3:]), _stream()), "Solve")
        else:
            raise TypeError("Solve expects 3 or 6 velocity tensors")

    def CutCoefficients(self, qx_out, qy_out, qz_out, qx, qy, qz, level):
        check(lib().whhd_recover3d_cut(self._h, *self._q(qx_out, qy_out, qz_out),
                                       *self._q(qx, qy, qz), int(level), _stream()),
              "CutCoefficients")

    def TraceBegin(self, max_solves):
        """Record per-stage HIP events during the next `max_solves` solves."""
        check(lib().whhd_recover3d_trace_begin(self._h, int(max_solves)), "TraceBegin")

    def TraceEnd(self):
        """Wait for the traced solves; returns ([{stage, kernel, us}], solves)."""
        cap = 64
        us = (ctypes.c_float * cap)()
        ns, nsol = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().whhd_recover3d_trace_end(self._h, us, cap, ctypes.byref(ns), ctypes.byref(nsol)),
              "TraceEnd")
        out = []
        for i in range(min(ns.value, cap)):
            label = (lib().whhd_recover3d_trace_label(self._h, i) or b"").decode()
            stage, _, kernel = label.partition(" | ")
            out.append({"stage": stage, "kernel": kernel, "us": float(us[i])})
        return out, nsol.value


class CuBatchRecover3D(_Handle):
    """Batched recovery of independent 3D grids of uneven sizes.  No reference
    counterpart: each problem gets CuWaveletPotentialRecover3D::Solve
    (WaveletPotentials/cuWaveletPotentialRecover3D.cu:16-61).  `dims` and `dxs`
    list (nx, ny, nz) and (dx, dy, dz) per problem.  The arrays of every problem
    live in one device arena owned by this object; :meth:`arrays` returns views
    of problem i's xv, yv, zv, qx, qy, qz shaped as :func:`mac_shapes3d`."""
    _destroy = "whhd_batch3d_destroy"
    _KEYS = ("xv", "yv", "zv", "qx", "qy", "qz")

    def __init__(self, dims, dxs, bc, type_w_curl, type_w_div):
        super().__init__()
        self.dims = [tuple(int(v) for v in d) for d in dims]
        n = len(self.dims)
        d = (ctypes.c_int * (3 * n))(*[v for t in self.dims for v in t])
        x = (ctypes.c_double * (3 * n))(*[float(v) for t in dxs for v in t])
        if len(x) != 3 * n:
            raise ValueError("one (dx, dy, dz) per problem")
        L = lib()
        check(L.whhd_batch3d_create(n, d, x, _b(bc), int(type_w_curl), int(type_w_div),
                                    ctypes.byref(self._h)), "CuBatchRecover3D")
        total = ctypes.c_longlong()
        offs = (ctypes.c_longlong * (6 * n))()
        check(L.whhd_batch3d_layout(self._h, ctypes.byref(total), offs), "CuBatchRecover3D.layout")
        self.arena = torch.zeros(int(total.value), dtype=torch.float32, device="cuda")
        self._views = []
        for i, (nx, ny, nz) in enumerate(self.dims):
            shp = mac_shapes3d(nx, ny, nz)
            v = {}
            for k, key in enumerate(self._KEYS):
                o, m = int(offs[6 * i + k]), int(torch.Size(shp[key]).numel())
                v[key] = self.arena[o:o + m].view(shp[key])
            self._views.append(v)

    def arrays(self, i):
        return self._views[i]

    def _arena(self):
        return ctypes.c_void_p(self.arena.data_ptr())

    def Solve(self):
        check(lib().whhd_batch3d_solve(self._h, self._arena(), _stream()), "CuBatchRecover3D.Solve")

    def Capture(self):
        """Capture the whole batch as one HIP graph (on this object's arena)."""
        check(lib().whhd_batch3d_capture(self._h, self._arena()), "CuBatchRecover3D.Capture")

    def Replay(self):
        check(lib().whhd_batch3d_replay(self._h, _stream()), "CuBatchRecover3D.Replay")

    def Autotune(self, reps=3):
        """Pick each size class's launch shapes by timing; overwrites the q arrays."""
        check(lib().whhd_batch3d_autotune(self._h, self._arena(), int(reps), _stream()),
              "CuBatchRecover3D.Autotune")

    def classes(self):
        L = lib()
        out = []
        for c in range(L.whhd_batch3d_num_classes(self._h)):
            info = (ctypes.c_int * 6)()
            check(L.whhd_batch3d_class_info(self._h, c, info), "CuBatchRecover3D.classes")
            out.append({"dims": tuple(info[:3]), "problems": info[3], "lp_waves": info[4],
                        "proj_zb": info[5]})
        return out

    def SetTiles(self, cls, lp_waves, proj_zb):
        check(lib().whhd_batch3d_set_tiles(self._h, int(cls), int(lp_waves), int(proj_zb)),
              "CuBatchRecover3D.SetTiles")


class CuWaveletPotentialRecover2D(_Handle):
    """CuWaveletPotentialRecover2D(int2 dim, DTYPE2 dx, char bc, WaveletType curl,
    WaveletType div) (WaveletPotentials/cuWaveletPotentialRecover2D.cuh)."""
    _destroy = "whhd_recover2d_destroy"

    def __init__(self, dim, dx, bc, type_w_curl, type_w_div):
        super().__init__()
        self.dim = tuple(int(v) for v in dim)
        check(lib().whhd_recover2d_create(*self.dim, float(dx[0]), float(dx[1]), _b(bc),
                                          int(type_w_curl), int(type_w_div),
                                          ctypes.byref(self._h)), "CuWaveletPotentialRecover2D")

    def Solve(self, qz, xv, yv):
        nx, ny = self.dim
        check(lib().whhd_recover2d_solve(self._h, _ptr(qz, numel=(nx + 1) * (ny + 1), name="qz"),
                                         _ptr(xv, numel=(nx + 1) * ny, name="xv"),
                                         _ptr(yv, numel=nx * (ny + 1), name="yv"), _stream()),
              "Solve")


class CuSolidLevelSet3D:
    """CuSolidLevelSet3D(int3 dim, DTYPE3 dx) (WaveletPotentials/cuSolidLevelSet3D.cuh:8-31):
    a cell-centred solid level set `ls` (shape (nz, nx, ny)), negative inside solids."""

    def __init__(self, dim, dx):
        self.dim = tuple(int(v) for v in dim)
        self.dx = tuple(float(v) for v in dx)
        nx, ny, nz = self.dim
        self.ls = torch.empty((nz, nx, ny), dtype=torch.float32, device="cuda")

    def InitLsSphere(self, center, radius):
        """cuInitLsSphere (WaveletPotentials/cuSolidLevelSet3D.cu:11-26, 370-374)."""
        self.InitLsSpheres([center], [radius])

    def InitLsSpheres(self, centers, radii):
        """Union of spheres: ls = min_i |p - c_i| - r_i (1..64 spheres)."""
        c = (ctypes.c_double * (3 * len(radii)))(*[float(v) for cc in centers for v in cc])
        r = (ctypes.c_double * len(radii))(*[float(v) for v in radii])
        check(lib().whhd_levelset_spheres(_ptr(self.ls, name="ls"), *self.dim, *self.dx, c, r,
                                          len(radii), _stream()), "InitLsSpheres")

    def GetFrac(self, ax, ay, az):
        """GetFrac(ax, ay, az) (WaveletPotentials/cuSolidLevelSet3D.cu:348-353)."""
        nx, ny, nz = self.dim
        check(lib().whhd_frac_from_ls(_ptr(ax, numel=(nx + 1) * ny * nz, name="ax"),
                                      _ptr(ay, numel=nx * (ny + 1) * nz, name="ay"),
                                      _ptr(az, numel=nx * ny * (nz + 1), name="az"),
                                      _ptr(self.ls, name="ls"), *self.dim, *self.dx, _stream()),
              "GetFrac")

    @staticmethod
    def GetFracDivide(ax, ay, az, ls, dim, dx):
        """The static GetFrac(ax, ay, az, ls, dim, dx)
        (WaveletPotentials/cuSolidLevelSet3D.cu:359-372): supersampled fractions
        from the cuGetA{x,y,z}_divide kernels (:164-317)."""
        nx, ny, nz = (int(v) for v in dim)
        check(lib().whhd_frac_divide(_ptr(ax, numel=(nx + 1) * ny * nz, name="ax"),
                                     _ptr(ay, numel=nx * (ny + 1) * nz, name="ay"),
                                     _ptr(az, numel=nx * ny * (nz + 1), name="az"),
                                     _ptr(ls, numel=nx * ny * nz, name="ls"), nx, ny, nz,
                                     *(float(v) for v in dx), _stream()), "GetFrac (static)")


class CuFastSweeping:
    """CuFastSweeping (WaveletPotentials/cuFastSweeping.cuh): eikonal relaxation."""
    _instance = None

    @classmethod
    def GetInstance(cls):
        if cls._instance is None:
            cls._instance = cls()
        return cls._instance

    def Solve(self, ls, dim, dx, n_iters):
        """Solve(ls, dim, dx, n_iters) (WaveletPotentials/cuFastSweeping.cu:250-257);
        the scratch iterate comes from torch's allocator."""
        nx, ny, nz = (int(v) for v in dim)
        p = _ptr(ls, numel=nx * ny * nz, name="ls")
        scratch = torch.empty_like(ls)
        check(lib().whhd_eikonal_sweep(p, _ptr(scratch, name="scratch"), nx, ny, nz, float(dx),
                                       int(n_iters), _stream()), "CuFastSweeping.Solve")


class CuMultigridFrac3D(_Handle):
    """CuMultigridFrac3D(DTYPE* Ax, DTYPE* Ay, DTYPE* Az, int3 dim, DTYPE3 dx, char bc='n')
    (WaveletPotentials/cuMultigridFrac3D.cuh:15-18): fractional-area multigrid."""
    _destroy = "whhd_mgfrac3d_destroy"

    def __init__(self, Ax, Ay, Az, dim, dx, bc="n"):
        super().__init__()
        self.dim = tuple(int(v) for v in dim)
        nx, ny, nz = self.dim
        self._n = nx * ny * nz
        check(lib().whhd_mgfrac3d_create(_ptr(Ax, numel=(nx + 1) * ny * nz, name="Ax"),
                                         _ptr(Ay, numel=nx * (ny + 1) * nz, name="Ay"),
                                         _ptr(Az, numel=nx * ny * (nz + 1), name="Az"),
                                         nx, ny, nz, *(float(v) for v in dx), _b(bc),
                                         ctypes.byref(self._h)), "CuMultigridFrac3D")

    def levels(self):
        dims = (ctypes.c_int * 96)()
        k = lib().whhd_mgfrac3d_levels(self._h, dims, 32)
        return [tuple(dims[3 * i:3 * i + 3]) for i in range(k)]

    def Solve(self, dst, pre_v, f, n_iters=10, nu1=2, nu2=2, weight=0.8):
        """Solve(dst, pre_v, f, n_iters, nu1, nu2, weight) (cuMultigridFrac3D.cu:249-265)."""
        pv = None if pre_v is None else _ptr(pre_v, numel=self._n, name="pre_v")
        check(lib().whhd_mgfrac3d_solve(self._h, _ptr(dst, numel=self._n, name="dst"), pv,
                                        _ptr(f, numel=self._n, name="f"), int(n_iters), int(nu1),
                                        int(nu2), float(weight), _stream()), "Solve")

    def SolveToTolerance(self, dst, dst_lo, pre_v, f, tol=1e-6, max_cycles=100, nu1=2, nu2=2,
                         weight=0.8):
        """V-cycles until ||f - A (dst + dst_lo)|| <= tol ||f|| (mixed-precision
        refinement; no reference counterpart).  Returns (cycles, residual,
        converged)."""
        pv = None if pre_v is None else _ptr(pre_v, numel=self._n, name="pre_v")
        cyc, rel = ctypes.c_int(0), ctypes.c_double(0.0)
        st = lib().whhd_mgfrac3d_solve_tol(self._h, _ptr(dst, numel=self._n, name="dst"),
                                           _ptr(dst_lo, numel=self._n, name="dst_lo"), pv,
                                           _ptr(f, numel=self._n, name="f"), int(max_cycles),
                                           float(tol), int(nu1), int(nu2), float(weight),
                                           ctypes.byref(cyc), ctypes.byref(rel), _stream())
        if st not in (0, 6):
            check(st, "SolveToTolerance")
        return cyc.value, rel.value, st == 0

    def SolvePCG(self, dst, dst_lo, pre_v, f, tol=1e-6, max_iters=100, nu1=2, nu2=2, weight=0.8):
        """Conjugate gradients preconditioned by one V-cycle until
        ||f - A (dst + dst_lo)|| <= tol ||f|| (no reference counterpart).
        Returns (iterations, residual, converged)."""
        pv = None if pre_v is None else _ptr(pre_v, numel=self._n, name="pre_v")
        it, rel = ctypes.c_int(0), ctypes.c_double(0.0)
        st = lib().whhd_mgfrac3d_solve_pcg(self._h, _ptr(dst, numel=self._n, name="dst"),
                                           _ptr(dst_lo, numel=self._n, name="dst_lo"), pv,
                                           _ptr(f, numel=self._n, name="f"), int(max_iters),
                                           float(tol), int(nu1), int(nu2), float(weight),
                                           ctypes.byref(it), ctypes.byref(rel), _stream())
        if st not in (0, 6):
            check(st, "SolvePCG")
        return it.value, rel.value, st == 0

    def GetResdualError(self, dst, f):
        """||f - A dst|| / ||f|| in FP64 (the reference's GetResdualError / GetFracRhs)."""
        rel = ctypes.c_double(0.0)
        check(lib().whhd_mgfrac3d_residual(self._h, _ptr(dst, numel=self._n, name="v"),
                                           _ptr(f, numel=self._n, name="f"), ctypes.byref(rel),
                                           _stream()), "GetResdualError")
        return rel.value


class UNet3D(_Handle):
    """Learned initial guess for CuMultigridFrac3D: a 3-level 3D U-Net (16/32/64
    channels, BF16 matrix-core convolutions with FP32 accumulation) from (f,
    solid level set) to the pressure; its prediction is the multigrid's pre_v.
    No reference counterpart (the reference's Solve takes any pre_v,
    WaveletPotentials/cuMultigridFrac3D.cu:249-258).  `dim` = (nx, ny, nz) cells of
    the solve; the network runs `factor` times coarser (block means in,
    trilinear interpolation out), so the extents are multiples of 4 * factor."""
    _destroy = "whhd_unet_destroy"
    MAGIC = b"WHUN"

    def __init__(self, dim, seed=0, factor=1):
        super().__init__()
        self.dim = tuple(int(v) for v in dim)
        self.factor = int(factor)
        check(lib().whhd_unet_create(*self.dim, self.factor, int(seed), ctypes.byref(self._h)),
              "UNet3D")
        self.n = int(self.dim[0] * self.dim[1] * self.dim[2])

    def num_params(self):
        return int(lib().whhd_unet_num_params(self._h))

    def get_params(self, which=0):
        """FP32 master weights (which=0) or the last gradients (which=1)."""
        import numpy as np
        out = np.empty(self.num_params(), dtype=np.float32)
        check(lib().whhd_unet_get_params(self._h, out.ctypes.data_as(ctypes.c_void_p), int(which)),
              "UNet3D.get_params")
        return out

    def set_params(self, params):
        import numpy as np
        a = np.ascontiguousarray(params, dtype=np.float32)
        if a.size != self.num_params():
            raise ValueError(f"expected {self.num_params()} parameters, got {a.size}")
        check(lib().whhd_unet_set_params(self._h, a.ctypes.data_as(ctypes.c_void_p)),
              "UNet3D.set_params")

    def predict(self, p0, f, ls, length):
        check(lib().whhd_unet_predict(self._h, _ptr(p0, numel=self.n, name="p0"),
                                      _ptr(f, numel=self.n, name="f"),
                                      _ptr(ls, numel=self.n, name="ls"), float(length), _stream()),
              "UNet3D.predict")

    def loss_grad(self, f, ls, p, length):
        loss = ctypes.c_double(0.0)
        check(lib().whhd_unet_loss_grad(self._h, _ptr(f, numel=self.n, name="f"),
                                        _ptr(ls, numel=self.n, name="ls"),
                                        _ptr(p, numel=self.n, name="p"), float(length),
                                        ctypes.byref(loss), _stream()), "UNet3D.loss_grad")
        return loss.value

    def adam(self, lr):
        check(lib().whhd_unet_adam(self._h, float(lr), _stream()), "UNet3D.adam")

    def set_output_scale(self, scale):
        check(lib().whhd_unet_set_output_scale(self._h, float(scale)), "UNet3D.set_output_scale")

    def output_scale(self):
        return float(lib().whhd_unet_output_scale(self._h))

    def save(self, path):
        """Checkpoint (little endian): b"WHUN", uint32 version 3, uint64 count,
        float64 output scale, uint32 factor, uint32 0, count FP32 weights."""
        import numpy as np
        w = self.get_params(0)
        with open(path, "wb") as fh:
            fh.write(self.MAGIC + np.array([3], "<u4").tobytes() +
                     np.array([w.size], "<u8").tobytes() +
                     np.array([self.output_scale()], "<f8").tobytes() +
                     np.array([self.factor, 0], "<u4").tobytes() + w.astype("<f4").tobytes())

    @staticmethod
    def checkpoint_factor(path):
        import numpy as np
        with open(path, "rb") as fh:
            raw = fh.read(32)
        if raw[:4] != UNet3D.MAGIC or np.frombuffer(raw[4:8], "<u4")[0] != 3:
            raise ValueError(f"{path}: not a U-Net checkpoint of version 3")
        return int(np.frombuffer(raw[24:28], "<u4")[0])

    def load(self, path):
        import numpy as np
        with open(path, "rb") as fh:
            raw = fh.read()
        if raw[:4] != self.MAGIC or np.frombuffer(raw[4:8], "<u4")[0] != 3:
            raise ValueError(f"{path}: not a U-Net checkpoint of version 3")
        count = int(np.frombuffer(raw[8:16], "<u8")[0])
        factor = int(np.frombuffer(raw[24:28], "<u4")[0])
        if factor != self.factor:
            raise ValueError(f"{path}: trained with factor {factor}, this network has {self.factor}")
        self.set_output_scale(float(np.frombuffer(raw[16:24], "<f8")[0]))
        self.set_params(np.frombuffer(raw[32:32 + 4 * count], "<f4"))


def _face_ptrs(dim, xv, yv, zv, names=("xv", "yv", "zv")):
    nx, ny, nz = dim
    return (_ptr(xv, numel=(nx + 1) * ny * nz, name=names[0]),
            _ptr(yv, numel=nx * (ny + 1) * nz, name=names[1]),
            _ptr(zv, numel=nx * ny * (nz + 1), name=names[2]))


class CuGradient:
    """CuGradient (WaveletPotentials/cuGradient.cuh): face gradients with fractions."""
    _instance = None

    @classmethod
    def GetInstance(cls):
        if cls._instance is None:
            cls._instance = cls()
        return cls._instance

    def Gradient3D_Frac_P(self, xv, yv, zv, p, ax, ay, az, dim, dx, bc="n"):
        """Gradient3D_Frac_P (WaveletPotentials/cuGradient.cu:2514-2545), uniform spacing."""
        dim = tuple(int(v) for v in dim)
        nx, ny, nz = dim
        check(lib().whhd_gradient3d_frac(*_face_ptrs(dim, xv, yv, zv),
                                         _ptr(p, numel=nx * ny * nz, name="p"),
                                         *_face_ptrs(dim, ax, ay, az, ("ax", "ay", "az")), *dim,
                                         *(float(v) for v in dx), _b(bc), _stream()),
              "Gradient3D_Frac_P")

    def SubtractGradient3D_Frac_P(self, xv, yv, zv, p, ax, ay, az, dim, dx, bc="n"):
        """u <- a * (u - grad p) in one pass."""
        dim = tuple(int(v) for v in dim)
        nx, ny, nz = dim
        check(lib().whhd_subtract_gradient_frac(*_face_ptrs(dim, xv, yv, zv),
                                                _ptr(p, numel=nx * ny * nz, name="p"),
                                                *_face_ptrs(dim, ax, ay, az, ("ax", "ay", "az")),
                                                *dim, *(float(v) for v in dx), _b(bc), _stream()),
              "SubtractGradient3D_Frac_P")

    def SubtractGradient3D_Frac_P2(self, xv, yv, zv, p_hi, p_lo, ax, ay, az, dim, dx, bc="n"):
        """u <- a * (u - grad p_hi - grad p_lo) in one pass, for a potential held
        as the sum of two floats (the multigrid's SolveToTolerance pair)."""
        dim = tuple(int(v) for v in dim)
        nx, ny, nz = dim
        check(lib().whhd_subtract_gradient_frac2(*_face_ptrs(dim, xv, yv, zv),
                                                 _ptr(p_hi, numel=nx * ny * nz, name="p_hi"),
                                                 _ptr(p_lo, numel=nx * ny * nz, name="p_lo"),
                                                 *_face_ptrs(dim, ax, ay, az, ("ax", "ay", "az")),
                                                 *dim, *(float(v) for v in dx), _b(bc), _stream()),
              "SubtractGradient3D_Frac_P2")


class CuConvergence:
    """CuConvergence (WaveletPotentials/cuConvergence.cuh): divergence of a*u."""
    _instance = None

    @classmethod
    def GetInstance(cls):
        if cls._instance is None:
            cls._instance = cls()
        return cls._instance

    def GetB_3d_frac_vel(self, b, xv, yv, zv, ax, ay, az, dim, dx, bc="n"):
        """GetB_3d_frac_vel (WaveletPotentials/cuConvergence.cu:435-447): b = -div(a u)."""
        dim = tuple(int(v) for v in dim)
        nx, ny, nz = dim
        check(lib().whhd_divergence3d_frac(_ptr(b, numel=nx * ny * nz, name="b"),
                                           *_face_ptrs(dim, xv, yv, zv),
                                           *_face_ptrs(dim, ax, ay, az, ("ax", "ay", "az")), *dim,
                                           *(float(v) for v in dx), _stream()),
              "GetB_3d_frac_vel")


def CudaMatAdd(dst, src, size, scale):
    """dst += src * scale (WaveletPotentials/cudaMath.cu:110-117)."""
    check(lib().whhd_axpy(_ptr(dst, name="dst"), _ptr(src, name="src"), int(size), float(scale),
                          _stream()), "CudaMatAdd")


def CudaMatMul(dst, lsrc, rsrc, size):
    """dst = lsrc * rsrc (WaveletPotentials/cudaMath.cu:161-168)."""
    check(lib().whhd_mul(_ptr(dst, name="dst"), _ptr(lsrc, name="lsrc"), _ptr(rsrc, name="rsrc"),
                         int(size), _stream()), "CudaMatMul")


def _reduce(src, size, op):
    scratch = torch.empty(lib().whhd_reduce_scratch_doubles(), dtype=torch.float64,
                          device=src.device)
    out = ctypes.c_double(0.0)
    check(lib().whhd_reduce(_ptr(src, name="src"), int(size), op,
                            ctypes.c_void_p(scratch.data_ptr()), ctypes.byref(out), _stream()),
          "reduce")
    return out.value


def CudaArraySum(src, size):
    """Sum of the first `size` elements, FP64 accumulation (cudaMath.cu:1230-1237)."""
    return _reduce(src, size, 0)


def CudaArrayNormSum(src, size):
    """Sum of squares (cudaMath.cu:1249-1257)."""
    return _reduce(src, size, 1)


def CudaArrayAbsSum(src, size):
    """Sum of absolute values (cudaMath.cu:1239-1247)."""
    return _reduce(src, size, 2)


def CudaFindMaxValue(src, size):
    """max |x| over the array (cudaMath.cu:1269-1280; reduced on the device)."""
    return _reduce(src, size, 3)


class CuDst3D(_Handle):
    """CuDst3D(int3 dim, DTYPE3 dx) (WaveletPotentials/cuDst3D.cuh): DST-I Poisson solver.
    dtype torch.float64 builds the FP64 solver (the reference's USE_DOUBLE)."""
    _destroy = "whhd_dst3d_destroy"

    def __init__(self, dim, dx, dtype=torch.float32):
        super().__init__()
        self.dim = tuple(int(v) for v in dim)
        self.dtype = dtype
        f = "whhd_dst3d_create_f64" if dtype == torch.float64 else "whhd_dst3d_create"
        check(getattr(lib(), f)(*self.dim, *(float(v) for v in dx), ctypes.byref(self._h)),
              "CuDst3D")

    def Solve(self, dst, f):
        """Solve(dst, f): (-L_h) dst = f (WaveletPotentials/cuDst3D.cu:265-341)."""
        n = 1
        for v in self.dim:
            n *= v
        fn = "whhd_dst3d_solve_f64" if self.dtype == torch.float64 else "whhd_dst3d_solve"
        check(getattr(lib(), fn)(self._h, _ptr(dst, self.dtype, numel=n, name="dst"),
                                 _ptr(f, self.dtype, numel=n, name="f"), _stream()),
              type(self).__name__ + ".Solve")


class CuDst(CuDst3D):
    """CuDst(int2 dim, DTYPE2 dx) (WaveletPotentials/cuDst.cuh): 2D DST-I Poisson solver
    on a dim[0] x dim[1] node grid (arrays (dim[0], dim[1]), y fastest)."""

    def __init__(self, dim, dx, dtype=torch.float32):
        _Handle.__init__(self)
        self.dim = tuple(int(v) for v in dim)
        self.dtype = dtype
        f = "whhd_dst2d_create_f64" if dtype == torch.float64 else "whhd_dst2d_create"
        check(getattr(lib(), f)(*self.dim, *(float(v) for v in dx), ctypes.byref(self._h)), "CuDst")


class CuParallelSweep3D(_Handle):
    """CuParallelSweep3D(int3 dim, DTYPE3 dx, char bc) (WaveletPotentials/cuParallelSweep3D.cuh):
    the sweep + DST comparator, velocity bc 'n', 'd' or 'z' ('z' as 'd'); a 2-element
    dim makes the 2D sweep.  dtype torch.float64 builds the FP64 comparator."""
    _destroy = "whhd_sweep3d_destroy"

    def __init__(self, dim, dx, bc="n", dtype=torch.float32):
        super().__init__()
        self.dim = tuple(int(v) for v in dim)
        self.dtype = dtype
        sfx = "_f64" if dtype == torch.float64 else ""
        if len(self.dim) == 2:  # CuParallelSweep3D(int2, DTYPE2, char): the 2D sweep
            check(getattr(lib(), "whhd_sweep2d_create" + sfx)(
                *self.dim, *(float(v) for v in dx), _b(bc), ctypes.byref(self._h)),
                "CuParallelSweep3D")
            return
        check(getattr(lib(), "whhd_sweep3d_create" + sfx)(
            *self.dim, *(float(v) for v in dx), _b(bc), ctypes.byref(self._h)), "CuParallelSweep3D")
        s = mac_shapes3d(*self.dim)
        self._size = {k: int(torch.Size(v).numel()) for k, v in s.items()}

    def _sfx(self):
        return "_f64" if self.dtype == torch.float64 else ""

    def _q(self, qx, qy, qz):
        s = self._size
        d = self.dtype
        return [_ptr(qx, d, numel=s["qx"], name="qx"), _ptr(qy, d, numel=s["qy"], name="qy"),
                _ptr(qz, d, numel=s["qz"], name="qz")]

    def Solve(self, *args):
        """(qx, qy, qz, xv, yv, zv) for a 3D handle, (qz, xv, yv) for a 2D one."""
        d = self.dtype
        if len(self.dim) == 2:
            qz, xv, yv = args
            nx, ny = self.dim
            check(getattr(lib(), "whhd_sweep2d_solve" + self._sfx())(
                self._h, _ptr(qz, d, numel=(nx + 1) * (ny + 1), name="qz"),
                _ptr(xv, d, numel=(nx + 1) * ny, name="xv"),
                _ptr(yv, d, numel=nx * (ny + 1), name="yv"), _stream()), "Solve")
            return
        qx, qy, qz, xv, yv, zv = args
        s = self._size
        check(getattr(lib(), "whhd_sweep3d_solve" + self._sfx())(
            self._h, *self._q(qx, qy, qz), _ptr(xv, d, numel=s["xv"], name="xv"),
            _ptr(yv, d, numel=s["yv"], name="yv"), _ptr(zv, d, numel=s["zv"], name="zv"),
            _stream()), "Solve")

    def Project(self, qx, qy, qz):
        check(getattr(lib(), "whhd_sweep3d_project" + self._sfx())(
            self._h, *self._q(qx, qy, qz), _stream()), "Project")


# ---- verification helpers on host arrays (WaveletPotentials/StreamInitial.cpp,
# WaveletPotentials/Utils.cpp:184-331) ------------------------------------------
def _f32c(a, name):
    import numpy as np
    if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]):
        raise TypeError(f"{name}: expected a C-contiguous float32 numpy array")
    return a


def SetSliceZero(val, dim, ns):
    """SetSliceZero(val, dim, ns) (StreamInitial.cpp:5-43): zero the first and last
    slices normal to axis ns of the (dim.x, dim.y, dim.z) float32 array `val`
    (shape (nz, nx, ny)), in place."""
    a = _f32c(val, "val")
    nx, ny, nz = (int(v) for v in dim)
    if a.size != nx * ny * nz:
        raise ValueError("SetSliceZero: array size does not match dim")
    check(lib().whhd_set_slice_zero(a.ctypes.data_as(ctypes.c_void_p), nx, ny, nz, _b(ns)),
          "SetSliceZero")


class StreamInitial:
    """StreamInitial::GetInstance()->GradientStream (StreamInitial.cpp:87-175): face
    velocities from an edge potential (3D) or stream function (2D), on float32
    numpy arrays written in place."""
    _inst = None

    @classmethod
    def GetInstance(cls):
        if cls._inst is None:
            cls._inst = cls()
        return cls._inst

    def GradientStream(self, *args):
        """(xv, yv, zv, qx, qy, qz, dim, dx) or (xv, yv, qz, dim, dx)."""
        if len(args) == 8:
            *arrs, dim, dx = args
            names = ("xv", "yv", "zv", "qx", "qy", "qz")
            nx, ny, nz = (int(v) for v in dim)
            sizes = ((nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1),
                     nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz)
            p = []
            for a, nm, sz in zip(arrs, names, sizes):
                a = _f32c(a, nm)
                if a.size != sz:
                    raise ValueError(f"GradientStream: {nm} has {a.size} elements, expected {sz}")
                p.append(a.ctypes.data_as(ctypes.c_void_p))
            check(lib().whhd_gradient_stream3d(*p, nx, ny, nz, *(float(v) for v in dx)),
                  "GradientStream")
        elif len(args) == 5:
            xv, yv, qz, dim, dx = args
            nx, ny = (int(v) for v in dim)
            p = []
            for a, nm, sz in zip((xv, yv, qz), ("xv", "yv", "qz"),
                                 ((nx + 1) * ny, nx * (ny + 1), (nx + 1) * (ny + 1))):
                a = _f32c(a, nm)
                if a.size != sz:
                    raise ValueError(f"GradientStream: {nm} has {a.size} elements, expected {sz}")
                p.append(a.ctypes.data_as(ctypes.c_void_p))
            check(lib().whhd_gradient_stream2d(*p, nx, ny, *(float(v) for v in dx)),
                  "GradientStream")
        else:
            raise TypeError("GradientStream takes (xv, yv, zv, qx, qy, qz, dim, dx) or "
                            "(xv, yv, qz, dim, dx)")


def CalculateErrorL2(*args):
    """CalculateErrorL2(base, check, size) or (base1, check1, size1, base2, check2,
    size2, base3, check3, size3) (Utils.cpp:184-207, 296-331): the relative L2
    error of `check` against `base` over float32 arrays."""
    if len(args) not in (3, 9):
        raise TypeError("CalculateErrorL2 takes 3 or 9 arguments")
    k = len(args) // 3
    B = (ctypes.c_void_p * k)()
    C = (ctypes.c_void_p * k)()
    N = (ctypes.c_longlong * k)()
    keep = []
    for i in range(k):
        b, c, size = args[3 * i: 3 * i + 3]
        n = 1
        for v in size:
            n *= int(v)
        b, c = _f32c(b, "base"), _f32c(c, "check")
        if b.size != n or c.size != n:
            raise ValueError("CalculateErrorL2: array size does not match size")
        keep += [b, c]
        B[i] = b.ctypes.data
        C[i] = c.ctypes.data
        N[i] = n
    out = ctypes.c_double(0.0)
    check(lib().whhd_error_l2(B, C, N, k, ctypes.byref(out)), "CalculateErrorL2")
    return out.value


# ---- field I/O on host arrays (WaveletPotentials/Utils.cpp:812-990) ------------
def _host(a, name):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 3:
        raise ValueError(f"{name}: expected a (nz, nx, ny) array")
    return a


def _dims(a):
    nz, nx, ny = a.shape
    return nx, ny, nz


def _path(p):
    return str(p).encode()


def WriteDataTxt(path, data, digits=12, scientific=False):
    """One row of ny values per (x, z), fixed notation with `digits` decimals."""
    a = _host(data, "data")
    check(lib().whhd_write_txt(_path(path), a.ctypes.data_as(ctypes.c_void_p), *_dims(a),
                               int(digits), int(bool(scientific))), "WriteDataTxt")


def WriteDataETxt(path, data):
    """WriteDataTxt in scientific notation with 14 digits."""
    WriteDataTxt(path, data, 14, True)


def ReadsDataTxt(path, dims):
    """Read an (nx, ny, nz) array written by WriteDataTxt; returns (nz, nx, ny) FP32."""
    import numpy as np
    nx, ny, nz = (int(v) for v in dims)
    out = np.empty((nz, nx, ny), dtype=np.float32)
    check(lib().whhd_read_txt(_path(path), out.ctypes.data_as(ctypes.c_void_p), nx, ny, nz),
          "ReadsDataTxt")
    return out


def WriteDataTxt_sparse(path, data, threshold=1e-6):
    a = _host(data, "data")
    check(lib().whhd_write_txt_sparse(_path(path), a.ctypes.data_as(ctypes.c_void_p), *_dims(a),
                                      float(threshold)), "WriteDataTxt_sparse")


def WriteVectorTxt_sparse(path, vx, vy, vz, threshold=1e-6):
    a, b, c = _host(vx, "vx"), _host(vy, "vy"), _host(vz, "vz")
    if not (a.shape == b.shape == c.shape):
        raise ValueError("WriteVectorTxt_sparse: the three arrays differ in shape")
    check(lib().whhd_write_vector_txt_sparse(_path(path), *(t.ctypes.data_as(ctypes.c_void_p)
                                                            for t in (a, b, c)), *_dims(a),
                                             float(threshold)), "WriteVectorTxt_sparse")


def WriteDataBinary(path, data):
    import numpy as np
    a = np.ascontiguousarray(data, dtype=np.float32)
    check(lib().whhd_write_binary(_path(path), a.ctypes.data_as(ctypes.c_void_p), a.size),
          "WriteDataBinary")


def ReadDataBinary(path, shape):
    import numpy as np
    out = np.empty(shape, dtype=np.float32)
    check(lib().whhd_read_binary(_path(path), out.ctypes.data_as(ctypes.c_void_p), out.size),
          "ReadDataBinary")
    return out


def WritePly(path, data, dx, threshold=0.0, binary=True):
    """PLY point cloud of a cell field (vertex at each cell centre with property
    `value`, where |value| >= threshold).  Returns the vertex count."""
    a = _host(data, "data")
    n = ctypes.c_longlong(0)
    check(lib().whhd_write_ply(_path(path), a.ctypes.data_as(ctypes.c_void_p), *_dims(a),
                               *(float(v) for v in dx), float(threshold), int(bool(binary)),
                               ctypes.byref(n)), "WritePly")
    return n.value
