"""Device-side preparation steps of the solid-obstacle and batched workloads.

`solid_velocity` is the solid treatment the reference's test driver applies
before a recovery (WaveletPotentials/TestLocalProjection.cpp:620-701: level set
of the spheres, face fractions, f = -div(a u), fractional multigrid,
u <- a (u - grad p)),
built from this package's kernels (CuSolidLevelSet3D, CuConvergence,
CuMultigridFrac3D, CuGradient).  It runs on the GPU only.
"""
import torch

from . import api

__all__ = ["face_fractions", "solid_velocity", "spectral_velocity", "solid_problem",
           "mg_solve_from"]


def face_fractions(dims, dx, centers, radii):
    """Face fractions (ax, ay, az) of the union of spheres and its level set."""
    nx, ny, nz = dims
    ls = api.CuSolidLevelSet3D(dims, dx)
    ls.InitLsSpheres(centers, radii)
    ax = torch.empty((nz, nx + 1, ny), device="cuda")
    ay = torch.empty((nz, nx, ny + 1), device="cuda")
    az = torch.empty((nz + 1, nx, ny), device="cuda")
    ls.GetFrac(ax, ay, az)
    return (ax, ay, az), ls.ls


def solid_velocity(xv, yv, zv, dims, dx, centers, radii, tol=1e-6, max_cycles=80):
    """In place: u <- a (u - grad p) with p from the fractional multigrid solve of
    -div(a grad p) = -div(a u), conjugate gradients preconditioned by a V-cycle
    (2 pre / 2 post weighted-Jacobi sweeps, w = 0.8) until the FP64 relative
    residual of the two-float iterate is <= tol.  Returns (V-cycles, residual)."""
    if len(radii) == 0:
        return 0, 0.0
    (ax, ay, az), _ = face_fractions(dims, dx, centers, radii)
    nx, ny, nz = dims
    b = torch.empty((nz, nx, ny), device="cuda")
    api.CuConvergence.GetInstance().GetB_3d_frac_vel(b, xv, yv, zv, ax, ay, az, dims, dx, "n")
    mg = api.CuMultigridFrac3D(ax, ay, az, dims, dx, "n")
    p, p_lo = torch.empty_like(b), torch.empty_like(b)
    cycles, res, _ = mg.SolvePCG(p, p_lo, None, b, tol, max_cycles)
    # u <- a (u - grad p_hi - grad p_lo) in one pass
    api.CuGradient.GetInstance().SubtractGradient3D_Frac_P2(xv, yv, zv, p, p_lo, ax, ay, az, dims,
                                                           dx, "n")
    return cycles, res


def spectral_velocity(n, kmax, seed, extent=(1.0, 1.0, 1.0), device="cuda"):
    """inputs.velocity3d's field (curl of inputs.spectral_potential3d: same modes,
    same seed) evaluated on the GPU in FP64 and returned as FP32 (xv, yv, zv)."""
    import numpy as np
    from . import inputs as I
    nx, ny, nz = n
    hx, hy, hz = extent[0] / nx, extent[1] / ny, extent[2] / nz
    r, amp, phase, pol = I._modes(kmax, I._rng(seed))
    base = amp * np.exp(1j * phase)
    R = torch.from_numpy(r.astype(np.float64)).to(device)

    def axis(coords):
        t = torch.from_numpy(np.asarray(coords, dtype=np.float64)).to(device)
        return torch.exp(2j * np.pi * torch.outer(t, R).to(torch.complex128))

    def comp(coef, xs, ys, zs):
        c = torch.from_numpy(np.ascontiguousarray(coef)).to(device)
        t = torch.tensordot(c, axis(zs).contiguous(), dims=([2], [1]))   # (Kx, Ky, nz)
        t = torch.tensordot(t, axis(ys).contiguous(), dims=([1], [1]))   # (Kx, nz, ny)
        t = torch.tensordot(axis(xs).contiguous(), t, dims=([1], [0]))   # (nx, nz, ny)
        return t.real.permute(1, 0, 2).contiguous()

    xc = (np.arange(nx) + 0.5) * hx / extent[0]
    yc = (np.arange(ny) + 0.5) * hy / extent[1]
    zc = (np.arange(nz) + 0.5) * hz / extent[2]
    xn, yn, zn = (np.arange(nx + 1) * hx / extent[0], np.arange(ny + 1) * hy / extent[1],
                  np.arange(nz + 1) * hz / extent[2])
    qx = comp(base * pol[..., 0], xc, yn, zn)
    qy = comp(base * pol[..., 1], xn, yc, zn)
    qz = comp(base * pol[..., 2], xn, yn, zc)
    qx[0] = 0; qx[-1] = 0; qx[:, :, 0] = 0; qx[:, :, -1] = 0
    qy[0] = 0; qy[-1] = 0; qy[:, 0, :] = 0; qy[:, -1, :] = 0
    qz[:, 0, :] = 0; qz[:, -1, :] = 0; qz[:, :, 0] = 0; qz[:, :, -1] = 0
    xv = (qz[:, :, 1:] - qz[:, :, :-1]) / hy - (qy[1:] - qy[:-1]) / hz
    yv = (qx[1:] - qx[:-1]) / hz - (qz[:, 1:, :] - qz[:, :-1, :]) / hx
    zv = (qy[:, 1:, :] - qy[:, :-1, :]) / hx - (qx[:, :, 1:] - qx[:, :, :-1]) / hy
    return (xv.float().contiguous(), yv.float().contiguous(), zv.float().contiguous())


def solid_problem(n, seed, kmax=12, count=8, sweep_iters=None):
    """The solid-obstacle workload on the GPU: `count` spheres (inputs.spheres),
    cubic cells of side 1 / max(n); the level set is exact within two cells of
    the spheres and fast-swept elsewhere (`sweep_iters` relaxations, default
    2 max(n)); face fractions from it; the spectral velocity (|k| <= kmax) and
    f = -div(a u).  Returns a dict of device tensors and scalars."""
    from . import inputs as I
    import numpy as np
    h = 1.0 / max(n)
    dx = (h, h, h)
    ext = tuple(v * h for v in n)
    c, r = I.spheres(count, seed)
    c = c * np.array(ext)
    nx, ny, nz = n
    sls = api.CuSolidLevelSet3D(n, dx)
    sls.InitLsSpheres(c, r)
    ls = sls.ls
    far = ls > 2 * h
    ls[far] = 4.0 * max(ext)
    iters = 2 * max(n) if sweep_iters is None else int(sweep_iters)
    api.CuFastSweeping.GetInstance().Solve(ls, n, h, iters)
    ax = torch.empty((nz, nx + 1, ny), device="cuda")
    ay = torch.empty((nz, nx, ny + 1), device="cuda")
    az = torch.empty((nz + 1, nx, ny), device="cuda")
    sls.GetFrac(ax, ay, az)
    xv, yv, zv = spectral_velocity(n, kmax, seed, ext)
    f = torch.empty((nz, nx, ny), device="cuda")
    api.CuConvergence.GetInstance().GetB_3d_frac_vel(f, xv, yv, zv, ax, ay, az, n, dx, "n")
    return {"dims": tuple(n), "dx": dx, "length": ext[0], "ls": ls, "ax": ax, "ay": ay, "az": az,
            "f": f, "xv": xv, "yv": yv, "zv": zv, "centers": c, "radii": r, "sweeps": iters}


def mg_solve_from(mg, f, pre_v=None, tol=1e-6, max_cycles=200, method="vcycles"):
    """Solve to `tol` from pre_v (None: zero) by the mixed-precision V-cycle
    refinement (method "vcycles") or by conjugate gradients with a V-cycle as
    the preconditioner ("pcg"); returns (p_hi, p_lo, V-cycles, residual)."""
    hi, lo = torch.empty_like(f), torch.empty_like(f)
    if method == "pcg":
        cycles, res, _ = mg.SolvePCG(hi, lo, pre_v, f, tol, max_cycles)
    elif method == "vcycles":
        cycles, res, _ = mg.SolveToTolerance(hi, lo, pre_v, f, tol, max_cycles)
    else:
        raise ValueError(f"unknown method {method!r}")
    return hi, lo, cycles, res
