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

__all__ = ["face_fractions", "solid_velocity"]


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
    -div(a grad p) = -div(a u), V-cycles (2 pre / 2 post weighted-Jacobi sweeps,
    w = 0.8) until the FP64 relative residual of the two-float iterate is
    <= tol.  Returns (cycles, residual)."""
    if len(radii) == 0:
        return 0, 0.0
    (ax, ay, az), _ = face_fractions(dims, dx, centers, radii)
    nx, ny, nz = dims
    b = torch.empty((nz, nx, ny), device="cuda")
    api.CuConvergence.GetInstance().GetB_3d_frac_vel(b, xv, yv, zv, ax, ay, az, dims, dx, "n")
    mg = api.CuMultigridFrac3D(ax, ay, az, dims, dx, "n")
    p, p_lo = torch.empty_like(b), torch.empty_like(b)
    cycles, res, _ = mg.SolveToTolerance(p, p_lo, None, b, tol, max_cycles)
    grad = api.CuGradient.GetInstance()
    # u <- a (u - grad p) for the high part, then u <- u - a grad p_lo
    grad.SubtractGradient3D_Frac_P(xv, yv, zv, p, ax, ay, az, dims, dx, "n")
    gx, gy, gz = torch.empty_like(xv), torch.empty_like(yv), torch.empty_like(zv)
    grad.Gradient3D_Frac_P(gx, gy, gz, p_lo, ax, ay, az, dims, dx, "n")
    xv.sub_(ax * gx)
    yv.sub_(ay * gy)
    zv.sub_(az * gz)
    return cycles, res
