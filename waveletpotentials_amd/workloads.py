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


def solid_velocity(xv, yv, zv, dims, dx, centers, radii, tol=1e-5, max_cycles=60):
    """In place: u <- a (u - grad p) with p from the fractional multigrid solve of
    -div(a grad p) = -div(a u), V-cycles (2 pre / 2 post weighted-Jacobi sweeps,
    w = 0.8) until the FP64 relative residual is <= tol.  Returns
    (cycles, residual).  An FP32 iterate has a residual floor: rounding the
    converged FP64 solution of a 128x64x64 problem with one sphere to FP32
    already leaves a relative residual of 1.5e-6, growing about as the square of
    the cell count per axis."""
    if len(radii) == 0:
        return 0, 0.0
    (ax, ay, az), _ = face_fractions(dims, dx, centers, radii)
    nx, ny, nz = dims
    b = torch.empty((nz, nx, ny), device="cuda")
    api.CuConvergence.GetInstance().GetB_3d_frac_vel(b, xv, yv, zv, ax, ay, az, dims, dx, "n")
    mg = api.CuMultigridFrac3D(ax, ay, az, dims, dx, "n")
    p = torch.zeros_like(b)
    res, cycles = float("inf"), 0
    while cycles < max_cycles:
        mg.Solve(p, p, b, 1, 2, 2, 0.8)
        cycles += 1
        res = mg.GetResdualError(p, b)
        if res <= tol:
            break
    api.CuGradient.GetInstance().SubtractGradient3D_Frac_P(xv, yv, zv, p, ax, ay, az, dims, dx, "n")
    return cycles, res
