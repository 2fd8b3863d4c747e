"""GPU parity of the solid-geometry kernels against the FP64 oracle.

Tolerances (FP32 kernels vs the FP64 oracle): level set and relaxed distance
within 2e-6 absolute on a unit-size domain (a few FP32 ulps of the values);
face fractions within 1e-4 absolute (the fraction divides by a face-local
length, so FP32 rounding is amplified where the interface is nearly normal
to the face).  Supersampled fractions (the static GetFrac) count sub-samples
by the sign of an FP32 interpolant: a sample within FP32 rounding of the zero
level can land on the other side, so a face may differ by a few sub-samples
(k/16384); at most 3 per face and on at most 1% of the cut faces.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def spheres(seed, k):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.2, 0.8, (k, 3)), rng.uniform(0.05, 0.15, k)


@pytest.mark.parametrize("n", [(16, 16, 16), (24, 16, 40), (64, 64, 64)])
def test_levelset_and_fractions(whhd, orc, n):
    dx = tuple(1.0 / v for v in n)
    c, r = spheres(20260002 + n[0], 8)
    s = whhd.CuSolidLevelSet3D(n, dx)
    s.InitLsSpheres(c, r)
    ref = orc.levelset_spheres(n, dx, c, r)
    got = s.ls.cpu().numpy().astype(np.float64)
    assert np.abs(got - ref).max() < 2e-6
    nx, ny, nz = n
    ax = torch.full((nz, nx + 1, ny), np.nan, device="cuda")
    ay = torch.full((nz, nx, ny + 1), np.nan, device="cuda")
    az = torch.full((nz + 1, nx, ny), np.nan, device="cuda")
    s.GetFrac(ax, ay, az)
    torch.cuda.synchronize()
    # fractions of the GPU level set itself, so only the fraction kernel is compared
    fref = orc.frac_from_ls(got, dx)
    for g, w in zip((ax, ay, az), fref):
        g = g.cpu().numpy().astype(np.float64)
        assert np.isfinite(g).all()
        assert np.abs(g - w).max() < 1e-4


def test_single_sphere_matches_reference_signature(whhd, orc):
    n = (16, 16, 16)
    dx = (1 / 16,) * 3
    s = whhd.CuSolidLevelSet3D(n, dx)
    s.InitLsSphere((0.5, 0.4, 0.6), 0.25)
    ref = orc.levelset_spheres(n, dx, [(0.5, 0.4, 0.6)], [0.25])
    assert np.abs(s.ls.cpu().numpy() - ref).max() < 2e-6


@pytest.mark.parametrize("iters", [1, 2, 7, 40])
def test_eikonal_sweep(whhd, orc, iters):
    n = (32, 24, 20)
    h = 1.0 / 32
    c, r = spheres(20260003, 4)
    ls0 = orc.levelset_spheres(n, (h,) * 3, c, r)
    ls = torch.tensor(ls0, dtype=torch.float32, device="cuda")
    whhd.CuFastSweeping.GetInstance().Solve(ls, n, h, iters)
    torch.cuda.synchronize()
    ref = orc.eikonal_sweep(ls0.astype(np.float32).astype(np.float64), h, iters)
    assert np.abs(ls.cpu().numpy().astype(np.float64) - ref).max() < 2e-6


@pytest.mark.parametrize("n", [(64, 64, 64), (40, 24, 32)])
def test_frac_divide_matches_oracle(whhd, orc, n):
    """Supersampled fractions of an 8-sphere level set (the static GetFrac,
    WaveletPotentials/cuSolidLevelSet3D.cu:164-317, 359-372) against the oracle
    on the same FP32 level set, two GPU calls in a row."""
    dx = tuple(1.0 / v for v in n)
    c, r = spheres(20260004 + n[0], 8)
    s = whhd.CuSolidLevelSet3D(n, dx)
    s.InitLsSpheres(c, r)
    ls = s.ls.cpu().numpy().astype(np.float64)
    ref = orc.frac_divide(ls, dx)
    nx, ny, nz = n
    outs = []
    for _ in range(2):
        a = (torch.full((nz, nx + 1, ny), np.nan, device="cuda"),
             torch.full((nz, nx, ny + 1), np.nan, device="cuda"),
             torch.full((nz + 1, nx, ny), np.nan, device="cuda"))
        whhd.CuSolidLevelSet3D.GetFracDivide(*a, s.ls, n, dx)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy().astype(np.float64) for t in a])
    for g1, g2 in zip(*outs):
        assert np.array_equal(g1, g2)
    for g, w in zip(outs[0], ref):
        assert np.isfinite(g).all()
        assert np.all(g * 16384 == np.round(g * 16384))
        diff = np.abs(g - w) * 16384
        cut = (w > 0) & (w < 1)
        assert cut.sum() > 100
        assert diff.max() <= 3.0
        assert (diff > 0).sum() <= max(1, 0.01 * cut.sum())


def test_solid_argument_errors(whhd):
    import ctypes
    L = whhd.lib()
    t = torch.zeros(8, device="cuda")
    p = ctypes.c_void_p(t.data_ptr())
    c = (ctypes.c_double * 3)(0.5, 0.5, 0.5)
    r = (ctypes.c_double * 1)(0.1)
    assert L.whhd_levelset_spheres(p, 2, 2, 2, 0.5, 0.5, 0.5, c, r, 0, None) == 1
    assert L.whhd_levelset_spheres(p, 2, 2, 2, 0.5, 0.5, 0.5, c, r, 65, None) == 1
    assert L.whhd_eikonal_sweep(p, p, 2, 2, 2, 0.5, 1, None) == 1
    assert L.whhd_eikonal_sweep(p, None, 2, 2, 2, 0.5, 1, None) == 1
    assert L.whhd_frac_divide(p, p, p, p, 1, 2, 2, 0.5, 0.5, 0.5, None) == 1
    assert L.whhd_frac_divide(p, None, p, p, 2, 2, 2, 0.5, 0.5, 0.5, None) == 1
