"""GPU parity of the sweep + DST comparator against the FP64 oracle.

Tolerances (FP32 vs FP64): sweep potentials 1e-5 relative to the largest
magnitude (running sums of up to 64 terms); DST solve 1e-5 relative (FP32
transforms: FFT lengths 2^L - 1, dense 255-term dot products otherwise); the
projected potential 1e-4 relative, and its node divergence below 2e-4 of the
largest edge term: the Laplacian of the FP32 roundoff left in phi by the
inverse transforms multiplies it by up to lambda_max / lambda_min ~ (N/pi)^2,
and an FP32 FFT of 32-128 points rounds about twice as much as a dense
product (1.8e-7 against 1.0e-7 relative per transform at n = 63, numpy
float32 emulation), so 64^3 reaches 1.1e-4 of the edge term.
"""
import numpy as np
import pytest
import torch

from waveletpotentials_amd import inputs as I

pytestmark = pytest.mark.gpu


def cuda(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("n", [(16, 16, 16), (24, 20, 28), (64, 64, 64)])
def test_sweep_and_project(whhd, orc, n):
    h = tuple(1.0 / v for v in n)
    q0 = I.spectral_potential3d(n, 4, I.seed_for(4, sum(n)))
    u = [a.astype(np.float32).astype(np.float64) for a in I.curl3d(*q0, h)]
    nx, ny, nz = n
    sw = whhd.CuParallelSweep3D(n, h, "n")
    q = [torch.full(s, np.nan, device="cuda") for s in ((nz + 1, nx, ny + 1), (nz + 1, nx + 1, ny),
                                                        (nz, nx + 1, ny + 1))]
    sw.Solve(*q, *(cuda(a) for a in u))
    ref = orc.sweep3d(u[0], u[1], n, h)
    for g, r in zip(q, ref):
        g = host(g)
        assert np.isfinite(g).all()
        assert np.abs(g - r).max() <= 1e-5 * max(np.abs(r).max(), 1e-30)
    sw.Project(*q)
    pref = orc.project_zero(*ref, n, h)
    scale = max(np.abs(r).max() for r in pref)
    got = [host(g) for g in q]
    for g, r in zip(got, pref):
        assert np.abs(g - r).max() <= 1e-4 * scale
    qx, qy, qz = got
    div = ((qx[1:-1, 1:, 1:-1] - qx[1:-1, :-1, 1:-1]) / h[0]
           + (qy[1:-1, 1:-1, 1:] - qy[1:-1, 1:-1, :-1]) / h[1]
           + (qz[1:, 1:-1, 1:-1] - qz[:-1, 1:-1, 1:-1]) / h[2])
    assert np.abs(div).max() <= 2e-4 * scale / min(h)


@pytest.mark.parametrize("shape,h", [((15, 15, 15), (0.1, 0.1, 0.1)), ((31, 17, 23), (0.1, 0.2, 0.15))])
def test_dst_solve(whhd, orc, shape, h):
    rng = np.random.default_rng(sum(shape))
    n1, n2, n3 = shape
    b = rng.standard_normal((n3, n1, n2)).astype(np.float32).astype(np.float64)
    d = whhd.CuDst3D(shape, h)
    phi = torch.empty((n3, n1, n2), device="cuda")
    d.Solve(phi, cuda(b))
    ref = orc.dst3d_solve(b, h)
    assert np.abs(host(phi) - ref).max() <= 1e-5 * np.abs(ref).max()
    # in place
    t = cuda(b)
    d.Solve(t, t)
    assert np.abs(host(t) - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("shape,h", [((255, 15, 31), (0.01, 0.1, 0.05)), ((63, 127, 7), (0.02, 0.01, 0.2)),
                                     ((31, 63, 255), (0.05, 0.03, 0.004))])
def test_dst_fft_lengths(whhd, orc, shape, h):
    """Lengths n with n + 1 a power of two take the FFT kernel on every axis:
    FP32 against the FP64 oracle at 1e-5, FP64 at 1e-12, out of place and in place."""
    rng = np.random.default_rng(sum(shape) + 7)
    n1, n2, n3 = shape
    b = rng.standard_normal((n3, n1, n2))
    ref = orc.dst3d_solve(b.astype(np.float32).astype(np.float64), h)
    d = whhd.CuDst3D(shape, h)
    phi = torch.empty((n3, n1, n2), device="cuda")
    d.Solve(phi, cuda(b))
    assert np.abs(host(phi) - ref).max() <= 1e-5 * np.abs(ref).max()
    t = cuda(b)
    d.Solve(t, t)
    assert np.abs(host(t) - ref).max() <= 1e-5 * np.abs(ref).max()
    ref64 = orc.dst3d_solve(b, h)
    d64 = whhd.CuDst3D(shape, h, dtype=torch.float64)
    phi64 = torch.empty((n3, n1, n2), device="cuda", dtype=torch.float64)
    d64.Solve(phi64, cuda64(b))
    assert np.abs(host(phi64) - ref64).max() <= 1e-12 * np.abs(ref64).max()


@pytest.mark.parametrize("shape,h", [((15, 15), (0.1, 0.1)), ((127, 33), (0.01, 0.03)), ((1, 7), (0.5, 0.2)),
                                     ((255, 63), (0.004, 0.02))])
def test_dst2d_solve(whhd, orc, shape, h):
    """CuDst (2D) against the oracle, out of place and in place."""
    rng = np.random.default_rng(sum(shape) + 1)
    b = rng.standard_normal(shape).astype(np.float32).astype(np.float64)
    d = whhd.CuDst(shape, h)
    phi = torch.empty(shape, device="cuda")
    d.Solve(phi, cuda(b))
    ref = orc.dst2d_solve(b, h)
    assert np.abs(host(phi) - ref).max() <= 1e-5 * np.abs(ref).max()
    t = cuda(b)
    d.Solve(t, t)
    assert np.abs(host(t) - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("bc", ["d", "z"])
def test_sweep_and_project_velocity_bc_d(whhd, orc, bc):
    """Velocity bc 'd' ('z' takes the same branches): the swept potential against
    the oracle, then ProjectD against the oracle and its interior node
    divergence."""
    n = (24, 20, 28)
    h = (0.04, 0.05, 0.035)
    nx, ny, nz = n
    rng = np.random.default_rng(21)
    q0 = [rng.standard_normal(s) for s in ((nz + 1, nx, ny + 1), (nz + 1, nx + 1, ny),
                                           (nz, nx + 1, ny + 1))]
    u = [a.astype(np.float32).astype(np.float64) for a in I.curl3d(*q0, h)]
    sw = whhd.CuParallelSweep3D(n, h, bc)
    q = [torch.full(s, np.nan, device="cuda") for s in ((nz + 1, nx, ny + 1), (nz + 1, nx + 1, ny),
                                                        (nz, nx + 1, ny + 1))]
    sw.Solve(*q, *(cuda(a) for a in u))
    ref = orc.sweep3d_d(*u, n, h)
    for g, r in zip(q, ref):
        g = host(g)
        assert np.isfinite(g).all()
        assert np.abs(g - r).max() <= 1e-5 * max(np.abs(r).max(), 1e-30)
    sw.Project(*q)
    pref = orc.project_d(*ref, n, h)
    scale = max(np.abs(r).max() for r in pref)
    got = [host(g) for g in q]
    for g, r in zip(got, pref):
        assert np.abs(g - r).max() <= 1e-4 * scale
    qx, qy, qz = got
    div = ((qx[1:-1, 1:, 1:-1] - qx[1:-1, :-1, 1:-1]) / h[0]
           + (qy[1:-1, 1:-1, 1:] - qy[1:-1, 1:-1, :-1]) / h[1]
           + (qz[1:, 1:-1, 1:-1] - qz[:-1, 1:-1, 1:-1]) / h[2])
    assert np.abs(div).max() <= 1e-4 * scale / min(h)


@pytest.mark.parametrize("bc", ["n", "d", "z"])
def test_sweep2d(whhd, orc, bc):
    """The 2D sweep (CuParallelSweep3D(int2, ...).Solve(qz, xv, yv)) against the oracle."""
    n = (64, 48)
    nx, ny = n
    dx = (1 / 64, 1 / 48)
    rng = np.random.default_rng(30)
    psi = rng.standard_normal((nx + 1, ny + 1))
    u = [a[0].astype(np.float32).astype(np.float64) for a in orc.curl2d(psi, n, dx)]
    sw = whhd.CuParallelSweep3D(n, dx, bc)
    qz = torch.full((nx + 1, ny + 1), np.nan, device="cuda")
    sw.Solve(qz, cuda(u[0]), cuda(u[1]))
    ref = orc.sweep2d(u[0], u[1], n, dx, bc)
    g = host(qz)
    assert np.isfinite(g).all()
    assert np.abs(g - ref).max() <= 1e-5 * np.abs(ref).max()


def cuda64(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("bc", ["n", "d"])
def test_fp64_comparator_matches_oracle(whhd, orc, bc):
    """The FP64 comparator (the reference's USE_DOUBLE build) against the FP64
    oracle at 1e-12: sweep, projection, and the 3D and 2D DST solves."""
    n = (24, 20, 28)
    nx, ny, nz = n
    h = (0.04, 0.05, 0.035)
    rng = np.random.default_rng(40)
    if bc == "n":
        q0 = I.spectral_potential3d(n, 4, I.seed_for(4, 41), tuple(a * b for a, b in zip(n, h)))
    else:
        q0 = [rng.standard_normal(s) for s in ((nz + 1, nx, ny + 1), (nz + 1, nx + 1, ny),
                                               (nz, nx + 1, ny + 1))]
    u = I.curl3d(*q0, h)
    sw = whhd.CuParallelSweep3D(n, h, bc, dtype=torch.float64)
    q = [torch.full(s, np.nan, device="cuda", dtype=torch.float64)
         for s in ((nz + 1, nx, ny + 1), (nz + 1, nx + 1, ny), (nz, nx + 1, ny + 1))]
    sw.Solve(*q, *(cuda64(a) for a in u))
    ref = orc.sweep3d(u[0], u[1], n, h) if bc == "n" else orc.sweep3d_d(*u, n, h)
    for g, r in zip(q, ref):
        assert np.abs(host(g) - r).max() <= 1e-12 * np.abs(r).max()
    sw.Project(*q)
    pref = orc.project_zero(*ref, n, h) if bc == "n" else orc.project_d(*ref, n, h)
    scale = max(np.abs(r).max() for r in pref)
    for g, r in zip(q, pref):
        assert np.abs(host(g) - r).max() <= 1e-11 * scale
    with pytest.raises(TypeError):
        sw.Solve(*[a.float() for a in q], *(cuda(a) for a in u))
    b = rng.standard_normal((7, 9, 5))
    d3 = whhd.CuDst3D((9, 5, 7), (0.1, 0.2, 0.15), dtype=torch.float64)
    phi = torch.empty((7, 9, 5), device="cuda", dtype=torch.float64)
    d3.Solve(phi, cuda64(b))
    r3 = orc.dst3d_solve(b, (0.1, 0.2, 0.15))
    assert np.abs(host(phi) - r3).max() <= 1e-12 * np.abs(r3).max()
    b2 = rng.standard_normal((33, 17))
    d2 = whhd.CuDst((33, 17), (0.03, 0.07), dtype=torch.float64)
    phi2 = torch.empty((33, 17), device="cuda", dtype=torch.float64)
    d2.Solve(phi2, cuda64(b2))
    r2 = orc.dst2d_solve(b2, (0.03, 0.07))
    assert np.abs(host(phi2) - r2).max() <= 1e-12 * np.abs(r2).max()
    s2 = whhd.CuParallelSweep3D((32, 24), (1 / 32, 1 / 24), bc, dtype=torch.float64)
    psi = rng.standard_normal((33, 25))
    v2 = [a[0] for a in orc.curl2d(psi, (32, 24), (1 / 32, 1 / 24))]
    qz = torch.full((33, 25), np.nan, device="cuda", dtype=torch.float64)
    s2.Solve(qz, cuda64(v2[0]), cuda64(v2[1]))
    rz = orc.sweep2d(v2[0], v2[1], (32, 24), (1 / 32, 1 / 24), bc)
    assert np.abs(host(qz) - rz).max() <= 1e-12 * np.abs(rz).max()
