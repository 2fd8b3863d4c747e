"""Pins of the fractional multigrid oracle.

* hierarchy: 192 -> 96 -> ... -> 3 -> 2 (ceil halving until a side reaches 1);
* operator: with open faces it is the 7-point Laplacian (A x^2 = -2 in the
  interior), the 'd' wall puts the boundary value on the face (A x = 0 in the
  wall cell), and a face fraction scales exactly that face's flux;
* the Jacobi sweep is v + w D^-1 (f - A v) for the same operator;
* the V-cycle converges to the dense direct solution of the assembled
  operator (brute force on a small grid), Neumann and Dirichlet, with random
  fractions and an anisotropic, non-power-of-two grid;
* residuals fall monotonically on a sphere problem.
"""
import numpy as np
import pytest


def ones(n):
    nx, ny, nz = n
    return (np.ones((nz, nx + 1, ny)), np.ones((nz, nx, ny + 1)), np.ones((nz + 1, nx, ny)))


def A_apply(orc, v, fr, h, bc):
    return -orc.wj_sweep(v, np.zeros_like(v), *fr, h, bc, residual=True)


def assemble(orc, n, fr, h, bc):
    nx, ny, nz = n
    N = nx * ny * nz
    A = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        A[:, j] = A_apply(orc, e.reshape(nz, nx, ny), fr, h, bc).ravel()
    return A


def test_hierarchy(orc):
    assert orc.mg_levels((192, 192, 192)) == [(192,) * 3, (96,) * 3, (48,) * 3, (24,) * 3,
                                              (12,) * 3, (6,) * 3, (3,) * 3, (2,) * 3]
    assert orc.mg_levels((8, 5, 4)) == [(8, 5, 4), (4, 3, 2)]


@pytest.mark.parametrize("bc", ["n", "d"])
def test_operator_is_the_laplacian(orc, bc):
    n, h = (9, 8, 7), (0.1, 0.1, 0.1)
    nx, ny, nz = n
    x = (np.arange(nx) + 0.5) * 0.1
    v = np.broadcast_to((x ** 2)[:, None], (nz, nx, ny)).copy()
    Av = A_apply(orc, v, ones(n), h, bc)
    # (the Dirichlet walls also act on the y and z boundary cells)
    inner = Av[:, 1:-1, :] if bc == "n" else Av[1:-1, 1:-1, 1:-1]
    assert np.allclose(inner, -2.0, atol=1e-10)


def test_dirichlet_wall_is_on_the_face(orc):
    n, h = (6, 5, 4), (0.2, 0.2, 0.2)
    nx, ny, nz = n
    x = (np.arange(nx) + 0.5) * 0.2
    v = np.broadcast_to(x[:, None], (nz, nx, ny)).copy()    # zero on the x = 0 wall
    Av = A_apply(orc, v, ones(n), h, "d")
    # the wall term 2 v0 / h^2 cancels the interior difference: exact for a linear field
    # (cells away from the y and z walls, which see v there)
    assert np.allclose(Av[1:-1, 0, 1:-1], 0.0, atol=1e-12)
    # ... while the Neumann operator sees a one-sided difference there
    Avn = A_apply(orc, v, ones(n), h, "n")
    assert np.allclose(Avn[:, 0, :], -0.2 / 0.04, atol=1e-12)


def test_fraction_scales_one_face(orc):
    n, h = (5, 4, 3), (0.25, 0.25, 0.25)
    nx, ny, nz = n
    rng = np.random.default_rng(5)
    v = rng.standard_normal((nz, nx, ny))
    full = ones(n)
    part = [a.copy() for a in full]
    part[0][1, 2, 3] = 0.3   # x-face between cells x=1 and x=2 at (y=3, z=1)
    d = A_apply(orc, v, part, h, "n") - A_apply(orc, v, full, h, "n")
    flux = (v[1, 1, 3] - v[1, 2, 3]) / 0.25 ** 2
    want = np.zeros_like(v)
    want[1, 1, 3] = -0.7 * flux
    want[1, 2, 3] = 0.7 * flux
    assert np.allclose(d, want, atol=1e-12)


@pytest.mark.parametrize("bc", ["n", "d"])
def test_jacobi_is_damped_diagonal_step(orc, bc):
    n, h = (5, 6, 4), (0.2, 0.15, 0.25)
    rng = np.random.default_rng(7)
    fr = [rng.uniform(0.2, 1.0, a.shape) for a in ones(n)]
    nx, ny, nz = n
    v, f = rng.standard_normal((2, nz, nx, ny))
    A = assemble(orc, n, fr, h, bc)
    w = 0.8
    got = orc.wj_sweep(v, f, *fr, h, bc, weight=w)
    want = v.ravel() + w * (f.ravel() - A @ v.ravel()) / np.diag(A)
    assert np.allclose(got.ravel(), want, atol=1e-11)


@pytest.mark.parametrize("bc,n", [("d", (6, 5, 4)), ("n", (6, 5, 4)), ("d", (7, 9, 5))])
def test_vcycle_converges_to_direct_solve(orc, bc, n):
    h = (0.3, 0.2, 0.25)
    rng = np.random.default_rng(11)
    fr = [rng.uniform(0.2, 1.0, a.shape) for a in ones(n)]
    nx, ny, nz = n
    A = assemble(orc, n, fr, h, bc)
    f = rng.standard_normal(A.shape[0])
    if bc == "n":
        # singular (constants): project f onto the range, compare modulo the nullspace
        U, S, Vt = np.linalg.svd(A)
        f = f - U[:, -1] * (U[:, -1] @ f)
        p = np.linalg.lstsq(A, f, rcond=None)[0]
    else:
        p = np.linalg.solve(A, f)
    got = orc.mg_solve(f.reshape(nz, nx, ny), *fr, h, bc, n_iters=60, nu1=2, nu2=2).ravel()
    if bc == "n":
        got = got - got.mean() + p.mean()
    assert np.abs(got - p).max() < 1e-9 * np.abs(p).max()


def test_residual_falls_on_sphere_problem(orc):
    n = (16, 16, 16)
    h = (1 / 16,) * 3
    ls = orc.levelset_spheres(n, h, [(0.5, 0.45, 0.55)], [0.2])
    fr = orc.frac_from_ls(ls, h)
    rng = np.random.default_rng(3)
    # f = -div(a u) for a random face field u that is zero on the walls satisfies
    # every solvability condition
    u = [rng.standard_normal(a.shape) for a in fr]
    u[0][:, [0, -1], :] = 0.0     # no flow through the domain walls
    u[1][:, :, [0, -1]] = 0.0
    u[2][[0, -1]] = 0.0
    au = [a * b for a, b in zip(fr, u)]
    f = -((au[0][:, 1:, :] - au[0][:, :-1, :]) / h[0] + (au[1][:, :, 1:] - au[1][:, :, :-1]) / h[1]
          + (au[2][1:] - au[2][:-1]) / h[2])
    res = [orc.mg_rel_residual(orc.mg_solve(f, *fr, h, "n", n_iters=k, nu1=2, nu2=2), f, *fr, h, "n")
           for k in (1, 2, 4, 8)]
    assert all(b < a for a, b in zip(res, res[1:]))
    assert res[-1] < 1e-4
