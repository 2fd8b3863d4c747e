"""Pins of the field-operator oracle.

* a linear pressure has unit gradient on open interior faces; the 'd' wall
  reproduces it (zero wall value on the face), 'n' walls carry none;
* summation by parts: sum over faces of a*u*grad(p) equals sum over cells of
  p * (-div(a u)) when u vanishes on the walls (pins signs and offsets of
  both operators against each other);
* -div(a grad p) is the multigrid operator A p, for 'n' and 'd'.
"""
import numpy as np
import pytest


def fracs(n, rng):
    nx, ny, nz = n
    return [rng.uniform(0.0, 1.0, s) for s in ((nz, nx + 1, ny), (nz, nx, ny + 1), (nz + 1, nx, ny))]


def test_gradient_of_linear_pressure(orc):
    n, h = (6, 5, 4), 0.2
    nx, ny, nz = n
    x = (np.arange(nx) + 0.5) * h
    p = np.broadcast_to(x[:, None], (nz, nx, ny)).copy()
    ones = [np.ones(s) for s in ((nz, nx + 1, ny), (nz, nx, ny + 1), (nz + 1, nx, ny))]
    gx, gy, gz = orc.gradient3d_frac(p, *ones, (h, h, h), "n")
    assert np.allclose(gx[:, 1:nx, :], 1.0, atol=1e-12)
    assert np.all(gx[:, [0, nx], :] == 0.0) and np.all(gy == 0.0) and np.all(gz == 0.0)
    gx, _, _ = orc.gradient3d_frac(p, *ones, (h, h, h), "d")
    assert np.allclose(gx[:, :nx, :], 1.0, atol=1e-12)   # the x = 0 wall face included
    # closed faces carry no gradient
    shut = [a.copy() for a in ones]
    shut[0][1, 3, 2] = 0.0
    gx, _, _ = orc.gradient3d_frac(p, *shut, (h, h, h), "n")
    assert gx[1, 3, 2] == 0.0 and gx[1, 2, 2] == pytest.approx(1.0)


def test_summation_by_parts(orc):
    n, h = (7, 6, 5), (0.3, 0.2, 0.25)
    rng = np.random.default_rng(3)
    fr = fracs(n, rng)
    nx, ny, nz = n
    p = rng.standard_normal((nz, nx, ny))
    u = [rng.standard_normal(a.shape) for a in fr]
    u[0][:, [0, nx], :] = 0.0
    u[1][:, :, [0, ny]] = 0.0
    u[2][[0, nz]] = 0.0
    g = orc.gradient3d_frac(p, *fr, h, "n")
    b = orc.divergence3d_frac(*u, *fr, h)
    lhs = sum(np.sum(a * uu * gg) for a, uu, gg in zip(fr, u, g))
    # faces with fraction <= eps carry no gradient but also no a*u
    rhs = np.sum(p * b)
    assert lhs == pytest.approx(rhs, rel=1e-12)


@pytest.mark.parametrize("bc", ["n", "d"])
def test_div_grad_is_the_multigrid_operator(orc, bc):
    n, h = (6, 5, 7), (0.2, 0.2, 0.2)
    rng = np.random.default_rng(4)
    fr = [np.maximum(a, 0.05) for a in fracs(n, rng)]   # all open: the gradient keeps every face
    p = rng.standard_normal((n[2], n[0], n[1]))
    g = orc.gradient3d_frac(p, *fr, h, bc)
    lap = orc.divergence3d_frac(*g, *fr, h)   # = -div(a grad p)
    Ap = -orc.wj_sweep(p, np.zeros_like(p), *fr, h, bc, residual=True)
    assert np.allclose(lap, Ap, atol=1e-10 * np.abs(Ap).max())
