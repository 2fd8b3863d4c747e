"""Pins of the solid-geometry oracle by closed forms and an independent solver.

* level set: sphere distances at hand-computable cells;
* face fractions: for a planar interface the fraction is the exact wet length
  of the face (0/1 for faces parallel to the interface, 1/2 + s/h for faces
  it crosses at signed distance s from the face centre);
* eikonal update: the 1-, 2- and 3-axis closed forms u = m + h,
  m + h/sqrt(2), m + h/sqrt(3) for equal neighbour minima;
* the relaxation keeps an axis-aligned signed distance unchanged and
  converges to the first-order fast-marching solution (computed here with a
  heap, independently of the oracle) on a sphere.
"""
import heapq
import math

import numpy as np
import pytest


def test_levelset_sphere_known_values(orc):
    n, h = (8, 8, 8), 0.125
    c = [(3.5 * h, 4.5 * h, 2.5 * h)]  # centre of cell (x=3, y=4, z=2)
    ls = orc.levelset_spheres(n, (h,) * 3, c, [0.2])
    assert ls[2, 3, 4] == pytest.approx(-0.2, abs=1e-15)
    for k in range(1, 4):
        assert ls[2, 3 + k, 4] == pytest.approx(k * h - 0.2, abs=1e-14)   # along x
        assert ls[2 + k, 3, 4] == pytest.approx(k * h - 0.2, abs=1e-14)   # along z
    assert ls[2, 4, 5] == pytest.approx(math.sqrt(2) * h - 0.2, abs=1e-14)
    # a union takes the nearer surface
    two = orc.levelset_spheres(n, (h,) * 3, c + [(0.9, 0.9, 0.9)], [0.2, 0.1])
    far = orc.levelset_spheres(n, (h,) * 3, [(0.9, 0.9, 0.9)], [0.1])
    assert np.array_equal(two, np.minimum(ls, far))


@pytest.mark.parametrize("s", [-0.7, -0.5, -0.3, -0.1, 0.0, 0.2, 0.45, 0.6])
def test_frac_of_planar_interface(orc, s):
    """ls = y - y0: x-faces are cut by the plane, wet length (1/2 + s/h) h."""
    nx, ny, nz, h = 4, 6, 3, 0.25
    y = (np.arange(ny) + 0.5) * h
    y0 = y[2] - s * h  # cell row 2 centre sits s*h above the plane
    ls = np.broadcast_to(y - y0, (nz, nx, ny)).copy()
    ax, ay, az = orc.frac_from_ls(ls, (h, h, h))
    want = np.clip(0.5 + (y - y0) / h, 0.0, 1.0)
    assert np.allclose(ax, np.broadcast_to(want, ax.shape), atol=1e-12)
    assert np.allclose(az, np.broadcast_to(want, az.shape), atol=1e-12)
    # y-faces are parallel to the plane: fully wet above it, dry below
    # (a face lying in the plane itself is half wet)
    yf = np.arange(1, ny) * h
    on = np.isclose(yf, y0, rtol=0, atol=1e-12)
    want_y = np.where(on, 0.5, (yf > y0).astype(float))
    assert np.array_equal(ay[:, :, 1:ny], np.broadcast_to(want_y, (nz, nx, ny - 1)))


def test_eikonal_update_closed_forms(orc):
    h, m, big = 0.1, 0.3, 1e8
    assert orc.eikonal_update(m, big, big, h, h, h) == pytest.approx(m + h, abs=1e-15)
    assert orc.eikonal_update(big, m, m, h, h, h) == pytest.approx(m + h / math.sqrt(2), abs=1e-14)
    assert orc.eikonal_update(m, m, m, h, h, h) == pytest.approx(m + h / math.sqrt(3), abs=1e-14)
    # the update solves sum over used axes of ((u - m_i)/h_i)^2 = 1
    mx, my, mz, hx, hy, hz = 0.31, 0.35, 0.38, 0.1, 0.12, 0.09
    u = orc.eikonal_update(mx, my, mz, hx, hy, hz)
    used = [(mi, hi) for mi, hi in ((mx, hx), (my, hy), (mz, hz)) if mi < u]
    assert len(used) == 3
    assert sum(((u - mi) / hi) ** 2 for mi, hi in used) == pytest.approx(1.0, abs=1e-12)


def test_sweep_keeps_planar_distance(orc):
    nx, ny, nz, h = 10, 6, 5, 0.1
    x = (np.arange(nx) + 0.5) * h
    ls = np.broadcast_to((x - 0.25)[:, None], (nz, nx, ny)).copy()  # x = 0, 1 solid
    out = orc.eikonal_sweep(ls, h, 12)
    assert np.allclose(out, ls, atol=1e-13)


def fast_marching(ls, h):
    """First-order fast marching from the fixed solid cells (ls <= -1e-12),
    with the same upwind update; neighbours outside the grid do not exist."""
    nz, nx, ny = ls.shape
    u = np.full(ls.shape, np.inf)
    done = ls <= -1e-12
    u[done] = ls[done]
    heap = []

    def update(k, i, j):
        ms = []
        for a, b in (((k, i - 1, j), (k, i + 1, j)), ((k, i, j - 1), (k, i, j + 1)),
                     ((k - 1, i, j), (k + 1, i, j))):
            vals = [u[p] for p in (a, b) if 0 <= p[0] < nz and 0 <= p[1] < nx and 0 <= p[2] < ny
                    and done[p]]
            ms.append(min(vals) if vals else np.inf)
        ms.sort()
        v = ms[0] + h
        if v > ms[1]:
            v = 0.5 * (ms[0] + ms[1] + math.sqrt(2 * h * h - (ms[0] - ms[1]) ** 2))
            if v > ms[2]:
                s = ms[0] + ms[1] + ms[2]
                q = sum(m * m for m in ms)
                v = (s + math.sqrt(s * s - 3 * (q - h * h))) / 3
        return v

    def push_neighbours(k, i, j):
        for dk, di, dj in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            p = (k + dk, i + di, j + dj)
            if 0 <= p[0] < nz and 0 <= p[1] < nx and 0 <= p[2] < ny and not done[p]:
                v = update(*p)
                if v < u[p]:
                    u[p] = v
                    heapq.heappush(heap, (v, p))

    for p in zip(*np.nonzero(done)):
        push_neighbours(*p)
    while heap:
        v, p = heapq.heappop(heap)
        if done[p] or v > u[p]:
            continue
        done[p] = True
        push_neighbours(*p)
    return u


def test_sweep_converges_to_fast_marching(orc):
    n, h = (10, 9, 8), 0.1
    ls = orc.levelset_spheres(n, (h,) * 3, [(0.42, 0.47, 0.38)], [0.22])
    got = orc.eikonal_sweep(ls, h, 80)
    want = fast_marching(ls, h)
    assert np.allclose(got, want, atol=1e-12)
    # and it is not the input: the sphere SDF is not a fixed point of the scheme
    assert np.abs(got - ls).max() > 1e-3
