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


# ---- supersampled fractions (the static GetFrac) and their samplers --------

def _grid_fn(n, f):
    """Values f(x, y, z) at the cell-centred grid nodes of an (nx, ny, nz) grid,
    in the (nz, nx, ny) layout."""
    nx, ny, nz = n
    z, x, y = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    return f(x.astype(float), y.astype(float), z.astype(float))


def test_linear_sampler_reproduces_trilinear_and_clamps(orc):
    """A trilinear function is reproduced exactly inside the grid; outside,
    cuGetIxFx clamps a coordinate below 1e-5 to 1e-5 and one at or beyond the
    last node to (n - 2) + (1 - 1e-5)."""
    n = (5, 7, 6)
    f = lambda x, y, z: 0.3 - 1.1 * x + 2.0 * y + 0.7 * z + 0.4 * x * y - 0.2 * y * z + 0.9 * x * z - 0.05 * x * y * z  # noqa: E731
    v = _grid_fn(n, f)
    rng = np.random.default_rng(5)
    for _ in range(50):
        g = rng.uniform([0.01, 0.01, 0.01], [n[0] - 1.01, n[1] - 1.01, n[2] - 1.01])
        assert orc.interp_linear3d(v, g) == pytest.approx(f(*g), abs=1e-12)
    assert orc.interp_linear3d(v, (-3.0, 2.5, 1.25)) == pytest.approx(f(1e-5, 2.5, 1.25), abs=1e-12)
    top = n[0] - 2 + (1 - 1e-5)
    for x in (n[0] - 1.0, n[0] + 4.0):
        assert orc.interp_linear3d(v, (x, 2.5, 1.25)) == pytest.approx(f(top, 2.5, 1.25), abs=1e-12)


def test_quadratic_sampler_is_the_quadratic_bspline(orc):
    """Inside the grid the quadratic B-spline reproduces linear functions and
    maps x^2 to x^2 + 1/4 (the uniform quadratic B-spline has variance 3/12),
    per axis and for products of axes."""
    n = (8, 9, 10)
    lin = lambda x, y, z: 1.5 - 0.5 * x + 0.25 * y + 2.0 * z  # noqa: E731
    vl = _grid_fn(n, lin)
    vx2 = _grid_fn(n, lambda x, y, z: x * x)
    vy2 = _grid_fn(n, lambda x, y, z: y * y)
    vxz = _grid_fn(n, lambda x, y, z: x * z)
    rng = np.random.default_rng(6)
    for _ in range(50):
        g = rng.uniform(0.5, [n[0] - 1.51, n[1] - 1.51, n[2] - 1.51])
        assert orc.interp_quadratic3d(vl, g) == pytest.approx(lin(*g), abs=1e-12)
        assert orc.interp_quadratic3d(vx2, g) == pytest.approx(g[0] ** 2 + 0.25, abs=1e-12)
        assert orc.interp_quadratic3d(vy2, g) == pytest.approx(g[1] ** 2 + 0.25, abs=1e-12)
        assert orc.interp_quadratic3d(vxz, g) == pytest.approx(g[0] * g[2], abs=1e-12)


@pytest.mark.parametrize("t128,sign", [(37.3, 1), (37.3, -1), (100.3, 1), (5.3, -1)])
def test_frac_divide_axis_aligned_plane(orc, t128, sign):
    """ls = sign * (y - y0): faces normal to y are 0 or 1 by the side of their
    plane; faces normal to x and z crossed by y = y0 at fraction t of their
    height get the midpoint-rule wet share of 128 rows, within 0.5/128 of the
    exact share (t sits 0.3 of a row above a row boundary, so a half-row shift
    of the sub-samples fails this)."""
    n, h = (6, 12, 7), 0.25
    row = 5
    y0 = (row + t128 / 128) * h
    ls = _grid_fn(n, lambda x, y, z: sign * ((y + 0.5) * h - y0))
    ax, ay, az = orc.frac_divide(ls, (h, h, h))
    nx, ny, nz = n
    ylo = np.arange(ny) * h
    wet = np.clip((ylo + h - y0) / h, 0, 1) if sign > 0 else np.clip((y0 - ylo) / h, 0, 1)
    for a in (ax, az):  # (nz', nx', ny): the y axis is last
        assert np.all(np.abs(a - wet[None, None, :]) <= 0.5 / 128 + 1e-12)
        crossed = a[..., row]
        assert np.all(np.abs(crossed * 16384 - np.round(crossed * 16384)) < 1e-9)
    yface = np.arange(ny + 1) * h
    exp_ay = (yface > y0) if sign > 0 else (yface < y0)
    assert np.array_equal(ay, np.broadcast_to(exp_ay.astype(float), ay.shape))


def _halfplane_area(y0, y1, z0, z1, a, b, d):
    """Exact area of {(y, z) in [y0,y1]x[z0,z1] : a*y + b*z > d} by clipping
    the rectangle with the half-plane and the shoelace formula."""
    poly = [(y0, z0), (y1, z0), (y1, z1), (y0, z1)]
    out = []
    for i in range(4):
        p, q = poly[i], poly[(i + 1) % 4]
        fp, fq = a * p[0] + b * p[1] - d, a * q[0] + b * q[1] - d
        if fp > 0:
            out.append(p)
        if (fp > 0) != (fq > 0):
            s = fp / (fp - fq)
            out.append((p[0] + s * (q[0] - p[0]), p[1] + s * (q[1] - p[1])))
    if len(out) < 3:
        return 0.0
    return 0.5 * abs(sum(out[i][0] * out[(i + 1) % len(out)][1] - out[(i + 1) % len(out)][0] * out[i][1]
                         for i in range(len(out))))


@pytest.mark.parametrize("theta", [0.3, 0.9, 1.2])
def test_frac_divide_tilted_plane(orc, theta):
    """A plane tilted in (y, z): on interior x-faces (no sampler clamping) the
    fraction is the exact wet area of the face up to the 128 x 128 midpoint
    rule (1/128); faces wholly on one side are exactly 0 or 1."""
    n, h = (4, 10, 10), 0.1
    a, b = math.cos(theta), math.sin(theta)
    d = a * 0.5 + b * 0.47
    ls = _grid_fn(n, lambda x, y, z: a * (y + 0.5) * h + b * (z + 0.5) * h - d)
    ax, _, _ = orc.frac_divide(ls, (h, h, h))
    nx, ny, nz = n
    checked = 0
    for z in range(1, nz - 1):
        for y in range(1, ny - 1):
            exact = _halfplane_area(y * h, (y + 1) * h, z * h, (z + 1) * h, a, b, d) / (h * h)
            for x in range(1, nx):
                got = ax[z, x, y]
                assert abs(got - exact) <= 1.0 / 128, (z, x, y, got, exact)
                if exact in (0.0, 1.0):
                    assert got == exact
                checked += 0 < exact < 1
    assert checked >= 10
