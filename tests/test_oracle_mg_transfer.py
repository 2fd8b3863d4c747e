"""Pins of the oracle's multigrid transfer operators and of the level cut, each
against what the geometry or the wavelet layout fixes, independently of the
formulas the oracle uses.

* cell restriction (CuRestrict3D::SolveP1RN with spacings,
  WaveletPotentials/cuRestrict3D.cu:801-970): the spacing-weighted average of a
  field linear in the cell-centre coordinates is that field at the coarse cell's
  centroid, for unequal spacings and odd extents (lone last child);
* hierarchy spacings (cuMultigridFrac2D.cu:4-26 via cuMultigridFrac3D.cu:9-134):
  every coarse cell boundary is a fine cell boundary, and the last one is the
  domain wall;
* prolongation (CuProlongate3D::SolveAddP1RN, cuProlongate3D.cu:353-423,
  594-616): each fine cell receives the value of the coarse cell that contains
  its centre, added to what it holds; restriction after prolongation is the
  identity;
* face-fraction restriction (SolveP1Ax/Ay/Az, cuRestrict3D.cu:972-1140): with the
  exact wet-area fractions of a tilted planar interface on the fine faces, every
  coarse face holds the exact wet-area fraction of its own face (all levels on
  power-of-two grids, level 1 on odd extents, where the last coarse face on the
  normal axis is the domain wall); uniform fractions stay uniform;
* level cut (CutCoefficients / cuCutCoefs, cuWaveletPotentialRecover3D.cu:123-239):
  a potential whose coefficients hold no detail in the `cut` finest bands comes
  back unchanged, one coefficient in those bands (on any axis of any component)
  is removed entirely, the last coefficient below the bands is kept, and the
  cut is idempotent.
"""
import numpy as np
import pytest


# ---- geometry helpers --------------------------------------------------------
def bounds(h):
    return np.concatenate([[0.0], np.cumsum(h)])


def centres(h):
    b = bounds(h)
    return 0.5 * (b[:-1] + b[1:])


def int_clamp(alpha, beta, lo, hi, t0, t1):
    """Exact integral over [t0, t1] of clamp(alpha + beta t, lo, hi)."""
    pts = [t0, t1]
    if beta != 0.0:
        for v in (lo, hi):
            t = (v - alpha) / beta
            if t0 < t < t1:
                pts.append(t)
    pts.sort()
    f = lambda t: min(max(alpha + beta * t, lo), hi)  # noqa: E731
    return sum(0.5 * (f(a) + f(b)) * (b - a) for a, b in zip(pts, pts[1:]))


# solid below the plane y = G0 + GX x + GZ z; fluid above
G0, GX, GZ = 0.07, 0.21, 0.13


def exact_fractions(hx, hy, hz):
    """Wet-area fractions of every face of the grid with these spacings."""
    bx, by, bz = bounds(hx), bounds(hy), bounds(hz)
    nx, ny, nz = len(hx), len(hy), len(hz)
    ax = np.empty((nz, nx + 1, ny))
    ay = np.empty((nz, nx, ny + 1))
    az = np.empty((nz + 1, nx, ny))
    for k in range(nz):
        for i in range(nx + 1):
            for j in range(ny):
                W = by[j + 1] - by[j]
                a = int_clamp(by[j + 1] - G0 - GX * bx[i], -GZ, 0.0, W, bz[k], bz[k + 1])
                ax[k, i, j] = a / (W * (bz[k + 1] - bz[k]))
    for k in range(nz):
        for i in range(nx):
            for j in range(ny + 1):
                W = bz[k + 1] - bz[k]
                a = int_clamp((by[j] - G0) / GZ - bz[k], -GX / GZ, 0.0, W, bx[i], bx[i + 1])
                ay[k, i, j] = a / (W * (bx[i + 1] - bx[i]))
    for k in range(nz + 1):
        for i in range(nx):
            for j in range(ny):
                W = by[j + 1] - by[j]
                a = int_clamp(by[j + 1] - G0 - GZ * bz[k], -GX, 0.0, W, bx[i], bx[i + 1])
                az[k, i, j] = a / (W * (bx[i + 1] - bx[i]))
    return ax, ay, az


def test_int_clamp_against_quadrature():
    rng = np.random.default_rng(1)
    for _ in range(20):
        al, be = rng.uniform(-1, 1, 2)
        t = np.linspace(-0.7, 1.3, 200001)
        v = np.clip(al + be * t, 0.0, 0.4)
        assert abs(int_clamp(al, be, 0.0, 0.4, -0.7, 1.3) - np.trapezoid(v, t)) < 1e-8


# ---- cell restriction and spacings -------------------------------------------
@pytest.mark.parametrize("n", [(8, 6, 4), (7, 5, 9), (1, 3, 2), (11, 4, 3)])
def test_restriction_of_linear_field_is_the_centroid_value(orc, n):
    nx, ny, nz = n
    rng = np.random.default_rng(sum(n))
    hx, hy, hz = (rng.uniform(0.3, 2.0, k) for k in n)
    cx, cy, cz = centres(hx), centres(hy), centres(hz)
    a, b, c, d = 0.4, 1.3, -0.7, 2.1
    f = a + b * cx[None, :, None] + c * cy[None, None, :] + d * cz[:, None, None]
    got = orc.mg_restrict(f, hx, hy, hz)
    # coarse cells are unions of one or two consecutive fine cells
    coarse = [np.add.reduceat(h, np.arange(0, len(h), 2)) for h in (hx, hy, hz)]
    Cx, Cy, Cz = (centres(h) for h in coarse)
    want = a + b * Cx[None, :, None] + c * Cy[None, None, :] + d * Cz[:, None, None]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [(16, 8, 16), (13, 7, 10)])
def test_hierarchy_cells_nest_in_the_fine_grid(orc, n):
    nx, ny, nz = n
    h = (0.1, 0.07, 0.05)
    fr = [np.ones(s) for s in ((nz, nx + 1, ny), (nz, nx, ny + 1), (nz + 1, nx, ny))]
    fine = [bounds(np.full(k, hk)) for k, hk in zip(n, h)]
    for level in range(len(orc.mg_levels(n))):
        *_, hx, hy, hz = orc.mg_level(*fr, h, level)
        for axis, hl in enumerate((hx, hy, hz)):
            idx = np.minimum(np.arange(len(hl) + 1) * 2 ** level, n[axis])
            np.testing.assert_allclose(bounds(hl), fine[axis][idx], rtol=0, atol=1e-14)


# ---- prolongation ------------------------------------------------------------
@pytest.mark.parametrize("n", [(8, 6, 4), (7, 5, 9), (3, 1, 2)])
def test_prolongation_gives_each_cell_its_parent_value(orc, n):
    nx, ny, nz = n
    rng = np.random.default_rng(3)
    fine0 = rng.standard_normal((nz, nx, ny))
    coarse = rng.standard_normal(((nz + 1) // 2, (nx + 1) // 2, (ny + 1) // 2))
    got = orc.mg_prolong_add(fine0, coarse)
    # parent = the coarse cell whose extent holds the fine cell centre (unit spacing)
    par = [np.searchsorted(bounds(np.full((k + 1) // 2, 2.0)), centres(np.ones(k))) - 1
           for k in (nx, ny, nz)]
    want = fine0 + coarse[np.ix_(par[2], par[0], par[1])]
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n", [(8, 6, 4), (7, 5, 9), (9, 3, 6)])
def test_restriction_after_prolongation_is_identity(orc, n):
    nx, ny, nz = n
    rng = np.random.default_rng(4)
    hx, hy, hz = (rng.uniform(0.3, 2.0, k) for k in n)
    coarse = rng.standard_normal(((nz + 1) // 2, (nx + 1) // 2, (ny + 1) // 2))
    fine = orc.mg_prolong_add(np.zeros((nz, nx, ny)), coarse)
    np.testing.assert_allclose(orc.mg_restrict(fine, hx, hy, hz), coarse, rtol=0, atol=1e-13)


# ---- face fractions ----------------------------------------------------------
@pytest.mark.parametrize("normal", [0, 1, 2])
def test_uniform_face_fractions_stay_uniform(orc, normal):
    ext = [9, 6, 5]
    ext[normal] += 1
    src = np.full((ext[2], ext[0], ext[1]), 0.37)
    np.testing.assert_array_equal(orc.mg_restrict_faces(src, normal), 0.37)


@pytest.mark.parametrize("n,levels", [((16, 8, 16), None), ((13, 7, 11), [1]),
                                      ((10, 9, 12), [1])])
def test_face_restriction_keeps_exact_wet_fractions(orc, n, levels):
    """A tilted planar interface: the fine faces hold exact wet-area fractions,
    and so must every coarse face.  On odd extents only level 1 is exact: from
    level 2 on a coarse face spans children of unequal width (the lone last
    cell), and the reference averages child fractions without area weights."""
    nx, ny, nz = n
    h = (1.0 / 16, 0.5 / 8, 1.0 / 16)
    fine = exact_fractions(np.full(nx, h[0]), np.full(ny, h[1]), np.full(nz, h[2]))
    assert all(0.05 < a.mean() < 0.95 for a in fine)   # the interface crosses every face set
    if levels is None:
        levels = range(1, len(orc.mg_levels(n)))
    for level in levels:
        gx, gy, gz, hx, hy, hz = orc.mg_level(*fine, h, level)
        wx, wy, wz = exact_fractions(hx, hy, hz)
        for got, want in ((gx, wx), (gy, wy), (gz, wz)):
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("normal", [0, 1, 2])
def test_last_coarse_face_is_the_wall_face(orc, normal):
    """On an odd extent the last coarse face on the normal axis lies on the
    domain wall, so it takes the fine wall face (after tangential averaging)."""
    ext = [7, 5, 9]
    ext[normal] += 1          # face array: odd cell count + 1 faces
    rng = np.random.default_rng(9)
    src = rng.uniform(0, 1, (ext[2], ext[0], ext[1]))
    got = orc.mg_restrict_faces(src, normal)
    ax_np = {0: 1, 1: 2, 2: 0}[normal]          # numpy axis of the normal
    wall = np.take(src, [-1], axis=ax_np)
    # tangential averaging of the wall plane: pairs, lone last child kept
    for a in range(3):
        if a == ax_np:
            continue
        k = wall.shape[a]
        idx = np.arange(0, k, 2)
        s = np.add.reduceat(wall, idx, axis=a)
        cnt = np.minimum(2, k - idx).astype(float)
        shp = [1, 1, 1]
        shp[a] = len(idx)
        wall = s / cnt.reshape(shp)
    np.testing.assert_allclose(np.take(got, [-1], axis=ax_np), wall, rtol=0, atol=1e-15)


# ---- level cut ---------------------------------------------------------------
# (type, bc) per axis of each potential component for velocity bc 'n': the
# component's own axis is the curl axis (CDF3.7, 'n'), the others the divergence
# axes (CDF4.6, 'd', one extra sample; ext 1
# axes), with (type, bc, ext, first live coefficient): the antisymmetric 'd'
# lines have zero end samples, so their first two coefficients carry no
# independent content and potentials with zero wall-tangential values are
# exactly those whose coefficients vanish there
def _layout(n):
    nx, ny, nz = n
    shapes = [(nz + 1, nx, ny + 1), (nz + 1, nx + 1, ny), (nz, nx + 1, ny + 1)]
    axes = []
    for comp in range(3):
        axes.append([("cdf3.7", "n", 0, 0) if a == comp else ("cdf4.6", "d", 1, 2)
                     for a in range(3)])
    return shapes, axes


_NP_AXIS = {0: 1, 1: 2, 2: 0}     # x, y, z -> numpy axis of a (z, x, y) array
_CH = "xyz"


def _synth(orc, coef, n, comp):
    """Potential component whose wavelet coefficients are `coef`."""
    _, axes = _layout(n)
    q = coef
    for a in range(3):
        wt, bc, _, _ = axes[comp][a]
        q, _ = orc.transform(q, int(np.log2(n[a])), _CH[a], bc, wt, forward=False)
    return q


def _kept_mask(n, comp, cut):
    """True where every axis index lies below the `cut` finest detail bands."""
    shapes, axes = _layout(n)
    m = np.ones(shapes[comp], dtype=bool)
    for a in range(3):
        _, _, ext, live = axes[comp][a]
        i = np.arange(shapes[comp][_NP_AXIS[a]])
        keep = (i >= live) & (i < 2 ** (int(np.log2(n[a])) - cut) + ext)
        sh = [1, 1, 1]
        sh[_NP_AXIS[a]] = -1
        m = m & keep.reshape(sh)
    return m


CUT_CASES = [((16, 16, 16), 1), ((32, 16, 32), 2), ((32, 32, 16), 3), ((16, 32, 32), 2)]


@pytest.mark.parametrize("n,cut", CUT_CASES)
def test_cut_keeps_a_potential_without_fine_details(orc, n, cut):
    shapes, _ = _layout(n)
    rng = np.random.default_rng(cut)
    q = []
    for c in range(3):
        coef = rng.standard_normal(shapes[c]) * _kept_mask(n, c, cut)
        q.append(_synth(orc, coef, n, c))
    out = orc.cut3d(*q, n, cut)
    for a, b in zip(q, out):
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-11 * np.abs(a).max())


@pytest.mark.parametrize("n,cut", CUT_CASES)
def test_cut_removes_any_single_fine_detail(orc, n, cut):
    """One coefficient on the first and on the last index of the cut bands, on
    every axis of every component, with the other indices in the kept block."""
    shapes, axes = _layout(n)
    rng = np.random.default_rng(7)
    for comp in range(3):
        for a in range(3):
            ext = axes[comp][a][2]
            L = int(np.log2(n[a]))
            for band_index in (2 ** (L - cut) + ext, shapes[comp][_NP_AXIS[a]] - 1):
                coef = np.zeros(shapes[comp])
                pos = [int(rng.integers(axes[comp][b][3],
                                        2 ** (int(np.log2(n[b])) - cut) + axes[comp][b][2]))
                       for b in range(3)]
                pos[a] = band_index
                coef[pos[2], pos[0], pos[1]] = 1.0
                q = [np.zeros(s) for s in shapes]
                q[comp] = _synth(orc, coef, n, comp)
                out = orc.cut3d(*q, n, cut)
                scale = np.abs(q[comp]).max()
                for o in out:
                    assert np.abs(o).max() < 1e-11 * scale, (comp, a, band_index)


@pytest.mark.parametrize("n,cut", CUT_CASES)
def test_cut_keeps_the_last_coefficient_below_the_bands(orc, n, cut):
    shapes, axes = _layout(n)
    for comp in range(3):
        for a in range(3):
            ext = axes[comp][a][2]
            coef = np.zeros(shapes[comp])
            pos = [axes[comp][b][3] for b in range(3)]
            pos[a] = 2 ** (int(np.log2(n[a])) - cut) + ext - 1
            coef[pos[2], pos[0], pos[1]] = 1.0
            q = [np.zeros(s) for s in shapes]
            q[comp] = _synth(orc, coef, n, comp)
            out = orc.cut3d(*q, n, cut)
            np.testing.assert_allclose(out[comp], q[comp], rtol=0,
                                       atol=1e-11 * np.abs(q[comp]).max())


@pytest.mark.parametrize("n,cut", CUT_CASES[:2])
def test_cut_is_idempotent(orc, n, cut):
    shapes, _ = _layout(n)
    rng = np.random.default_rng(11)
    q = [_synth(orc, rng.standard_normal(s) * _kept_mask(n, c, 0), n, c)
         for c, s in enumerate(shapes)]
    once = orc.cut3d(*q, n, cut)
    twice = orc.cut3d(*once, n, cut)
    for a, b in zip(once, twice):
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-11 * np.abs(a).max())
