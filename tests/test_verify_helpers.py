"""Host verification helpers of the reference's tests, exported by the library:
SetSliceZero, StreamInitial::GradientStream (3D and 2D) and CalculateErrorL2
(WaveletPotentials/StreamInitial.cpp:5-175, WaveletPotentials/Utils.cpp:184-331).
Checked against numpy restatements of the same differences and sums and
against the fact that the discrete curl of a gradient-free edge potential is
divergence free."""
import numpy as np

import waveletpotentials_amd as W
from waveletpotentials_amd import inputs as I


def test_gradient_stream3d_is_the_discrete_curl():
    n = (9, 7, 11)
    nx, ny, nz = n
    dx = (0.1, 0.13, 0.07)
    rng = np.random.default_rng(3)
    q = [rng.standard_normal(s).astype(np.float32) for s in
         ((nz + 1, nx, ny + 1), (nz + 1, nx + 1, ny), (nz, nx + 1, ny + 1))]
    v = [np.empty(s, np.float32) for s in ((nz, nx + 1, ny), (nz, nx, ny + 1), (nz + 1, nx, ny))]
    W.StreamInitial.GetInstance().GradientStream(*v, *q, n, dx)
    want = I.curl3d(*[a.astype(np.float64) for a in q], dx)
    for g, w in zip(v, want):
        np.testing.assert_allclose(g, w, rtol=1e-5, atol=1e-5 * np.abs(w).max())
    d = I.div3d(*[a.astype(np.float64) for a in v], dx)
    assert np.abs(d).max() < 1e-3 * max(np.abs(a).max() for a in v) / min(dx)


def test_gradient_stream2d():
    nx, ny = 12, 9
    dx = (0.2, 0.1)
    psi = np.random.default_rng(4).standard_normal((nx + 1, ny + 1)).astype(np.float32)
    xv = np.empty((nx + 1, ny), np.float32)
    yv = np.empty((nx, ny + 1), np.float32)
    W.StreamInitial.GetInstance().GradientStream(xv, yv, psi, (nx, ny), dx)
    p = psi.astype(np.float64)
    np.testing.assert_allclose(xv, (p[:, 1:] - p[:, :-1]) / dx[1], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(yv, -(p[1:, :] - p[:-1, :]) / dx[0], rtol=1e-5, atol=1e-5)


def test_set_slice_zero():
    nx, ny, nz = 5, 6, 4
    for axis, sl in (("x", lambda a: (a[:, 0, :], a[:, -1, :])),
                     ("y", lambda a: (a[:, :, 0], a[:, :, -1])),
                     ("z", lambda a: (a[0], a[-1]))):
        a = np.ones((nz, nx, ny), np.float32)
        W.SetSliceZero(a, (nx, ny, nz), axis)
        lo, hi = sl(a)
        assert not lo.any() and not hi.any()
        assert a.sum() == a.size - 2 * lo.size


def test_calculate_error_l2():
    rng = np.random.default_rng(5)
    a = [rng.standard_normal(s).astype(np.float32) for s in ((3, 4, 5), (2, 6, 3), (4, 4, 4))]
    b = [x + 1e-3 * rng.standard_normal(x.shape).astype(np.float32) for x in a]
    dims = [(s[1], s[2], s[0]) for s in (x.shape for x in a)]
    e1 = W.CalculateErrorL2(a[0], b[0], dims[0])
    A, B = a[0].astype(np.float64), b[0].astype(np.float64)
    assert abs(e1 - np.sqrt(((A - B) ** 2).sum() / (A ** 2).sum())) < 1e-12
    e3 = W.CalculateErrorL2(a[0], b[0], dims[0], a[1], b[1], dims[1], a[2], b[2], dims[2])
    num = sum(((x.astype(np.float64) - y) ** 2).sum() for x, y in zip(a, b))
    den = sum((x.astype(np.float64) ** 2).sum() for x in a)
    assert abs(e3 - np.sqrt(num / (den + 1e-12))) < 1e-12
