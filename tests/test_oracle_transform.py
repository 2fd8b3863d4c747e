"""CPU checks of the oracle's line transforms (WaveletPotentials/cuWHHDForward.cu,
cuWaveletCompute.cuh) against what fixes them independently:

* a lane-level emulation of the reference's 32-lane kernels (window, halo, shuffle
  and shared-memory ping-pong semantics) must give identical numbers;
* inverse(forward(x)) == x for boundary-consistent data;
* the CDF(p, q) structure: details of degree < p polynomials vanish and the
  synthesis wavelet has q vanishing moments (pins every predict/update tap);
* the Haar member (CDF1.1) gives an orthonormal basis: its Gram table is I.
"""
import numpy as np
import pytest

from _cases import CDF_PQ, ODD_FAMILIES, REGISTERED, bc_consistent, line_len


@pytest.mark.parametrize("wt,bc", REGISTERED)
def test_clean_equals_lane_emulation(orc, wt, bc):
    rng = np.random.default_rng(7)
    for levels in range(2, 11):
        n = line_len(wt, levels)
        if n < (5 if wt in ODD_FAMILIES else 4):
            continue
        x = rng.standard_normal(n)
        f, rc = orc.transform(x.reshape(1, 1, n), levels, "y", bc, wt, True)
        fe, rce = orc.emulate_line(x, levels, bc, wt, True)
        assert rc == 0 and rce == 0
        np.testing.assert_array_equal(f.ravel(), fe)
        b, rc = orc.transform(f, levels, "y", bc, wt, False)
        be, rce = orc.emulate_line(f.ravel(), levels, bc, wt, False)
        assert rc == 0 and rce == 0
        np.testing.assert_array_equal(b.ravel(), be)


@pytest.mark.parametrize("wt,bc", REGISTERED)
def test_perfect_reconstruction(orc, wt, bc):
    rng = np.random.default_rng(11)
    for levels in range(2, 12):
        n = line_len(wt, levels)
        x = bc_consistent(rng.standard_normal((3, 2, n)), wt, bc)
        f, _ = orc.transform(x, levels, "y", bc, wt, True)
        b, _ = orc.transform(f, levels, "y", bc, wt, False)
        assert np.abs(b - x).max() < 1e-12 * max(1.0, np.abs(x).max()), (n, np.abs(b - x).max())


@pytest.mark.parametrize("wt,bc", REGISTERED)
def test_wrong_parity_is_undefined(orc, wt, bc):
    n = (1 << 6) + (0 if wt in ODD_FAMILIES else 1)
    x = np.random.default_rng(0).standard_normal((1, 1, n))
    _, rc = orc.transform(x, 6, "y", bc, wt, True, allow_undefined=True)
    assert rc == orc.UNDEFINED


@pytest.mark.parametrize("wt", sorted(CDF_PQ))
def test_cdf_vanishing_moments(orc, wt):
    p, q = CDF_PQ[wt]
    levels = 8
    n = line_len(wt, levels)
    W = n // 2 + (n & 1)
    t = np.arange(n) / n - 0.3
    for k in range(p + 1):
        f, _ = orc.transform((t ** k).reshape(1, 1, n), levels, "y", "n", wt, True)
        det = np.abs(f.ravel()[W + 8:W + n // 2 - 8]).max()
        if k < p:
            assert det < 1e-13, (k, det)
        else:
            assert det > 1e-10, (k, det)  # the next degree is not annihilated
    e = np.zeros(n)
    e[W + n // 4] = 1.0
    b, _ = orc.transform(e.reshape(1, 1, n), levels, "y", "n", wt, False)
    x = np.arange(n) - n / 2.0
    for k in range(q + 1):
        m = abs(np.sum(b.ravel() * x ** k))
        if k < q:
            assert m < 1e-8 * (n / 2.0) ** k, (k, m)
        else:
            assert m > 1e-3, (k, m)


def test_haar_gram_is_identity(orc):
    for n in (16, 64):
        g = orc.gram_table("cdf1.1", "n", n)
        np.testing.assert_allclose(g[0], 1.0, atol=1e-13)
        np.testing.assert_allclose(g[1:5], 0.0, atol=1e-13)


@pytest.mark.parametrize("wt,bc,n", [("cdf4.6", "d", 33), ("cdf3.7", "n", 32), ("cdf2.8", "n", 33)])
def test_gram_symmetric_positive(orc, wt, bc, n):
    g = orc.gram_table(wt, bc, n)
    assert (g[0] > 0).all()
    np.testing.assert_allclose(g[3][:-1], g[2][1:], atol=1e-14)  # G[i,i+1] == G[i+1,i]
    np.testing.assert_allclose(g[4][:-2], g[1][2:], atol=1e-14)


def test_axes_agree(orc):
    """The same lines transformed along x, y and z give the same numbers."""
    rng = np.random.default_rng(3)
    n, m = 33, 8
    lines = rng.standard_normal((m, n))
    lines[:, 0] = lines[:, -1] = 0.0
    ref, _ = orc.transform(lines.reshape(1, m, n), 5, "y", "d", "cdf4.6", True)
    ax, _ = orc.transform(np.ascontiguousarray(lines.T.reshape(1, n, m)), 5, "x", "d", "cdf4.6", True)
    az, _ = orc.transform(np.ascontiguousarray(lines.T.reshape(n, 1, m)), 5, "z", "d", "cdf4.6", True)
    np.testing.assert_array_equal(ref.reshape(m, n), ax.reshape(n, m).T)
    np.testing.assert_array_equal(ref.reshape(m, n), az.reshape(n, m).T)
