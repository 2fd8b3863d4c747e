"""GPU parity of the line transforms (libwhhd.so) against the FP64 oracle.

FP32 kernels vs the FP64 oracle: relative L2 <= 2e-6 per array (FP32 roundoff
over ~10 lifting levels).  FP64 kernels vs the oracle: <= 1e-13.
"""
import numpy as np
import pytest
import torch

from _cases import ODD_FAMILIES, REGISTERED, bc_consistent, line_len

pytestmark = pytest.mark.gpu

TOL32 = 2e-6
TOL64 = 1e-13


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def shape_for(axis, n, m1, m2):
    # (nz, nx, ny) with the transformed axis of length n
    return {"y": (m2, m1, n), "x": (m2, n, m1), "z": (n, m1, m2)}[axis]


def dims_of(shape):
    nz, nx, ny = shape
    return (nx, ny, nz)


def make(rng, shape, axis, wt, bc):
    x = rng.standard_normal(shape)
    if wt in ODD_FAMILIES and bc == "d":
        idx = {"y": (slice(None), slice(None), [0, -1]), "x": (slice(None), [0, -1], slice(None)),
               "z": ([0, -1], slice(None), slice(None))}[axis]
        x[idx] = 0.0
    return x


@pytest.mark.parametrize("wt,bc", REGISTERED)
@pytest.mark.parametrize("axis", ["y", "x", "z"])
def test_transform_matches_oracle(whhd, orc, wt, bc, axis):
    rng = np.random.default_rng(hash((wt, bc, axis)) % 2**32)
    T = whhd.WaveletType
    wtype = T[f"WT_CDF{wt[3:].replace('.', '_')}"]
    fw = whhd.CuWHHDForward.GetInstance()
    for levels in (2, 3, 5, 6, 8):
        n = line_len(wt, levels)
        shape = shape_for(axis, n, 5, 3)
        x = make(rng, shape, axis, wt, bc)
        ref_f, _ = orc.transform(x, levels, axis, bc, wt, True)
        ref_b, _ = orc.transform(ref_f, levels, axis, bc, wt, False)
        for dt, tol in ((torch.float32, TOL32), (torch.float64, TOL64)):
            src = torch.tensor(x, dtype=dt, device="cuda")
            dst = torch.empty_like(src)
            fw.Solve1D(dst, src, dims_of(shape), levels, axis, bc, wtype, True)
            torch.cuda.synchronize()
            assert rel(dst.cpu().numpy(), ref_f) < tol, (n, dt, "forward")
            back = torch.tensor(ref_f, dtype=dt, device="cuda")
            fw.Solve1D(back, None, dims_of(shape), levels, axis, bc, wtype, False)  # in place
            torch.cuda.synchronize()
            assert rel(back.cpu().numpy(), ref_b) < tol, (n, dt, "inverse")


@pytest.mark.parametrize("wt,bc", [("cdf3.7", "n"), ("cdf4.6", "d")])
@pytest.mark.parametrize("axis", ["y", "x", "z"])
def test_transform_config_sizes(whhd, orc, kernel_mode, wt, bc, axis):
    """The two (type, bc) pairs of the recovery path at lengths 2^2..2^10 (+1),
    out of place (line-parallel kernels for 2^5..2^8) with ragged line counts
    so partial tiles and partial wavefronts are exercised."""
    rng = np.random.default_rng(5)
    T = whhd.WaveletType
    wtype = T.WT_CDF3_7 if wt == "cdf3.7" else T.WT_CDF4_6
    fw = whhd.CuWHHDForward.GetInstance()
    for levels in (2, 3, 4, 5, 6, 7, 8, 10):
        n = line_len(wt, levels)
        shape = shape_for(axis, n, 37, 3)
        x = make(rng, shape, axis, wt, bc)
        ref_f, _ = orc.transform(x, levels, axis, bc, wt, True)
        src = torch.tensor(x, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        fw.Solve1D(dst, src, dims_of(shape), levels, axis, bc, wtype, True)
        back = torch.empty_like(src)
        fw.Solve1D(back, dst, dims_of(shape), levels, axis, bc, wtype, False)
        torch.cuda.synchronize()
        assert rel(dst.cpu().numpy(), ref_f) < TOL32
        assert rel(back.cpu().numpy(), x) < TOL32


def test_gram_tables_match_oracle(whhd, orc):
    ip = whhd.CuGetInnerProduct.GetInstance()
    T = whhd.WaveletType
    for wt, bc, n, t in (("cdf4.6", "d", 33, T.WT_CDF4_6), ("cdf3.7", "n", 32, T.WT_CDF3_7),
                         ("cdf4.6", "d", 257, T.WT_CDF4_6), ("cdf3.7", "n", 256, T.WT_CDF3_7),
                         ("cdf2.8", "n", 65, T.WT_CDF2_8), ("cdf1.1", "n", 16, T.WT_CDF1_1)):
        g = ip.GenerateIp(t, bc, n).cpu().numpy()
        ref = orc.gram_table(wt, bc, n)
        np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-13)
    g2 = ip.GenerateIp("d_cdf4.6", 33)
    assert g2 is ip.GenerateIp(T.WT_CDF4_6, "d", 33)  # cached like ips_map_


def test_invalid_arguments_raise(whhd):
    fw = whhd.CuWHHDForward.GetInstance()
    T = whhd.WaveletType
    x = torch.zeros((1, 1, 12), device="cuda")
    with pytest.raises(whhd.WhhdError, match="NOT_POW2"):
        fw.Solve1D(x, x, (1, 12, 1), 3, "y", "n", T.WT_CDF3_7, True)
    x = torch.zeros((1, 1, 33), device="cuda")
    with pytest.raises(whhd.WhhdError, match="UNSUPPORTED"):
        fw.Solve1D(x, x, (1, 33, 1), 5, "y", "n", T.WT_CDF3_7, True)  # even family, odd line
    with pytest.raises(whhd.WhhdError, match="UNSUPPORTED"):
        fw.Solve1D(x, x, (1, 33, 1), 5, "y", "n", T.WT_CDF3_9, True)  # no kernel in the reference
    with pytest.raises(whhd.WhhdError, match="INVALID"):
        fw.Solve1D(x, x, (1, 33, 1), 4, "y", "d", T.WT_CDF4_6, True)  # wrong level
