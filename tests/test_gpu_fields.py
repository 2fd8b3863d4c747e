"""GPU parity of the field operators and reductions against the FP64 oracle.

Tolerances (FP32 kernels vs FP64): gradients and divergences within 1e-5
relative to the largest magnitude; FP64-accumulated reductions of FP32 data
within 1e-12 relative of the numpy FP64 sums of the same FP32 values.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def cuda(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.float64)


def fracs(n, rng):
    nx, ny, nz = n
    fr = [rng.uniform(0.0, 1.0, s) for s in ((nz, nx + 1, ny), (nz, nx, ny + 1), (nz + 1, nx, ny))]
    for a in fr:
        a[a < 0.2] = 0.0   # some closed faces
    return [a.astype(np.float32).astype(np.float64) for a in fr]


@pytest.mark.parametrize("bc", ["n", "d", "z"])
@pytest.mark.parametrize("n", [(16, 12, 10), (64, 64, 64)])
def test_gradient_and_divergence(whhd, orc, bc, n):
    rng = np.random.default_rng(sum(n) + ord(bc))
    h = tuple(1.0 / v for v in n)
    fr = fracs(n, rng)
    nx, ny, nz = n
    p = rng.standard_normal((nz, nx, ny)).astype(np.float32).astype(np.float64)
    dfr = [cuda(a) for a in fr]
    g = [torch.full(a.shape, np.nan, device="cuda") for a in fr]
    whhd.CuGradient.GetInstance().Gradient3D_Frac_P(*g, cuda(p), *dfr, n, h, bc)
    ref = orc.gradient3d_frac(p, *fr, h, bc)
    for a, b in zip(g, ref):
        got = host(a)
        assert np.isfinite(got).all()
        assert np.abs(got - b).max() <= 1e-5 * np.abs(b).max()
    u = [rng.standard_normal(a.shape).astype(np.float32).astype(np.float64) for a in fr]
    b = torch.empty((nz, nx, ny), device="cuda")
    whhd.CuConvergence.GetInstance().GetB_3d_frac_vel(b, *(cuda(a) for a in u), *dfr, n, h, bc)
    bref = orc.divergence3d_frac(*u, *fr, h)
    assert np.abs(host(b) - bref).max() <= 1e-5 * np.abs(bref).max()
    # fused u <- a (u - grad p) equals the three separate steps
    du = [cuda(a) for a in u]
    whhd.CuGradient.GetInstance().SubtractGradient3D_Frac_P(*du, cuda(p), *dfr, n, h, bc)
    for got, uu, gg, a in zip(du, u, ref, fr):
        want = a * (uu - gg)
        assert np.abs(host(got) - want).max() <= 1e-5 * max(np.abs(want).max(), 1.0)
    # two-float potential: u <- a (u - grad p_hi - grad p_lo) against the FP64
    # oracle gradient of p_hi + p_lo (the low part is not rounded away)
    p_lo = (rng.standard_normal((nz, nx, ny)) * 1e-4).astype(np.float32).astype(np.float64)
    ref_sum = orc.gradient3d_frac(p + p_lo, *fr, h, bc)
    du = [cuda(a) for a in u]
    whhd.CuGradient.GetInstance().SubtractGradient3D_Frac_P2(*du, cuda(p), cuda(p_lo), *dfr, n, h, bc)
    for got, uu, gg, a in zip(du, u, ref_sum, fr):
        want = a * (uu - gg)
        assert np.abs(host(got) - want).max() <= 1e-5 * max(np.abs(want).max(), 1.0)


def test_elementwise_and_reductions(whhd):
    rng = np.random.default_rng(9)
    x = rng.standard_normal(3_000_001).astype(np.float32)
    y = rng.standard_normal(3_000_001).astype(np.float32)
    dx, dy = cuda(x), cuda(y)
    whhd.CudaMatAdd(dx, dy, x.size, -0.5)
    want = x + np.float32(-0.5) * y
    assert np.allclose(host(dx), want, rtol=1e-6, atol=1e-6)
    z = torch.empty_like(dx)
    whhd.CudaMatMul(z, dx, dy, x.size)
    assert np.allclose(host(z), want * y, rtol=1e-6, atol=1e-6)
    v = want.astype(np.float64)
    assert whhd.CudaArraySum(dx, x.size) == pytest.approx(v.sum(), rel=1e-9, abs=1e-6)
    assert whhd.CudaArrayNormSum(dx, x.size) == pytest.approx((v * v).sum(), rel=1e-12)
    assert whhd.CudaArrayAbsSum(dx, x.size) == pytest.approx(np.abs(v).sum(), rel=1e-12)
    assert whhd.CudaFindMaxValue(dx, x.size) == np.abs(v).max()
    assert whhd.CudaArraySum(dx, 0) == 0.0
