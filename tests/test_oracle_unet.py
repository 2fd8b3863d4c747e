"""Pins of the U-Net oracle (FP64 convolutions and their composition).

* the 3x3x3 convolution equals scipy.ndimage.correlate summed over input
  channels (zero padding), and the fused upsampled input equals an explicitly
  upsampled one;
* the input gradient is the adjoint of the convolution (dot-product test) and
  the weight gradient is the derivative of the (linear in w) convolution;
* the composed network's gradient matches central finite differences of its
  loss, layer by layer, biases included;
* BF16 rounding is round-to-nearest-even on the FP32 bit pattern;
* the resampling around a coarser network keeps constants, and the trilinear
  interpolation reproduces a linear field away from the walls.
"""
import numpy as np
import pytest
from scipy import ndimage


def test_conv3_equals_correlate(orc):
    rng = np.random.default_rng(1)
    dims = (4, 6, 5)
    ca, cb, co = 3, 2, 4
    a = rng.standard_normal((np.prod(dims), ca))
    b = rng.standard_normal((np.prod(dims), cb))
    w = rng.standard_normal(co * 27 * (ca + cb))
    bias = rng.standard_normal(co)
    y = orc.conv3(a, ca, b, cb, False, dims, w, bias, co)
    x = np.concatenate([a, b], axis=1).reshape(*dims, ca + cb)
    W = w.reshape(co, 3, 3, 3, ca + cb)     # taps (dz, dx, dy) in 0..2
    for o in range(co):
        ref = bias[o] + sum(ndimage.correlate(x[..., c], W[o, :, :, :, c], mode="constant")
                            for c in range(ca + cb))
        assert np.allclose(y[:, o].reshape(dims), ref, atol=1e-12)


def test_conv3_upsampled_input(orc):
    rng = np.random.default_rng(2)
    dims = (4, 4, 6)
    coarse = rng.standard_normal((2 * 2 * 3, 3))
    skip = rng.standard_normal((np.prod(dims), 2))
    w = rng.standard_normal(2 * 27 * 5)
    y = orc.conv3(coarse, 3, skip, 2, True, dims, w, None, 2)
    c = coarse.reshape(2, 2, 3, 3)
    up = np.repeat(np.repeat(np.repeat(c, 2, 0), 2, 1), 2, 2).reshape(-1, 3)
    ref = orc.conv3(up, 3, skip, 2, False, dims, w, None, 2)
    assert np.allclose(y, ref, atol=1e-12)


def test_conv3_gradients_are_adjoints(orc):
    rng = np.random.default_rng(3)
    dims = (3, 5, 4)
    V, cin, co = np.prod(dims), 6, 5
    x = rng.standard_normal((V, cin))
    g = rng.standard_normal((V, co))
    w1, w2 = rng.standard_normal(co * 27 * cin), rng.standard_normal(co * 27 * cin)
    y1 = orc.conv3(x, cin, None, 0, False, dims, w1, None, co)
    dx = orc.conv3_bwd_data(g, co, dims, w1, cin)
    assert abs(np.sum(y1 * g) - np.sum(x * dx)) < 1e-10 * np.sum(np.abs(y1 * g))
    y2 = orc.conv3(x, cin, None, 0, False, dims, w2, None, co)
    dw, db = orc.conv3_bwd_weight(g, co, x, cin, None, 0, False, dims)
    assert abs(np.sum((y2 - y1) * g) - np.sum((w2 - w1) * dw)) < 1e-10 * np.sum(np.abs(y1 * g))
    assert np.allclose(db, g.sum(axis=0))


def test_unet_gradient_matches_finite_differences(orc):
    rng = np.random.default_rng(4)
    dims = (8, 8, 8)
    lay, n = orc.unet_layout()
    params = np.zeros(n)
    for (ca, cb, _, co, _), (wo, bo) in zip(orc.UNET_LAYERS, lay):
        rows = 1 if wo == lay[-1][0] else co
        w = params[wo:bo].reshape(co, 27, ca + cb)
        w[:rows] = rng.standard_normal((rows, 27, ca + cb)) * np.sqrt(2.0 / (27 * (ca + cb)))
        params[bo:bo + rows] = 0.1 * rng.standard_normal(rows)
    f = rng.standard_normal(dims)
    ls = rng.uniform(-0.2, 0.5, dims)
    p = rng.standard_normal(dims) * 1e-3
    length = 1.0
    loss, grads, _ = orc.unet_loss_grad(params, f, ls, p, length)
    assert loss > 0
    checked = 0
    for l, (wo, bo) in enumerate(lay):
        co = orc.UNET_LAYERS[l][3]
        rows_end = bo if l < len(lay) - 1 else wo + (bo - wo) // co   # live rows only
        for idx in list(rng.choice(np.arange(wo, rows_end), 3, replace=False)) + [bo]:
            h = 1e-5 * max(1.0, abs(params[idx]))
            pp, pm = params.copy(), params.copy()
            pp[idx] += h
            pm[idx] -= h
            lp = orc.unet_loss_grad(pp, f, ls, p, length)[0]
            lm = orc.unet_loss_grad(pm, f, ls, p, length)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[idx]) <= 1e-5 * max(abs(fd), 1e-3 * loss), (l, idx, fd, grads[idx])
            checked += 1
    assert checked == 4 * len(lay)


def test_bf16_round(orc):
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1e-30, 3.140625])
    r = orc.bf16_round(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                    # tie: to even
    assert r[2] == 1.0 + 4 * 2.0 ** -8    # tie: to even (upward)
    assert r[3] == -2.5
    assert abs(r[4] - 1e-30) <= 1e-30 * 2.0 ** -8
    assert r[5] == 3.140625


def test_resampling(orc):
    c = np.full((3, 4, 5), 2.5)
    assert np.allclose(orc.unet_up(c, 4), 2.5)
    assert np.allclose(orc.unet_down(np.full((8, 12, 4), -1.0), 4), -1.0)
    # linear field sampled at coarse cell centres; interior fine cells exact
    F, nc = 2, (4, 5, 6)
    zc, xc, yc = np.meshgrid(*[np.arange(n) + 0.5 for n in nc], indexing="ij")
    coarse = 0.3 * zc - 1.2 * xc + 0.7 * yc
    fine = orc.unet_up(coarse, F)
    zf, xf, yf = np.meshgrid(*[(np.arange(n * F) + 0.5) / F for n in nc], indexing="ij")
    ref = 0.3 * zf - 1.2 * xf + 0.7 * yf
    inner = (slice(1, -1),) * 3
    assert np.allclose(fine[inner], ref[inner], atol=1e-12)
    # block means of the linear fine field are the coarse values
    assert np.allclose(orc.unet_down(ref, F), coarse, atol=1e-12)
