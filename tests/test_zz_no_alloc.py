"""whhd_reduce_noalloc (the reduction whose partials live in a static device
array) against whhd_reduce with a caller-owned workspace: the same kernels, so
every result must be identical, on two calls in a row with the second smaller."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = {0: "sum", 1: "sum of squares", 2: "sum of |x|", 3: "max |x|"}


def _reduce(L, x, n, op, torch):
    scratch = torch.empty(L.whhd_reduce_scratch_doubles(), dtype=torch.float64, device="cuda")
    out = ctypes.c_double()
    assert L.whhd_reduce(x.data_ptr(), n, op, scratch.data_ptr(), ctypes.byref(out), None) == 0
    return out.value


def _noalloc(L, x, n, op):
    out = ctypes.c_double()
    assert L.whhd_reduce_noalloc(x.data_ptr(), n, op, ctypes.byref(out), None) == 0
    return out.value


@pytest.mark.parametrize("op", sorted(OPS))
def test_noalloc_reduce_equals_workspace_reduce(whhd, op):
    import torch
    L = whhd.lib()
    g = torch.Generator(device="cpu").manual_seed(17 + op)
    big = torch.randn(3_000_017, generator=g).to("cuda")
    small = torch.randn(1_003, generator=g).to("cuda")
    for x in (big, small):      # second call smaller: nothing of the first may leak in
        n = x.numel()
        print(f"reduce_noalloc op={OPS[op]} n={n}", flush=True)
        got = _noalloc(L, x, n, op)
        want = _reduce(L, x, n, op, torch)
        assert got == want, (op, n, got, want)
        ref = x.double().cpu().numpy()
        exact = {0: ref.sum(), 1: (ref * ref).sum(), 2: np.abs(ref).sum(), 3: np.abs(ref).max()}[op]
        assert abs(got - exact) <= 1e-9 * max(1.0, abs(exact))


def test_noalloc_reduce_of_empty_and_bad_arguments(whhd):
    import torch
    L = whhd.lib()
    x = torch.ones(8, device="cuda")
    out = ctypes.c_double(-1.0)
    assert L.whhd_reduce_noalloc(x.data_ptr(), 0, 0, ctypes.byref(out), None) == 0
    assert out.value == 0.0
    assert L.whhd_reduce_noalloc(x.data_ptr(), 8, 4, ctypes.byref(out), None) == 1
    assert L.whhd_reduce_noalloc(None, 8, 0, ctypes.byref(out), None) == 1
    assert L.whhd_reduce_noalloc(x.data_ptr(), -1, 0, ctypes.byref(out), None) == 1
