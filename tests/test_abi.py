"""The C-ABI library loads without a GPU and exports every symbol include/whhd.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "whhd.h")).read()
    return sorted(set(re.findall(r"WHHD_API\s+[\w\s\*]+?\b(whhd_\w+)\s*\(", txt)))


@pytest.fixture(scope="module")
def L():
    from waveletpotentials_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from waveletpotentials_amd import build
        build.build()
    return _lib.lib()


def test_header_parses():
    syms = declared_symbols()
    assert "whhd_recover3d_solve" in syms and "whhd_transform1d" in syms
    assert len(syms) >= 20


def test_every_declared_symbol_is_exported(L):
    from waveletpotentials_amd import _lib
    for s in declared_symbols():
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"


def test_status_strings(L):
    assert L.whhd_status_string(0) == b"ok"
    assert b"power of two" in L.whhd_status_string(2)


def test_wavelet_registry_matches_oracle(L, orc):
    from _cases import REGISTERED
    reg = set(REGISTERED)
    for i, name in enumerate(orc.WAVELETS):
        for bc in "ndz":
            assert bool(L.whhd_wavelet_supported(i, bc.encode())) == ((name, bc) in reg), (name, bc)
            assert bool(L.whhd_wavelet_supported(i, bc.encode())) == orc.wavelet_supported(name, bc)


def test_proj_helpers(L):
    from waveletpotentials_amd import api
    assert L.whhd_proj_bc(b"n") == b"d" and L.whhd_proj_bc(b"d") == b"n"
    assert L.whhd_proj_bc(b"z") == b"z"
    assert L.whhd_proj_wavelet(api.WaveletType.WT_CDF4_6) == api.WaveletType.WT_CDF3_7
    assert L.whhd_proj_wavelet(api.WaveletType.WT_CDF3_7) == api.WaveletType.WT_CDF2_8
    assert L.whhd_proj_wavelet(api.WaveletType.WT_CDF1_7) == -1
    assert api.type_from_string("d_cdf4.6") == api.WaveletType.WT_CDF4_6
    assert api.string_from_type(api.WaveletType.WT_CDF3_7, "n") == "n_cdf3.7"


def test_argument_errors_without_gpu(L):
    """Argument checks run before any device call."""
    h = ctypes.c_void_p()
    st = L.whhd_recover3d_create(12, 16, 16, 0.1, 0.1, 0.1, b"n", 11, 7, ctypes.byref(h))
    assert st == 2  # NOT_POW2
    assert b"power" in L.whhd_last_error()
    st = L.whhd_recover3d_create(16, 16, 16, 0.1, 0.1, 0.1, b"n", 0, 7, ctypes.byref(h))
    assert st == 3  # CDF3.9 has no kernel
    st = L.whhd_transform1d(None, None, 4, 4, 4, 2, b"y", b"n", 11, 1, None)
    assert st == 1


def test_trace_calls_reject_null_handle(L):
    assert L.whhd_recover3d_trace_begin(None, 4) == 1
    assert L.whhd_recover3d_trace_end(None, None, 0, None, None) == 1
    assert L.whhd_recover3d_trace_label(None, 0) is None


@pytest.mark.parametrize("n", [(16, 8, 32), (128, 32, 64), (16, 32, 8), (64, 256, 256), (256, 256, 256),
                               (4, 64, 4)])
def test_recover3d_scratch_holds_every_pass(L, n):
    """The handle's ping-pong scratch of component c receives the middle pass of
    velocity c's chain (MAC extents with +1 on axis c) and of potential c's
    chain (+1 on the other two axes), so on non-cubic grids it must hold the
    larger of the two; the transformed velocity needs its own array."""
    nx, ny, nz = n
    vel = [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)]
    pot = [nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz]
    want = 4 * sum(v + max(v, p) for v, p in zip(vel, pot))
    got = ctypes.c_ulonglong()
    assert L.whhd_recover3d_scratch_bytes(nx, ny, nz, ctypes.byref(got)) == 0
    assert got.value == want
    # the rule sized by the potential alone (the old sizing) is short wherever a
    # velocity array is the larger one
    short = 4 * sum(v + p for v, p in zip(vel, pot))
    assert (got.value > short) == any(v > p for v, p in zip(vel, pot))
    assert L.whhd_recover3d_scratch_bytes(0, 4, 4, ctypes.byref(got)) == 1
