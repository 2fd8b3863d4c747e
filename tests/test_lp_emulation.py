"""CPU check of the line-parallel transform algorithm (the register-resident
kernels of csrc/line_parallel.h, compiled for the host by
tests/cpp/lp_emulation.cpp) against the FP64 oracle, for both wavelets of the
recovery path, levels 5-8, both directions, every look-ahead depth of the
single-phase loads and the two-phase (fetch, then unpack one block later)
loads of the contiguous-line kernels.

The host build runs the same templates as the GPU kernels (level and block
structure, reflections, streamed loads, run stores) one line at a time; the
device-only I/O (buffer addressing, LDS transposition) is covered by the GPU
tests.  Tolerance: FP32 arithmetic vs FP64, 2e-6 relative to the largest
coefficient.
"""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emu():
    import __graft_entry__ as G
    stale = (not os.path.exists(G.LP_EMU_LIB)
             or os.path.getmtime(G.LP_EMU_LIB) < max(
                 os.path.getmtime(p) for p in (G.LP_EMU_SRC,
                                               os.path.join(ROOT, "waveletpotentials_amd", "csrc",
                                                            "line_parallel.h"),
                                               os.path.join(ROOT, "waveletpotentials_amd", "csrc",
                                                            "lifting.h"))))
    if stale:
        G.build_lp_emulation()
    L = ctypes.CDLL(G.LP_EMU_LIB)
    L.lp_emulate.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_longlong, ctypes.c_int,
                                                     ctypes.c_longlong] + [ctypes.c_int] * 4
    return L


@pytest.mark.parametrize("wt,name,bc,ext", [(0, "cdf3.7", "n", 0), (1, "cdf4.6", "d", 1)])
@pytest.mark.parametrize("lev", [5, 6, 7, 8])
@pytest.mark.parametrize("fwd", [1, 0])
def test_lp_algorithm_matches_oracle(emu, orc, wt, name, bc, ext, lev, fwd):
    n, lines = (1 << lev) + ext, 6
    rng = np.random.default_rng(1000 * lev + 10 * wt + fwd)
    src = rng.standard_normal((1, lines, n))
    ref, _ = orc.transform(src, lev, "y", bc, name, bool(fwd))
    for look in (0, 16, 32, -1):
        s32 = src.astype(np.float32)
        d32 = np.full_like(s32, np.nan)
        assert emu.lp_emulate(d32.ctypes.data, s32.ctypes.data, lines, n, n, lev, wt, fwd, look) == 0
        assert np.isfinite(d32).all(), "a coefficient was never written"
        assert np.abs(d32 - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("n", [33, 65, 129, 257])
def test_aligned_final_stores_cover_each_sample_once(emu, n):
    """The inverse's aligned final stores for lines of 2^L + 1 samples (whole
    128-byte lines, the line boundaries stored at the last step): on the
    kernel's own index functions, every sample of every existing line of a
    wavefront is stored exactly once with its value, nothing is stored past
    the last existing line, and every LDS word read lies in the three staging
    areas, for full and partial wavefronts."""
    emu.lp_tail_model.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    for lines_here in (64, 63, 62, 33, 32, 31, 30, 17, 2, 1):
        out = np.full(64 * n, np.nan, dtype=np.float32)
        assert emu.lp_tail_model(n, lines_here, out.ctypes.data) == 0, lines_here
