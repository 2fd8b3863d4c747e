"""The reference-style C++ classes of include/whhd_compat.hpp, driven from a
compiled C++ program (tests/cpp/compat_recover3d.cpp, built by
__graft_entry__.build()), against the FP64 oracle: a translation unit written
for the reference's headers gets the same results after switching the include.

Tolerance: FP32 kernels vs the FP64 oracle, relative L2 <= 1e-5.
"""
import os
import subprocess

import numpy as np
import pytest

from waveletpotentials_amd import inputs as I

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "compat_recover3d")


def test_compat_classes_match_oracle(orc, tmp_path):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run __graft_entry__.build() first")
    N = 16
    n = (N, N, N)
    dx = (1.0 / N,) * 3
    (xv, yv, zv), _, _ = I.velocity3d(n, 4, I.seed_for(1, 99))
    for name, a in (("xv", xv), ("yv", yv), ("zv", zv)):
        np.ascontiguousarray(a, dtype=np.float32).tofile(tmp_path / f"{name}.f32")
    r = subprocess.run([BIN, str(N), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = orc.recover3d(xv, yv, zv, n, dx)
    got = [np.fromfile(tmp_path / f"{c}.f32", dtype=np.float32).astype(np.float64).reshape(q.shape)
           for c, q in zip(("qx", "qy", "qz"), ref)]
    assert orc.rel_l2(ref, got) < 1e-5
    t_ref, _ = orc.transform(xv, 4, "z", "n", "cdf3.7", True)
    t_got = np.fromfile(tmp_path / "solve1d.f32", dtype=np.float32).astype(np.float64)
    assert orc.rel_l2(t_ref, t_got.reshape(t_ref.shape)) < 1e-5
