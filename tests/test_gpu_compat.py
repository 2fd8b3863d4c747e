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
    # CuGetInnerProduct: the program checked the cache (nullptr before
    # generation, one pointer per key); both tables stay readable afterwards
    # and equal the oracle's Gram tables rounded to FP32
    for fname, wt, bc, m in (("gram_d", "cdf4.6", "d", N + 1), ("gram_n", "cdf3.7", "n", N)):
        g_ref = np.asarray(orc.gram_table(wt, bc, m)).reshape(-1)
        g_got = np.fromfile(tmp_path / f"{fname}.f32", dtype=np.float32).astype(np.float64)
        assert g_got.shape == g_ref.shape
        assert np.abs(g_got - g_ref).max() <= 1e-6 * np.abs(g_ref).max()


def test_compat_solid_classes_match_oracle(orc, tmp_path):
    """CuSolidLevelSet3D, CuConvergence, CuMultigridFrac3D, CuGradient, the Cuda*
    helpers and CuParallelSweep3D through the C++ classes, against the oracle:
    level set and f to FP32 rounding, 10 V-cycles to 1e-4 (the multigrid
    tolerance), the residual it reports to the oracle's residual of the same
    iterate, the sweep + DST potential to 1e-5, and the gradient correction
    reducing max |div| by the multigrid's residual."""
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run __graft_entry__.build() first")
    N = 16
    n = (N, N, N)
    h = (1.0 / N,) * 3
    (xv, yv, zv), _, _ = I.velocity3d(n, 4, I.seed_for(2, 99))
    for name, a in (("xv", xv), ("yv", yv), ("zv", zv)):
        np.ascontiguousarray(a, dtype=np.float32).tofile(tmp_path / f"{name}.f32")
    c, r = (0.45, 0.55, 0.5), 0.2
    out = subprocess.run([BIN, "solids", str(N), str(tmp_path), *map(str, c), str(r)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    res, fmax0, fmax1 = map(float, out.stdout.split())
    load = lambda name, shape: np.fromfile(tmp_path / f"{name}.f32", np.float32).astype(np.float64).reshape(shape)  # noqa: E731
    ls_ref = orc.levelset_spheres(n, h, np.array([c]), np.array([r]))
    assert np.abs(load("ls", ls_ref.shape) - ls_ref).max() < 1e-6
    fr = orc.frac_from_ls(ls_ref, h)
    u32 = [a.astype(np.float32).astype(np.float64) for a in (xv, yv, zv)]
    f_ref = orc.divergence3d_frac(*u32, *fr, h)
    f_got = load("f", f_ref.shape)
    assert orc.rel_l2(f_ref, f_got) < 1e-5
    assert abs(fmax0 - np.abs(f_got).max()) <= 1e-6 * fmax0
    p_ref = orc.mg_solve(f_got, *fr, h, "n", n_iters=10, nu1=2, nu2=2, weight=0.8)
    p_got = load("p", p_ref.shape)
    assert orc.rel_l2(p_ref - p_ref.mean(), p_got - p_got.mean()) < 1e-4
    r_ref = orc.mg_rel_residual(p_got, f_got, *fr, h, "n")
    assert abs(res - r_ref) <= 1e-3 * r_ref + 1e-7
    assert fmax1 < 0.5 * fmax0
    q_ref = orc.project_zero(*orc.sweep3d(xv, yv, n, h), n, h)
    q_got = [load(k, q.shape) for k, q in zip(("sx", "sy", "sz"), q_ref)]
    assert orc.rel_l2(list(q_ref), q_got) < 1e-5
    # the static GetFrac overload runs the supersampled fractions, not the
    # two-point formula of the member overload
    d_ref = orc.frac_divide(load("ls", ls_ref.shape), h)
    for k, a in zip(("dax", "day", "daz"), d_ref):
        got = load(k, a.shape)
        assert np.abs(got - a).max() <= 4.0 / 16384, k
    assert max(np.abs(a - b).max() for a, b in zip(d_ref, fr)) > 1e-2
