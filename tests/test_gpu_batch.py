"""Batched recovery of many grids of uneven sizes (one launch per stage and size
class, classes as parallel branches, one HIP graph for the batch).

Batching changes which workgroup handles a line, never the arithmetic, so every
problem's potential must equal the single-grid solve bit for bit; the single-grid
solve is pinned against the FP64 oracle in test_gpu_recover.py, and one
anisotropic problem here is also checked against the oracle at the 1e-5 bar.  A
captured graph must reproduce the eager result on every replay, with eager work
and new inputs between replays; autotuned launch shapes must leave the results
unchanged.
"""
import numpy as np
import pytest
import torch

from waveletpotentials_amd import inputs as I

pytestmark = pytest.mark.gpu

# sizes below 32 cells run the windowed kernels one problem at a time
DIMS = [(32, 32, 32), (64, 32, 32), (32, 32, 32), (16, 16, 16), (32, 64, 32), (64, 32, 32),
        (32, 32, 32), (8, 8, 8), (32, 32, 64)]


def curl_t(qx, qy, qz, dx):
    hx, hy, hz = dx
    xv = (qz[:, :, 1:] - qz[:, :, :-1]) / hy - (qy[1:] - qy[:-1]) / hz
    yv = (qx[1:] - qx[:-1]) / hz - (qz[:, 1:, :] - qz[:, :-1, :]) / hx
    zv = (qy[:, 1:, :] - qy[:, :-1, :]) / hx - (qx[:, :, 1:] - qx[:, :, :-1]) / hy
    return xv, yv, zv


def fill(batch, seed):
    for i, d in enumerate(batch.dims):
        dx = tuple(1.0 / v for v in d)
        q = I.spectral_potential3d(d, max(2, min(d) // 8), I.seed_for(4, seed + i))
        a = batch.arrays(i)
        for key, v in zip(("xv", "yv", "zv"), I.curl3d(*q, dx)):
            a[key].copy_(torch.from_numpy(v).to(device="cuda", dtype=torch.float32))


def make(whhd, dims, bc="n"):
    T = whhd.WaveletType
    return whhd.CuBatchRecover3D(dims, [tuple(1.0 / v for v in d) for d in dims], bc,
                                 T.WT_CDF3_7, T.WT_CDF4_6)


def snapshot(batch):
    return [[batch.arrays(i)[k].clone() for k in ("qx", "qy", "qz")] for i in range(len(batch.dims))]


def single(whhd, d, a, bc="n"):
    T = whhd.WaveletType
    dx = tuple(1.0 / v for v in d)
    rec = whhd.CuWaveletPotentialRecover3D(d, dx, bc, T.WT_CDF3_7, T.WT_CDF4_6)
    q = [torch.full_like(a[k], float("nan")) for k in ("qx", "qy", "qz")]
    rec.Solve(*q, a["xv"], a["yv"], a["zv"])
    return q


def test_batch_equals_single_solves(whhd, kernel_mode):
    b = make(whhd, DIMS)
    cls = b.classes()
    assert sorted(c["problems"] for c in cls) == sorted([3, 2, 1, 1, 1, 1])
    assert sum(c["problems"] for c in cls) == len(DIMS)
    fill(b, 0)
    for k in ("qx", "qy", "qz"):
        for i in range(len(DIMS)):
            b.arrays(i)[k].fill_(float("nan"))
    b.Solve()
    got = snapshot(b)
    for i, d in enumerate(DIMS):
        ref = single(whhd, d, b.arrays(i))
        for g, r in zip(got[i], ref):
            assert torch.isfinite(g).all(), f"problem {i} {d}: coefficient never written"
            assert torch.equal(g, r), f"problem {i} {d} differs from the single-grid solve"


def test_batch_velocity_bc_d_equals_single_solves(whhd):
    """Velocity bc 'd' runs the windowed transforms and the divergence bc 'n'
    projection per size class; still bit for bit the single-grid solves."""
    dims = [(32, 32, 32), (16, 16, 16), (32, 32, 32), (16, 32, 8)]
    b = make(whhd, dims, "d")
    fill(b, 3)
    b.Solve()
    got = snapshot(b)
    for i, d in enumerate(dims):
        ref = single(whhd, d, b.arrays(i), "d")
        for g, r in zip(got[i], ref):
            assert torch.equal(g, r), f"problem {i} {d} differs from the single-grid solve"


def test_batch_problem_matches_oracle(whhd, orc):
    b = make(whhd, DIMS)
    fill(b, 0)
    b.Solve()
    i = 4
    d = DIMS[i]
    dx = tuple(1.0 / v for v in d)
    a = b.arrays(i)
    v = [a[k].double().cpu().numpy() for k in ("xv", "yv", "zv")]
    ref = orc.recover3d(*v, d, dx)
    torch.cuda.synchronize()
    got = [a[k].double().cpu().numpy() for k in ("qx", "qy", "qz")]
    num = sum(float(np.sum((g - r) ** 2)) for g, r in zip(got, ref))
    den = sum(float(np.sum(r ** 2)) for r in ref)
    assert np.sqrt(num / den) < 1e-5


def test_graph_replays_match_eager(whhd):
    b = make(whhd, DIMS)
    fill(b, 0)
    b.Capture()
    for r in range(3):
        fill(b, 100 * (r + 1))          # new inputs, written eagerly between replays
        b.Replay()
        torch.cuda.synchronize()
        replay = snapshot(b)
        for i in range(len(DIMS)):
            for k in ("qx", "qy", "qz"):
                b.arrays(i)[k].zero_()
        b.Solve()
        torch.cuda.synchronize()
        eager = snapshot(b)
        for i in range(len(DIMS)):
            for g, e in zip(replay[i], eager[i]):
                assert torch.equal(g, e), f"replay {r}, problem {i}"


def test_autotune_keeps_results(whhd):
    b = make(whhd, DIMS)
    fill(b, 7)
    b.Solve()
    before = snapshot(b)
    b.Capture()
    b.Autotune(reps=2)                 # recaptures the graph with the chosen shapes
    for c in b.classes():
        assert c["lp_waves"] in (1, 2, 4) and c["proj_zb"] in (1, 2, 4, 8)
    b.Replay()
    torch.cuda.synchronize()
    after = snapshot(b)
    for i in range(len(DIMS)):
        for g, e in zip(after[i], before[i]):
            assert torch.equal(g, e), f"problem {i} changed after autotuning"
    with pytest.raises(whhd.WhhdError):
        b.SetTiles(0, 3, 4)
    with pytest.raises(whhd.WhhdError):
        b.Replay()                      # a failed SetTiles keeps the graph
        b.SetTiles(0, 2, 2)             # a good one drops it
        b.Replay()


def test_batched_workload_curl(whhd):
    """The batched workload's problem list (64 problems of 32^3 .. 128^3 with
    spheres made solid-compatible by the fractional multigrid, solved to 1e-5,
    above the FP32 iterate's residual floor): every problem's
    recovered potential reproduces its velocity, curl q = u, to 1e-5."""
    from waveletpotentials_amd import workloads
    probs = I.batch_problems(64)
    b = whhd.CuBatchRecover3D([p["dims"] for p in probs], [p["dx"] for p in probs], "n",
                              whhd.WaveletType.WT_CDF3_7, whhd.WaveletType.WT_CDF4_6)
    assert {c["dims"] for c in b.classes()} <= set(I.BATCH_SIZES)
    for i, p in enumerate(probs):
        q = I.spectral_potential3d(p["dims"], p["kmax"], p["seed"], p["extent"])
        a = b.arrays(i)
        for key, v in zip(("xv", "yv", "zv"), I.curl3d(*q, p["dx"])):
            a[key].copy_(torch.from_numpy(v).to(device="cuda", dtype=torch.float32))
        cycles, res = workloads.solid_velocity(a["xv"], a["yv"], a["zv"], p["dims"], p["dx"],
                                               p["centers"], p["radii"])
        assert res <= 1e-5, f"problem {i}: multigrid stopped at {res:.2e} after {cycles} cycles"
    b.Solve()
    for i, p in enumerate(probs):
        a = b.arrays(i)
        c = curl_t(a["qx"].double(), a["qy"].double(), a["qz"].double(), p["dx"])
        u = [a[k].double() for k in ("xv", "yv", "zv")]
        num = sum(float(((x - y) ** 2).sum()) for x, y in zip(c, u))
        den = sum(float((y ** 2).sum()) for y in u)
        assert (num / den) ** 0.5 < 1e-5, f"problem {i} {p['dims']}"
