"""Benchmark: 3D vector-potential recovery on a 256^3 MAC grid, FP32, one MI355X.

A step is one CuWaveletPotentialRecover3D::Solve (9 forward line transforms,
the per-cell projection, 9 inverse line transforms) over a synthetic
divergence-free velocity (curl of a seeded Fourier potential, |k| <= 16,
|k|^(-11/6) amplitudes) that is resident in HBM before timing starts.

Prints one JSON line.  `value` is Mcells/s of the whole solve; `roofline`
describes the kernel with the largest share of the solve, timed per launch
with HIP events on the stream the library launches on; `cpu_baseline` times
the FP64 C oracle (a port of the reference algorithm) on this host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "potential-recovery Mcells/s @256³ fp32 (% HBM peak); learned-init MG speedup"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=256, help="cells per axis")
    p.add_argument("--stage-reps", type=int, default=20)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-runs", type=int, default=5)
    return p.parse_args()


def stage_plan(N):
    """The 19 launches of one solve: (name, array, dims (nx,ny,nz), kind, args)."""
    xv, yv, zv = (N + 1, N, N), (N, N + 1, N), (N, N, N + 1)
    qx, qy, qz = (N, N + 1, N + 1), (N + 1, N, N + 1), (N + 1, N + 1, N)
    C, D = ("n", 11), ("d", 7)  # (bc, wavelet): curl CDF3.7/'n', div CDF4.6/'d'
    fwd = [("xv", xv, "x", D), ("xv", xv, "y", C), ("xv", xv, "z", C),
           ("yv", yv, "x", C), ("yv", yv, "y", D), ("yv", yv, "z", C),
           ("zv", zv, "x", C), ("zv", zv, "y", C), ("zv", zv, "z", D)]
    inv = [("qx", qx, "x", C), ("qx", qx, "y", D), ("qx", qx, "z", D),
           ("qy", qy, "x", D), ("qy", qy, "y", C), ("qy", qy, "z", D),
           ("qz", qz, "x", D), ("qz", qz, "y", D), ("qz", qz, "z", C)]
    return fwd, inv


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); the single-GPU number is the whole-job value.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import waveletpotentials_amd as W
    from waveletpotentials_amd import _lib
    from waveletpotentials_amd import inputs as I

    N = args.n
    n = (N, N, N)
    cells = N ** 3
    T = W.WaveletType
    torch.cuda.set_device(0)

    (xv, yv, zv), _, dx = I.velocity3d(n, 16, I.seed_for(1))
    dev = lambda a: torch.tensor(a, dtype=torch.float32, device="cuda")  # noqa: E731
    d_xv, d_yv, d_zv = dev(xv), dev(yv), dev(zv)
    qx = torch.empty((N + 1, N, N + 1), device="cuda")
    qy = torch.empty((N + 1, N + 1, N), device="cuda")
    qz = torch.empty((N, N + 1, N + 1), device="cuda")
    rec = W.CuWaveletPotentialRecover3D(n, dx, "n", T.WT_CDF3_7, T.WT_CDF4_6)

    for _ in range(args.warmup):
        rec.Solve(qx, qy, qz, d_xv, d_yv, d_zv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rec.Solve(qx, qy, qz, d_xv, d_yv, d_zv)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = cells * args.steps / elapsed / 1e6

    # accuracy of the timed result
    q_gpu = [t.cpu().numpy().astype(np.float64) for t in (qx, qy, qz)]
    curl = I.curl3d(*q_gpu, dx)
    curl_err = float(np.sqrt(sum(np.sum((a - b) ** 2) for a, b in zip(curl, (xv, yv, zv)))
                             / sum(np.sum(b ** 2) for b in (xv, yv, zv))))
    max_div = float(np.abs(I.div3d(*curl, dx)).max())

    # ---- per-launch timing of every stage with HIP events on the library's stream ----
    L = _lib.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fwd, inv = stage_plan(N)
    bufs = {"xv": d_xv.clone(), "yv": d_yv.clone(), "zv": d_zv.clone(), "qx": qx, "qy": qy,
            "qz": qz}
    lv = int(np.log2(N))
    stages = []

    def time_launch(fn, reps):
        for _ in range(2):
            fn()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        e.synchronize()
        return s.elapsed_time(e) / reps  # ms

    scratch = {k: torch.empty_like(v) for k, v in bufs.items()}
    for name, dims, axis, (bc, wt) in fwd + inv:
        src = ctypes.c_void_p(bufs[name].data_ptr())
        dst = ctypes.c_void_p(scratch[name].data_ptr())
        is_fwd = (name[0] != "q")

        def fn(src=src, dst=dst, dims=dims, axis=axis, bc=bc, wt=wt, is_fwd=is_fwd):
            _lib.check(L.whhd_transform1d(dst, src, *dims, lv, axis.encode(), bc.encode(), wt,
                                          int(is_fwd), stream), "transform1d")
        ms = time_launch(fn, args.stage_reps)
        nbytes = 2 * 4 * int(np.prod(dims))
        stages.append({"stage": f"{'fwd' if is_fwd else 'inv'} {name} {axis}",
                       "kernel": f"lp_kernel<{'CDF4.6 d' if wt == 7 else 'CDF3.7 n'},L={lv},"
                                 f"{'fwd' if is_fwd else 'inv'},"
                                 f"{'contiguous' if axis == 'y' else 'strided'}>",
                       "us": ms * 1e3, "bytes": nbytes})
    proj = W.CuDivLocalProject3D((N + 1,) * 3, (lv,) * 3, dx, T.WT_CDF4_6, "d")
    ms = time_launch(lambda: proj.ProjectLocal_q_ccc(qx, qy, qz, bufs["xv"], bufs["yv"],
                                                     bufs["zv"], (lv,) * 3), args.stage_reps)
    pbytes = 4 * (sum(int(np.prod(s)) for s in ((N + 1, N, N),) * 3)
                  + sum(int(np.prod(s)) for s in ((N + 1, N + 1, N),) * 3))
    stages.append({"stage": "project", "kernel": "project3d_kernel", "us": ms * 1e3,
                   "bytes": pbytes})
    # calibration copy of a known size (for scaling rocprof byte counters)
    cal_src = torch.empty(16 * 1024 * 1024, device="cuda")
    cal_dst = torch.empty_like(cal_src)
    for _ in range(3):
        torch.mul(cal_src, 1.0, out=cal_dst)  # an elementwise kernel distinct from the copies
    torch.cuda.synchronize()
    stage_total_us = sum(s["us"] for s in stages)
    stage_bytes = sum(s["bytes"] for s in stages)
    # dominant kernel = the instantiation with the largest summed time
    by_kernel = {}
    for s in stages:
        k = by_kernel.setdefault(s["kernel"], {"us": 0.0, "bytes": 0, "launches": 0})
        k["us"] += s["us"]
        k["bytes"] += s["bytes"]
        k["launches"] += 1
    dom_name, dom = max(by_kernel.items(), key=lambda kv: kv[1]["us"])
    dom_us = dom["us"] / dom["launches"]
    dom_bytes = dom["bytes"] / dom["launches"]
    achieved = dom_bytes / (dom_us * 1e-6) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (curl of a seeded Fourier potential, |k|<=16, |k|^-11/6)",
        "config": {"workload": f"3D vector-potential recovery {N}^3 MAC grid, curl CDF3.7 'n', "
                               f"div CDF4.6 'd', 19 stages", "cells": cells},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "algorithmic_bytes_per_launch": int(dom_bytes),
                     "avg_launch_us": round(dom_us, 2)},
        "stage_model": {"bytes_per_solve": int(stage_bytes),
                        "solve_GBs": round(stage_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                        "solve_frac_of_peak": round(stage_bytes / (ms_per_step * 1e-3) / 1e9
                                                    / HBM_PEAK_GBS, 4),
                        "sum_of_isolated_stage_us": round(stage_total_us, 1),
                        "stages": [{k: (round(v, 2) if isinstance(v, float) else v)
                                    for k, v in s.items()} for s in stages]},
        "accuracy": {"curl_q_vs_u_rel_l2": curl_err, "max_abs_div_curl_q": max_div},
    }

    if not args.no_cpu_baseline:
        import oracle as O
        O.build()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        runs = []
        ref = None
        for i in range(args.cpu_runs + 1):
            t = time.perf_counter()
            ref = O.recover3d(xv, yv, zv, n, dx)
            if i > 0:
                runs.append(time.perf_counter() - t)
        med = float(np.median(runs))
        out["cpu_baseline"] = {"value": round(cells / med / 1e6, 2), "unit": "Mcells/s",
                               "cores": threads, "kind": "port",
                               "sample": f"full {N}^3 FP64 recovery by the C oracle (OpenMP), "
                                         f"1 warm-up + median of {args.cpu_runs}"}
        out["accuracy"]["q_vs_fp64_oracle_rel_l2"] = O.rel_l2(ref, q_gpu)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
