"""Benchmark: 3D vector-potential recovery on a 256^3 MAC grid, FP32, one MI355X.

A step is one CuWaveletPotentialRecover3D::Solve (9 forward line transforms,
the per-cell projection, 9 inverse line transforms) over a synthetic
divergence-free velocity (curl of a seeded Fourier potential, |k| <= 16,
|k|^(-11/6) amplitudes) that is resident in HBM before timing starts.

Prints one JSON line.  `value` is Mcells/s of the whole solve (K solves
between device synchronizations).  `roofline` describes the kernel with the
largest share of the solve: a second run of K solves records HIP events after
every launch on the stream the library launches on (the library's stage
trace), giving each launch's duration in context.  `cpu_baseline` times the
C oracle (a port of the reference algorithm) built in FP32 on this host, with
the FP64 oracle's figure beside it.  The line also carries `learned_init`
(the learned initial guess against the multigrid from zero on held-out 192^3
seeds, with the V-cycle refinement as the solver and, under `pcg`, with
conjugate gradients) and `secondary` (the batch, solids and sweep workloads'
figures from a few steps each; the solids step solves by conjugate gradients
with a V-cycle preconditioner; `--workload` runs any of them alone).
"""
import argparse
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
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-runs", type=int, default=5)
    p.add_argument("--no-learned", action="store_true",
                   help="recover workload: skip the learned-init multigrid measurement")
    p.add_argument("--learned-seeds", type=int, default=4)
    p.add_argument("--no-secondary", action="store_true",
                   help="recover workload: skip the batch, solids and sweep figures")
    p.add_argument("--workload", choices=("recover", "batch", "solids", "learned", "sweep"),
                   default="recover",
                   help="recover: one N^3 grid (the headline); batch: 64 grids of 32^3..128^3 "
                        "with spheres, one HIP graph per batch; solids: 192^3 with 8 spheres, "
                        "fractional multigrid (conjugate gradients, V-cycle preconditioner) to "
                        "1e-6 and the velocity correction; learned: "
                        "the same solve from the U-Net's prediction against from zero; sweep: "
                        "the sweep + DST comparator on the same N^3 input as recover")
    return p.parse_args()


def array_bytes(N):
    """Bytes of each MAC array of an N^3 grid (FP32)."""
    shapes = {"xv": (N + 1, N, N), "yv": (N, N + 1, N), "zv": (N, N, N + 1),
              "qx": (N, N + 1, N + 1), "qy": (N + 1, N, N + 1), "qz": (N + 1, N + 1, N)}
    return {k: 4 * v[0] * v[1] * v[2] for k, v in shapes.items()}


def stage_bytes(stage, ab):
    """Algorithmic bytes of one traced stage: a line pass reads and writes its
    array once; the projection, alone or fused with the potentials' inverse z
    pass, reads three velocity and writes three potential arrays once."""
    if stage.startswith("project"):
        return sum(ab[k] for k in ("xv", "yv", "zv", "qx", "qy", "qz"))
    return 2 * ab[stage.split()[1]]


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprof counter
    summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by the calibration
    kernel's reading), or None when the summary does not cover it."""
    path = os.path.join(ROOT, "profiles", "traffic_recover3d_256.json")
    try:
        with open(path) as f:
            rows = json.load(f)
    except (OSError, ValueError):
        return None
    row = rows.get("kernels", {}).get(kernel)
    if not row:
        return None
    return {"bytes": int(row["fetch_bytes"] + row["write_bytes"]),
            "source": "profiles/traffic_recover3d_256.json"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); the single-GPU number is the whole-job value.", file=sys.stderr)
        sys.exit(2)

    others = {"batch": bench_batch, "solids": bench_solids, "learned": bench_learned,
              "sweep": bench_sweep}
    if args.workload in others:
        print(json.dumps(others[args.workload](args)))
        return

    import numpy as np
    import torch

    import waveletpotentials_amd as W
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

    # ---- per-launch durations in context: HIP events recorded by the library
    # after every launch of K further solves, on the stream it launches on ----
    rec.TraceBegin(args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rec.Solve(qx, qy, qz, d_xv, d_yv, d_zv)
    torch.cuda.synchronize()
    traced_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    stages, solves = rec.TraceEnd()
    ab = array_bytes(N)
    for st in stages:
        st["bytes"] = stage_bytes(st["stage"], ab)
    calibrate()
    stage_total_us = sum(st["us"] for st in stages)
    bytes_per_solve = sum(st["bytes"] for st in stages)
    # dominant kernel = the instantiation with the largest summed time
    by_kernel = {}
    for st in stages:
        k = by_kernel.setdefault(st["kernel"], {"us": 0.0, "bytes": 0, "launches": 0})
        k["us"] += st["us"]
        k["bytes"] += st["bytes"]
        k["launches"] += 1
    dom_name, dom = max(by_kernel.items(), key=lambda kv: kv[1]["us"])
    dom_us = dom["us"] / dom["launches"]
    dom_bytes = dom["bytes"] / dom["launches"]
    achieved = dom_bytes / (dom_us * 1e-6) / 1e9
    traffic = measured_traffic(dom_name)

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
                               f"div CDF4.6 'd', {len(stages)} launches", "cells": cells},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes"] if traffic else None,
                     "kernel": dom_name, "launches_per_solve": dom["launches"],
                     "algorithmic_bytes_per_launch": int(dom_bytes),
                     "avg_launch_us": round(dom_us, 2),
                     "traffic_source": traffic["source"] if traffic else None},
        "stage_model": {"bytes_per_solve": int(bytes_per_solve),
                        "solve_GBs": round(bytes_per_solve / (ms_per_step * 1e-3) / 1e9, 1),
                        "solve_frac_of_peak": round(bytes_per_solve / (ms_per_step * 1e-3) / 1e9
                                                    / HBM_PEAK_GBS, 4),
                        "traced_ms_per_step": round(traced_ms, 4), "traced_solves": solves,
                        "sum_of_stage_us": round(stage_total_us, 1),
                        "stages": [{k: (round(v, 2) if isinstance(v, float) else v)
                                    for k, v in st.items()} for st in stages]},
        "accuracy": {"curl_q_vs_u_rel_l2": curl_err, "max_abs_div_curl_q": max_div},
    }

    if not args.no_cpu_baseline:
        import oracle as O
        O.build()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        def timed(fn):
            runs, res = [], None
            for i in range(args.cpu_runs + 1):
                t = time.perf_counter()
                res = fn(xv, yv, zv, n, dx)
                if i > 0:
                    runs.append(time.perf_counter() - t)
            return cells / float(np.median(runs)) / 1e6, res

        v32, _ = timed(O.recover3d_f32)
        v64, ref = timed(O.recover3d)
        out["cpu_baseline"] = {"value": round(v32, 2), "unit": "Mcells/s",
                               "cores": threads, "kind": "port",
                               "sample": f"full {N}^3 recovery by the C oracle built in FP32 "
                                         f"(OpenMP), 1 warm-up + median of {args.cpu_runs}",
                               "fp64_value": round(v64, 2),
                               "fp64_sample": "the same with the FP64 oracle (the parity checker)"}
        out["accuracy"]["q_vs_fp64_oracle_rel_l2"] = O.rel_l2(ref, q_gpu)
    else:
        out["cpu_baseline"] = None
    if not args.no_learned:
        out["learned_init"] = learned_summary(args.learned_seeds)
    if not args.no_secondary:
        out["secondary"] = secondary(args)
    print(json.dumps(out))


def secondary(args):
    """The other workloads' lines (batch, solids, sweep), measured in the same
    run with a few steps each and no CPU legs, condensed to their figures."""
    sub = argparse.Namespace(**{**vars(args), "steps": 5, "warmup": 1, "no_cpu_baseline": True})
    out = {}
    for name, fn in (("batch", bench_batch), ("solids", bench_solids), ("sweep", bench_sweep)):
        r = fn(sub)
        row = {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "steps")}
        rf = r.get("roofline") or {}
        row["roofline_frac"] = rf.get("frac")
        for k in ("v_cycle_ms", "dst_solve_ms", "kernel"):
            if k in rf:
                row[k] = rf[k]
        if "accuracy" in r:
            row["accuracy"] = r["accuracy"]
        out[name] = row
    return out


def calibrate():
    """A kernel of known size (16 Mi floats read, 16 Mi written) for scaling the
    rocprof byte counters of a profiled run."""
    import torch
    cal_src = torch.empty(16 * 1024 * 1024, device="cuda")
    cal_dst = torch.empty_like(cal_src)
    for _ in range(3):
        torch.mul(cal_src, 1.0, out=cal_dst)
    torch.cuda.synchronize()


def learned_summary(n_seeds):
    """The second half of the metric: the fractional multigrid on held-out 192^3
    solid-obstacle seeds (8 spheres), GPU time and V-cycles to an FP64 relative
    residual of 1e-6 from zero and from the U-Net's prediction (prediction
    included); speedup = summed plain time / summed learned time, with the
    V-cycle refinement as the solver, and under `pcg` the same with conjugate
    gradients preconditioned by a V-cycle."""
    import statistics

    import waveletpotentials_amd as W
    from waveletpotentials_amd import train_unet as TU

    n = (192, 192, 192)
    ckpt = os.path.join(ROOT, "waveletpotentials_amd", "weights", "unet_solids192.bin")
    factor = W.UNet3D.checkpoint_factor(ckpt)
    net = W.UNet3D(n, factor=factor)
    net.load(ckpt)
    rows = TU.evaluate(net, n, [TU.EVAL_SEED0 + i for i in range(max(1, n_seeds))])
    pm = sum(r["plain_ms"] for r in rows)
    lm = sum(r["learned_ms"] for r in rows)
    return {"speedup": round(pm / lm, 3), "target": 2.0, "seeds": len(rows),
            "network_grid": f"{192 // factor}^3 (factor {factor})",
            "plain_cycles_mean": statistics.mean(r["plain_cycles"] for r in rows),
            "learned_cycles_mean": statistics.mean(r["learned_cycles"] for r in rows),
            "plain_ms_mean": round(pm / len(rows), 3), "learned_ms_mean": round(lm / len(rows), 3),
            "plain_res_max": max(r["plain_res"] for r in rows),
            "learned_res_max": max(r["learned_res"] for r in rows),
            "prediction_rel_err_mean": statistics.mean(r["prediction_rel_err"] for r in rows),
            # the same comparison with conjugate gradients (a V-cycle as the
            # preconditioner) as the solver of both arms
            "pcg": {"speedup": round(sum(r["pcg_plain_ms"] for r in rows)
                                     / sum(r["pcg_learned_ms"] for r in rows), 3),
                    "plain_iters_mean": statistics.mean(r["pcg_plain_iters"] for r in rows),
                    "learned_iters_mean": statistics.mean(r["pcg_learned_iters"] for r in rows),
                    "plain_ms_mean": round(statistics.mean(r["pcg_plain_ms"] for r in rows), 3),
                    "learned_ms_mean": round(statistics.mean(r["pcg_learned_ms"] for r in rows), 3)},
            "rows": rows}


def bench_batch(args):
    """64 independent grids (extents drawn from 32^3, 64^3, 128^3, 128x64x64; 0-4
    spheres each, the velocity made solid-compatible by the fractional multigrid
    before timing), recovered by one captured HIP graph per step after the
    autotuner picked each size class's launch shapes."""
    import numpy as np
    import torch

    import waveletpotentials_amd as W
    from waveletpotentials_amd import inputs as I
    from waveletpotentials_amd import workloads

    T = W.WaveletType
    torch.cuda.set_device(0)
    probs = I.batch_problems(64)
    b = W.CuBatchRecover3D([p["dims"] for p in probs], [p["dx"] for p in probs], "n",
                           T.WT_CDF3_7, T.WT_CDF4_6)
    mg_cycles = 0
    for i, p in enumerate(probs):
        q = I.spectral_potential3d(p["dims"], p["kmax"], p["seed"], p["extent"])
        a = b.arrays(i)
        for key, v in zip(("xv", "yv", "zv"), I.curl3d(*q, p["dx"])):
            a[key].copy_(torch.from_numpy(v).to(device="cuda", dtype=torch.float32))
        mg_cycles += workloads.solid_velocity(a["xv"], a["yv"], a["zv"], p["dims"], p["dx"],
                                              p["centers"], p["radii"])[0]
    torch.cuda.synchronize()
    print(f"batch inputs ready ({len(probs)} problems, {mg_cycles} multigrid cycles)",
          file=sys.stderr, flush=True)
    b.Capture()
    b.Autotune(reps=3)
    calibrate()
    for _ in range(args.warmup):
        b.Replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.Replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cells = sum(p["dims"][0] * p["dims"][1] * p["dims"][2] for p in probs)
    ms_per_step = elapsed * 1e3 / args.steps
    # stage model: each of the 9 forward passes reads and writes a velocity
    # array, the projection reads the three velocity and writes the three
    # potential arrays, each of the 9 inverse passes reads and writes a
    # potential array: 7 (S_v + S_q) floats per problem
    model_bytes = 0
    for p in probs:
        nx, ny, nz = p["dims"]
        sv = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
        sq = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
        model_bytes += 4 * 7 * (sv + sq)
    achieved = model_bytes / (ms_per_step * 1e-3) / 1e9

    def rel(ref, got):
        num = sum(float(np.sum((g - r) ** 2)) for g, r in zip(got, ref))
        return (num / sum(float(np.sum(r ** 2)) for r in ref)) ** 0.5

    # per-problem accuracy of the timed result: curl q = u; with the CPU
    # baseline, also q against the FP64 oracle run on the same inputs
    vs, qs = [], []
    worst_curl = 0.0
    for i, p in enumerate(probs):
        a = b.arrays(i)
        vs.append([a[k].double().cpu().numpy() for k in ("xv", "yv", "zv")])
        qs.append([a[k].double().cpu().numpy() for k in ("qx", "qy", "qz")])
        worst_curl = max(worst_curl, rel(vs[-1], I.curl3d(*qs[-1], p["dx"])))
    out = {
        "metric": "batched potential-recovery Mcells/s (64 grids, one HIP graph)",
        "value": round(cells * args.steps / elapsed / 1e6, 1),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (per problem: curl of a seeded Fourier potential, |k|<=min(n)/8, "
                "spheres made solid-compatible by the fractional multigrid)",
        "config": {"workload": "batched 3D recovery, 64 grids of 32^3/64^3/128^3/128x64x64, "
                               "curl CDF3.7 'n', div CDF4.6 'd'", "cells": cells,
                   "classes": b.classes()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batch (stage-model bytes / step time)",
                     "algorithmic_bytes_per_step": int(model_bytes)},
        "accuracy": {"worst_curl_q_vs_u_rel_l2": worst_curl},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        import oracle as O
        O.build()
        cpu_s, worst_q = 0.0, 0.0
        for i, p in enumerate(probs):
            t = time.perf_counter()
            ref = O.recover3d(*vs[i], p["dims"], p["dx"])
            cpu_s += time.perf_counter() - t
            worst_q = max(worst_q, rel(ref, qs[i]))
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["accuracy"]["worst_q_vs_fp64_oracle_rel_l2"] = worst_q
        out["cpu_baseline"] = {"value": round(cells / cpu_s / 1e6, 2), "unit": "Mcells/s",
                               "cores": threads, "kind": "port",
                               "sample": "one pass of the C oracle over all 64 problems "
                                         "(FP64, OpenMP)"}
    return out


def _line(metric, value, unit, args, ms, data, config, higher=True, dtype="f32"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": higher,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": data,
            "config": config}


def bench_sweep(args):
    """The sweep + DST comparator (CuParallelSweep3D::Solve + Project, velocity
    bc 'n') on the recover workload's N^3 input: a step is the z sweep, the
    interior node divergence, the DST-I Poisson solve on the (N-1)^3 interior
    nodes (six line passes and the eigenvalue scaling) and the gradient
    correction."""
    import numpy as np
    import torch

    import waveletpotentials_amd as W
    from waveletpotentials_amd import inputs as I
    from waveletpotentials_amd import workloads as WL

    torch.cuda.set_device(0)
    N = args.n
    n = (N, N, N)
    cells = N ** 3
    dx = (1.0 / N,) * 3
    xv, yv, zv = WL.spectral_velocity(n, 16, I.seed_for(1))
    sw = W.CuParallelSweep3D(n, dx, "n")
    q = [torch.empty(sh, device="cuda") for sh in ((N + 1, N, N + 1), (N + 1, N + 1, N),
                                                    (N, N + 1, N + 1))]

    def step():
        sw.Solve(*q, xv, yv, zv)
        sw.Project(*q)

    calibrate()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # the DST solve alone, timed with events on the current stream
    b = torch.randn((N - 1, N - 1, N - 1), device="cuda")
    phi = torch.empty_like(b)
    dst = W.CuDst3D((N - 1, N - 1, N - 1), dx)
    dst.Solve(phi, b)
    e0, e1 = torch.cuda.Event(True), torch.cuda.Event(True)
    e0.record()
    for _ in range(5):
        dst.Solve(phi, b)
    e1.record()
    torch.cuda.synchronize()
    dst_ms = e0.elapsed_time(e1) / 5
    # the DST's algorithmic bytes: six line passes and the scaling, each reading
    # and writing the (N-1)^3 array once
    dst_bytes = 7 * 2 * 4 * (N - 1) ** 3
    achieved = dst_bytes / (dst_ms * 1e-3) / 1e9
    # accuracy: divergence of the projected potential's curl against the input
    import oracle as O
    got = [t.double().cpu().numpy() for t in q]
    u = [t.double().cpu().numpy() for t in (xv, yv, zv)]
    curl = O.curl3d(*got, n, dx)
    out = _line("sweep + DST comparator Mcells/s @%d^3" % N, round(cells / (ms * 1e-3) / 1e6, 1),
                "Mcells/s", args, ms, "synthetic (curl of a seeded Fourier potential, |k|<=16)",
                {"workload": f"CuParallelSweep3D Solve + Project {N}^3, velocity bc 'n', DST-I on "
                             f"{N - 1}^3 interior nodes", "cells": cells})
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "kernel": "DST-I Poisson solve (7 streaming stages / solve time)",
                       "dst_solve_ms": round(dst_ms, 4), "algorithmic_bytes": dst_bytes}
    out["accuracy"] = {"curl_q_vs_u_rel_l2": O.rel_l2(u, list(curl))}
    out["cpu_baseline"] = None
    return out


def bench_solids(args):
    """192^3 with 8 spheres: a step is the fractional multigrid solve from zero
    to a relative residual of 1e-6 (conjugate gradients preconditioned by one
    FP32 V-cycle with 2 + 2 weighted-Jacobi sweeps, two-float iterate, FP64
    residuals) plus the velocity correction u <- a (u - grad p).  The level
    set, fractions and f are built before timing.  One solve by the V-cycle
    refinement alone is timed after the steps, for comparison."""
    import numpy as np
    import torch

    import waveletpotentials_amd as W
    from waveletpotentials_amd import inputs as I
    from waveletpotentials_amd import workloads as WL

    torch.cuda.set_device(0)
    n = (192, 192, 192)
    cells = 192 ** 3
    P = WL.solid_problem(n, I.seed_for(2))
    mg = W.CuMultigridFrac3D(P["ax"], P["ay"], P["az"], n, P["dx"], "n")
    grad = W.CuGradient.GetInstance()
    f = P["f"]
    u0 = [P[k].clone() for k in ("xv", "yv", "zv")]
    u = [t.clone() for t in u0]

    calibrate()

    def step():
        for a, b in zip(u, u0):
            a.copy_(b)
        hi, lo, cyc, res = WL.mg_solve_from(mg, f, None, 1e-6, 400, "pcg")
        grad.SubtractGradient3D_Frac_P2(*u, hi, lo, P["ax"], P["ay"], P["az"], n, P["dx"], "n")
        return cyc, res

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cycles, res = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    t = time.perf_counter()
    _, _, ref_cycles, ref_res = WL.mg_solve_from(mg, f, None, 1e-6, 400, "vcycles")
    torch.cuda.synchronize()
    ref_ms = (time.perf_counter() - t) * 1e3
    # V-cycle time and the stage model's bytes per cycle (per finest cell: 4
    # smoothing sweeps x 6 + residual with its restriction 5.125 + the coarse
    # correction the first post-smoothing sweep prolongs 0.125 = 29.25 floats,
    # x 8/7 for the coarser levels)
    v = torch.zeros_like(f)
    e0, e1 = torch.cuda.Event(True), torch.cuda.Event(True)
    e0.record()
    mg.Solve(v, v, f, 20, 2, 2, 0.8)
    e1.record()
    torch.cuda.synchronize()
    vc_ms = e0.elapsed_time(e1) / 20
    cyc_bytes = 29.25 * 8.0 / 7.0 * 4 * cells
    achieved = cyc_bytes / (vc_ms * 1e-3) / 1e9
    # divergence of the corrected velocity a (u - grad p), which the solve makes
    # zero (the reference checks div(a (u - grad p)) the same way,
    # TestLocalProjection.cpp:690-696)
    b = torch.empty_like(f)
    ones = [torch.ones_like(t) for t in u]
    W.CuConvergence.GetInstance().GetB_3d_frac_vel(b, *u, *ones, n, P["dx"], "n")
    out = _line("fractional-multigrid solve to 1e-6 + velocity correction, Mcells/s @192^3",
                round(cells / (ms * 1e-3) / 1e6, 1), "Mcells/s", args, ms,
                "synthetic (8 seeded spheres, fast-swept level set, spectral velocity |k|<=12)",
                {"workload": "solid-obstacle projection 192^3, 8 spheres, conjugate gradients "
                             "preconditioned by a V-cycle (2+2 weighted Jacobi w=0.8) to "
                             "relative residual 1e-6", "cells": cells,
                 "v_cycles": cycles, "fast_sweeps": P["sweeps"]})
    out["v_cycle_refinement"] = {"v_cycles": ref_cycles, "ms": round(ref_ms, 3),
                                 "relative_residual": ref_res}
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "kernel": "one V-cycle (stage-model bytes / V-cycle time)",
                       "v_cycle_ms": round(vc_ms, 4), "algorithmic_bytes_per_launch": int(cyc_bytes)}
    out["accuracy"] = {"relative_residual": res,
                       "max_abs_div_after_over_before": float(b.abs().max() / f.abs().max())}
    out["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        import oracle as O
        O.build()
        fr = [t.double().cpu().numpy() for t in (P["ax"], P["ay"], P["az"])]
        fh = f.double().cpu().numpy()
        t = time.perf_counter()
        O.mg_solve(fh, *fr, P["dx"], "n", n_iters=2, nu1=2, nu2=2, weight=0.8)
        cpu_cycle = (time.perf_counter() - t) / 2
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": round(cells / (cpu_cycle * cycles) / 1e6, 3),
                               "unit": "Mcells/s", "cores": threads, "kind": "port",
                               "sample": f"2 FP64 V-cycles of the C oracle at 192^3, "
                                         f"scaled to the {cycles} V-cycles the GPU's "
                                         f"conjugate-gradient solve applies"}
    return out


def bench_learned(args):
    """The learned initial guess on held-out 192^3 solid-obstacle seeds: GPU time
    and V-cycles to a relative residual of 1e-6 from the U-Net's prediction
    (prediction included) against from zero; value = plain time / learned time."""
    import statistics

    import torch

    import waveletpotentials_amd as W
    from waveletpotentials_amd import train_unet as TU

    torch.cuda.set_device(0)
    n = (192, 192, 192)
    ckpt = os.path.join(ROOT, "waveletpotentials_amd", "weights", "unet_solids192.bin")
    factor = W.UNet3D.checkpoint_factor(ckpt)
    net = W.UNet3D(n, factor=factor)
    net.load(ckpt)
    seeds = [TU.EVAL_SEED0 + i for i in range(max(1, args.steps))]
    rows = TU.evaluate(net, n, seeds)
    pm = statistics.mean(r["plain_ms"] for r in rows)
    lm = statistics.mean(r["learned_ms"] for r in rows)
    f = torch.empty(n[::-1], device="cuda")
    e0, e1 = torch.cuda.Event(True), torch.cuda.Event(True)
    p0 = torch.empty_like(f)
    ls = torch.zeros_like(f)
    f.normal_()
    net.predict(p0, f, ls, 1.0)
    e0.record()
    for _ in range(5):
        net.predict(p0, f, ls, 1.0)
    e1.record()
    torch.cuda.synchronize()
    pred_ms = e0.elapsed_time(e1) / 5
    flops = 2.0 * (192 // factor) ** 3 * 27 * sum(
        (ca + cb) * co / (8 ** lev) for ca, cb, co, lev in
        ((16, 0, 16, 0), (16, 0, 16, 0), (16, 0, 32, 1), (32, 0, 32, 1), (32, 0, 64, 2),
         (64, 0, 64, 2), (64, 32, 32, 1), (32, 0, 32, 1), (32, 16, 16, 0), (16, 0, 16, 0),
         (16, 0, 1, 0)))
    out = _line("learned-init MG speedup (GPU time to 1e-6, plain / learned) @192^3",
                round(pm / lm, 3), "x", args, lm,
                "synthetic held-out seeds of the solid-obstacle workload",
                {"workload": "fractional multigrid 192^3, 8 spheres, to relative residual 1e-6, "
                             "from zero vs from the 3-level U-Net prediction",
                 "network_grid": f"{192 // factor}^3 (factor {factor})",
                 "seeds": len(rows),
                 "plain_cycles_mean": statistics.mean(r["plain_cycles"] for r in rows),
                 "learned_cycles_mean": statistics.mean(r["learned_cycles"] for r in rows),
                 "plain_ms_mean": round(pm, 3), "learned_ms_mean": round(lm, 3),
                 "prediction_rel_err_mean": statistics.mean(r["prediction_rel_err"] for r in rows)},
                dtype="bf16")
    achieved = flops / (pred_ms * 1e-3) / 1e12
    out["roofline"] = {"bound": "mfma", "achieved": round(achieved, 1), "peak": 2500.0,
                       "unit": "TFLOP/s", "frac": round(achieved / 2500.0, 4), "traffic": None,
                       "kernel": "U-Net forward (algorithmic FLOPs / prediction time)",
                       "prediction_ms": round(pred_ms, 3)}
    out["cpu_baseline"] = None
    out["rows"] = rows
    return out


if __name__ == "__main__":
    main()
