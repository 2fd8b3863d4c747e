"""Train and evaluate the learned initial guess for the solid-obstacle multigrid.

Training data are solves the method itself produces: for each seed, the
solid-obstacle workload (workloads.solid_problem: 8 spheres, fast-swept level
set, face fractions, spectral velocity, f = -div(a u)) is solved by the
fractional multigrid to a relative residual of 1e-6, and the network learns
(f, level set) -> p.  Everything runs on the GPU with this package's kernels;
the network's forward and backward passes are the library's own (UNet3D).

    python -m waveletpotentials_amd.train_unet --n 192 --samples 2000 --epochs 2 \\
        --out unet.bin

Evaluation compares, on held-out seeds, V-cycles and GPU time to 1e-6 from
zero against from the network's prediction (prediction time included).
"""
import argparse
import json
import math
import sys
import time

import torch

from . import api
from . import inputs as I
from . import workloads as WL

TRAIN_SEED0 = I.seed_for(3)              # training seeds: TRAIN_SEED0 + i
EVAL_SEED0 = I.seed_for(3, 500000)       # held-out seeds


def make_sample(n, seed, tol=1e-6):
    """(f, level set, p, length, cycles) for one seed; p solves the problem to tol."""
    P = WL.solid_problem(n, seed)
    mg = api.CuMultigridFrac3D(P["ax"], P["ay"], P["az"], n, P["dx"], "n")
    hi, lo, cycles, res = WL.mg_solve_from(mg, P["f"], None, tol)
    p = hi + lo
    return P["f"], P["ls"], p, P["length"], cycles


def scale_set(net):
    return net.output_scale() != 1.0


def train(net, n, samples, epochs, lr, log=sys.stderr, bf16_cache=True, flips=True):
    """Adam on `samples` seeds for `epochs` passes; the first pass generates and
    solves the samples (kept on the device, BF16 when bf16_cache).  With flips,
    each step mirrors (f, level set, p) along a random subset of the axes: the
    walls are closed on every side, so a mirrored problem is another valid
    problem with the mirrored solution."""
    gen = torch.Generator().manual_seed(12345)
    cache = []
    losses = []
    t0 = time.time()
    for ep in range(epochs):
        for i in range(samples):
            if ep == 0:
                f, ls, p, length, cyc = make_sample(n, TRAIN_SEED0 + i)
                if i == 0 and not scale_set(net):
                    # output units from the first sample: |p| / (rms(f) L^2)
                    pc = p - p.mean()
                    ratio = float(pc.pow(2).mean().sqrt() / (f.pow(2).mean().sqrt() * length ** 2))
                    net.set_output_scale(ratio)
                    print(f"output scale {ratio:.4e}", file=log, flush=True)
                dt = torch.bfloat16 if bf16_cache else torch.float32
                cache.append((f.to(dt), ls.to(dt), p.to(dt), length))
            else:
                fb, lb, pb, length = cache[i]
                f, ls, p = fb.float(), lb.float(), pb.float()
            # cosine decay over all steps
            step = ep * samples + i
            total = epochs * samples
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            if flips:
                axes = [d for d in range(3) if torch.rand(1, generator=gen).item() < 0.5]
                if axes:
                    f, ls, p = (t.flip(axes).contiguous() for t in (f, ls, p))
            loss = net.loss_grad(f, ls, p, length)
            net.adam(lr_t)
            losses.append(loss)
            if (i + 1) % 100 == 0 or i == samples - 1:
                k = min(100, len(losses))
                print(f"epoch {ep} sample {i + 1}/{samples} loss(avg of last {k}) "
                      f"{sum(losses[-k:]) / k:.4e}  {time.time() - t0:.0f}s", file=log, flush=True)
    return losses


def evaluate(net, n, seeds, tol=1e-6, max_cycles=400):
    """Per held-out seed: V-cycles and GPU milliseconds to tol from zero and from
    the prediction (prediction included), plus the prediction's own residual;
    by the V-cycle refinement and (keys pcg_*) by conjugate gradients with a
    V-cycle as the preconditioner."""
    rows = []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    for s in seeds:
        P = WL.solid_problem(n, s)
        mg = api.CuMultigridFrac3D(P["ax"], P["ay"], P["az"], n, P["dx"], "n")
        f = P["f"]
        p0 = torch.empty_like(f)
        # warm both paths once (kernel loading), then time
        net.predict(p0, f, P["ls"], P["length"])
        WL.mg_solve_from(mg, f, None, 1e-2, 5)
        torch.cuda.synchronize()
        ev[0].record()
        hi, lo, c_plain, r_plain = WL.mg_solve_from(mg, f, None, tol, max_cycles)
        ev[1].record()
        net.predict(p0, f, P["ls"], P["length"])
        _, _, c_learn, r_learn = WL.mg_solve_from(mg, f, p0, tol, max_cycles)
        ev[2].record()
        _, _, g_plain, _ = WL.mg_solve_from(mg, f, None, tol, max_cycles, "pcg")
        ev[3].record()
        net.predict(p0, f, P["ls"], P["length"])
        _, _, g_learn, _ = WL.mg_solve_from(mg, f, p0, tol, max_cycles, "pcg")
        ev[4].record()
        torch.cuda.synchronize()
        r0 = mg.GetResdualError(p0, f)
        ps = hi.double() + lo.double()
        ps = ps - ps.mean()
        e = p0.double() - p0.double().mean() - ps
        rows.append({"seed": s, "plain_cycles": c_plain, "plain_ms": ev[0].elapsed_time(ev[1]),
                     "plain_res": r_plain, "learned_cycles": c_learn,
                     "learned_ms": ev[1].elapsed_time(ev[2]), "learned_res": r_learn,
                     "prediction_res": r0,
                     "prediction_rel_err": float(e.norm() / ps.norm()),
                     "pcg_plain_iters": g_plain, "pcg_plain_ms": ev[2].elapsed_time(ev[3]),
                     "pcg_learned_iters": g_learn, "pcg_learned_ms": ev[3].elapsed_time(ev[4])})
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=192)
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--factor", type=int, default=4,
                    help="the network runs on the grid this many times coarser")
    ap.add_argument("--eval", type=int, default=8, help="held-out seeds to evaluate")
    ap.add_argument("--init", default=None, help="start from this checkpoint")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    n = (a.n, a.n, a.n)
    net = api.UNet3D(n, seed=a.seed, factor=a.factor)
    if a.init:
        net.load(a.init)
    losses = train(net, n, a.samples, a.epochs, a.lr) if a.samples > 0 else []
    net.save(a.out)
    rows = evaluate(net, n, [EVAL_SEED0 + i for i in range(a.eval)])
    print(json.dumps({"n": a.n, "factor": a.factor, "samples": a.samples, "epochs": a.epochs,
                      "final_loss": losses[-1] if losses else None, "eval": rows}))


if __name__ == "__main__":
    main()
