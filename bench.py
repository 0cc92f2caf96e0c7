"""Benchmark of the gfx950 implicit PPFV step (Sayram-2D Solver::update).

Headline (default) workload: the mixed-precision batch -- 512 independent
instances with uneven grids (each dimension uniform in 32..512, per-instance D
seed and amplitude 0.1..10x) plus one 4096 x 4096 grid, advanced together by
the batched path. The default precision is SG_PC32: FP64 BiCGSTAB (operator,
vectors, residual, convergence test) preconditioned by an FP32 multigrid
V-cycle, with FP64 assembly and flux weights. --precision mixed measures the
FP32 Krylov solve inside FP64 iterative refinement, --precision fp64 the FP64
solve throughout. A step is one implicit time step of every instance; inputs
are resident in HBM before the timed region starts.

Beside it, on the same batch:
  * the batched step in each of the three precisions (a shorter run each);
  * the plain FP64 path -- the same kernels, one instance at a time, FP64
    throughout -- timed for the speed-up;
  * the error of the measured result against the plain FP64 GPU result, and
    of both against the CPU oracle on the smallest instances;
  * the CPU oracle (banded-LU restatement of Solver::update, one thread) timed
    on a bounded sample of the batch (cpu_baseline);
  * HIP-event timing of the dominant kernel (the level-0 zebra line sweeps)
    over the timed region, for the roofline.

Prints one JSON line on stdout (progress on stderr). --gpus above 1 is refused.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sayram-2d_amd"))

import sayram_cases as sc   # noqa: E402
import sayram_gpu as sg     # noqa: E402

METRIC = "cell-steps/s, batched mixed-prec vs plain FP64 at ≤1e-10 rel err; p99 latency"
DTYPES = {"pc32": "f64+f32", "mixed": "f32+f64", "fp64": "f64"}
PREC_NAMES = {"pc32": "FP64 BiCGSTAB with an FP32 multigrid preconditioner",
              "mixed": "FP32 BiCGSTAB inside FP64 iterative refinement",
              "fp64": "FP64 throughout"}
HBM_PEAK_GBS = 8000.0
KERNEL_DESC = "k_line<T,5,*,false,*> level-0 zebra line sweep dispatches (every tile shape)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_batch_problems(n_inst, lo, hi, big, seed):
    probs = [sc.seeded_case(s["nx"], s["ny"], seed=s["seed"], scale=s["scale"])
             for s in sc.batch_specs(n_inst, lo, hi, seed)]
    if big:
        probs.append(sc.seeded_case(big, big, seed=seed * 7 + 1, scale=1.0))
    return probs


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_batched(probs, warmup, steps, precision=sg.SG_MIXED, graph=True, time_kernel=True,
                opts=None, streams=0):
    """Batched steps. time_kernel brackets every level-0 sweep dispatch with
    HIP events; such a run launches eagerly, on one stream and in one instance
    group (streams=1, group_cells=-1), so that each dispatch's interval is
    that kernel alone."""
    if opts is None:
        opts = sg.make_opts(precision=precision, use_graph=1 if graph else -1, streams=streams)
    if time_kernel:
        opts.use_graph, opts.streams, opts.group_cells = -1, 1, -1
    b = sg.Batch(opts)
    slots = [b.add(p) for p in probs]
    b.step(warmup)
    if time_kernel:
        b.time_kernel("line_l0")
    t0 = time.perf_counter()
    b.step(steps)
    dt = time.perf_counter() - t0
    kms, klaunch, kbytes = b.kernel_time()
    st = [b.stats(s) for s in slots]
    b.close()
    return dt, (kms, klaunch, kbytes), st


def run_plain_fp64(probs, warmup, steps):
    """The plain FP64 path: the same kernels, one instance at a time."""
    total = 0.0
    fs = []
    for k, p in enumerate(probs):
        s = sg.Solver(p, sg.make_opts(precision=sg.SG_FP64))
        s.update(warmup)
        t0 = time.perf_counter()
        s.update(steps)
        total += time.perf_counter() - t0
        fs.append(s.f())
        s.close()
        if k % 128 == 127:
            log(f"[bench] plain FP64: {k + 1}/{len(probs)} instances")
    return total, fs


def batched_states(probs, nsteps, precision):
    b = sg.Batch(sg.make_opts(precision=precision))
    slots = [b.add(p) for p in probs]
    b.step(nsteps)
    fs = [b.f(s) for s in slots]
    b.close()
    return fs


def serving(n_req, seed, precision, pname):
    """Continuous batching under Poisson arrivals, on a
    bounded seeded trace. The full serving workload (200 req/s, grids
    log-uniform 32..2048 per side, 10..1000 steps) offers ~2.4e10 cell-steps/s,
    far beyond one GPU, so its latency grows with the trace length; the trace
    here keeps the arrival rate but draws grids in 32..512 and 10..100 steps,
    and the sustainable rate of the full distribution is derived from the
    measured serving throughput."""
    import sayram_serve as sv
    reqs = sv.poisson_trace(n_req, 200.0, 32, 512, 10, 100, seed=seed + 17)
    loop = sv.ServeLoop(hbm_budget=64 << 30, precision=precision)
    done, wall = loop.run(reqs, realtime=True)
    loop.close()
    s = sv.summarize(done, wall)
    e_cells = ((2048 - 32) / math.log(2048 / 32)) ** 2      # log-uniform sides
    e_steps = (10 + 1000) / 2
    s.update({"trace": f"{n_req} requests, Poisson 200 req/s, sides log-uniform 32..512, "
                       f"steps 10..100, precision {pname}, HBM budget 64 GiB",
              "sustainable_rate_full_distribution_req_per_s":
                  s["throughput_cell_steps_per_s"] / (e_cells * e_steps)})
    return s


def cpu_baseline(probs, budget_s):
    """CPU oracle (banded LU, 1 thread) on instances of the batch in order,
    one step each, until about budget_s seconds of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    cells = 0
    t_total = 0.0
    n = 0
    for p in probs:
        # banded LU costs N*nx^2: keep single steps under a few seconds
        if p.ncells * p.nx * p.nx > 4e10:
            continue
        op = orc.OracleProblem(p)
        t0 = time.process_time()
        orc.step(op, p.f0)
        t_total += time.process_time() - t0
        cells += p.ncells
        n += 1
        if t_total >= budget_s:
            break
    return cells / t_total, n, cells, t_total


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
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--precision", choices=["pc32", "mixed", "fp64"], default="pc32",
                    help="precision of the batched step that is measured")
    ap.add_argument("--graph", choices=["on", "off"], default="on",
                    help="replay the Krylov chunks from captured hipGraphs")
    ap.add_argument("--streams", type=int, choices=[0, 1, 2], default=0,
                    help="HIP streams of the line sweeps' tile shapes (0: library default, 2)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch parameters on the workload first (sayram_tune) and "
                         "time the step with the best set")
    ap.add_argument("--skip-precisions", action="store_true",
                    help="do not time the batched step in the other precisions")
    ap.add_argument("--oracle-checks", type=int, default=3)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=64,
                    help="requests in the serving trace (0 skips the serving measurement)")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the timed batched steps (for rocprofv3); with --graph off, only "
                         "the kernel-timing run (eager, one stream, HIP events per dispatch)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times: the known-size kernel that "
                         "scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {a.numel() * 8} bytes 3 times")
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the step of one GPU is what is measured "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if sg.lib().sg_device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    t0 = time.perf_counter()
    probs = make_batch_problems(args.instances, args.lo, args.hi, args.big, args.seed)
    cells = sum(p.ncells for p in probs)
    log(f"[bench] {len(probs)} instances, {cells / 1e6:.2f} M cells, inputs in "
        f"{time.perf_counter() - t0:.1f}s")

    precs = {"pc32": sg.SG_PC32, "mixed": sg.SG_MIXED, "fp64": sg.SG_FP64}
    prec = precs[args.precision]
    graph = args.graph == "on"
    tuned = None
    if args.autotune:
        import sayram_tune as tu
        best, bv, trials = tu.tune(probs, prec, steps=1, warmup=1, log=log)
        tuned = {"best": best, "value": bv, "trials": len(trials)}
        graph = best["use_graph"] == 1
    def tuned_opts(use_graph):
        if not tuned:
            return None
        o = tu.opts_for(prec, dict(tuned["best"], use_graph=use_graph))
        o.streams = args.streams
        return o

    def kernel_timing_run():
        # HIP-event timing of the dominant kernel records events from the host
        # around every dispatch: eager launches on one stream (same kernels,
        # same launch sizes as the headline run)
        dte, kt, _ = run_batched(probs, args.warmup, args.steps, prec, graph=False,
                                 time_kernel=True, opts=tuned_opts(-1))
        log(f"[bench] eager launches, one stream (kernel timing): {dte / args.steps * 1e3:.2f} ms/step")
        if kt[1] and kt[0] > 0:
            kms, kl, kbytes = kt
            log(f"[bench] {KERNEL_DESC}: {kl} dispatches, HIP-event mean {kms / kl * 1e3:.1f} us, "
                f"{kbytes / kl / 1e6:.1f} MB algorithmic per dispatch, "
                f"{kbytes / (kms * 1e-3) / 1e9:.0f} GB/s")
        return dte, kt

    if args.profile_only and not graph:
        kernel_timing_run()
        return
    dt, _, st = run_batched(probs, args.warmup, args.steps, prec, graph=graph, time_kernel=False,
                            opts=tuned_opts(1 if graph else -1), streams=args.streams)
    value = cells * args.steps / dt
    iters = [s["last_iters"] for s in st]
    refs = [s["last_refinements"] for s in st]
    log(f"[bench] {args.precision} batched: {dt / args.steps * 1e3:.2f} ms/step, {value:.3e} "
        f"cell-steps/s, Krylov iterations mean {np.mean(iters):.1f} max {max(iters)}, "
        f"refinements mean {np.mean(refs):.2f}")
    if args.profile_only:
        return

    out = {"metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": DTYPES[args.precision],
           "data": "synthetic: seeded smooth log-normal D tensor, G = sin cos p^2",
           "config": {"workload": f"batched, {PREC_NAMES[args.precision]}: {args.instances} "
                                  f"instances with dims uniform in {args.lo}..{args.hi} plus one "
                                  f"{args.big}^2 grid, dt = 1e-3 day",
                      "precision": args.precision,
                      "instances": len(probs), "cells": cells, "seed": args.seed},
           "launch": {"graph": graph, "streams": args.streams or 2, "autotuned": tuned},
           "solver": {"inner_iterations_mean": float(np.mean(iters)), "inner_iterations_max": int(max(iters)),
                      "refinements_mean": float(np.mean(refs)),
                      "min_f": float(min(s["min_f"] for s in st)),
                      "clamped_cells": int(sum(s["clamp_count"] for s in st)),
                      "max_rel_residual": float(max(s["last_rel_residual"] for s in st))}}
    dte, (kms, kl, kbytes) = kernel_timing_run()
    out["launch"]["eager_one_stream_ms_per_step"] = dte / args.steps * 1e3
    if kl and kms > 0:
        avg_ms = kms / kl
        achieved = kbytes / (kms * 1e-3) / 1e9
        out["roofline"] = {"kernel": KERNEL_DESC,
                           "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                           "avg_launch_ms": avg_ms, "launches": kl,
                           "timing": "HIP events around every dispatch on its stream, in an "
                                     "eager run on one stream and one instance group so that each "
                                     "interval is the kernel alone (the headline run overlaps tile "
                                     "shapes and instance groups)",
                           "algorithmic_bytes_per_launch": kbytes / kl}
        # HBM bytes per dispatch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
        # runs recorded beside the kernel-trace summary
        tf = os.path.join(ROOT, "profiles", "line_l0_traffic.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                t = json.load(fh)
            out["roofline"]["traffic"] = t["traffic_bytes_per_dispatch"]
            out["roofline"]["traffic_basis"] = (
                f"{t['basis']}; {t['traffic_bytes_per_dispatch'] / t['algorithmic_bytes_per_dispatch']:.2f}x "
                f"the algorithmic bytes there ({t['summary']})")

    if not args.skip_precisions:
        out["precisions"] = {}
        for name, pv in precs.items():
            if name == args.precision:
                out["precisions"][name] = {"value": value, "ms_per_step": dt / args.steps * 1e3,
                                           "krylov_iterations_mean": float(np.mean(iters))}
                continue
            d2, _, st2 = run_batched(probs, 1, 3, pv)
            out["precisions"][name] = {
                "value": cells * 3 / d2, "ms_per_step": d2 / 3 * 1e3,
                "krylov_iterations_mean": float(np.mean([s["last_iters"] for s in st2])),
                "refinements_mean": float(np.mean([s["last_refinements"] for s in st2]))}
            log(f"[bench] {name} batched: {d2 / 3 * 1e3:.2f} ms/step")

    if not args.skip_plain:
        pw, ps = 1, 1
        t_plain, f_plain = run_plain_fp64(probs, pw, ps)
        plain_value = cells * ps / t_plain
        log(f"[bench] plain FP64: {t_plain / ps * 1e3:.1f} ms/step, {plain_value:.3e} cell-steps/s")
        f_batched = batched_states(probs, pw + ps, prec)
        errs = [rel_l2(a, b) for a, b in zip(f_batched, f_plain)]
        out["plain_fp64"] = {"value": plain_value, "unit": "cell-steps/s",
                             "ms_per_step": t_plain / ps * 1e3,
                             "path": "same kernels, one instance at a time, FP64"}
        out["speedup_vs_plain_fp64"] = value / plain_value
        acc = {"steps_compared": pw + ps, "precision": args.precision,
               "max_rel_l2_batched_vs_plain_fp64": float(max(errs)),
               "min_f_batched": float(min(f.min() for f in f_batched))}
        if args.oracle_checks:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as orc
            small = sorted(range(len(probs)), key=lambda k: probs[k].ncells)[:args.oracle_checks]
            eo = []
            for k in small:
                ref = orc.run(orc.OracleProblem(probs[k]), probs[k].f0, pw + ps)
                eo.append((rel_l2(f_batched[k], ref), rel_l2(f_plain[k], ref)))
            acc["oracle_instances"] = [list(probs[k].f0.shape) for k in small]
            acc["max_rel_l2_batched_vs_oracle"] = float(max(e[0] for e in eo))
            acc["max_rel_l2_fp64_vs_oracle"] = float(max(e[1] for e in eo))
        out["accuracy"] = acc
        log(f"[bench] accuracy: {acc}")

    if args.serve_requests > 0:
        out["serving"] = serving(args.serve_requests, args.seed, prec, args.precision)
        log(f"[bench] serving: {out['serving']}")

    if args.cpu_budget > 0:
        v, n, c, t = cpu_baseline(probs, args.cpu_budget)
        out["cpu_baseline"] = {"value": v, "unit": "cell-steps/s", "cores": 1, "kind": "port",
                               "sample": f"first {n} instances of the batch (banded-LU oracle "
                                         f"small enough for single steps), 1 step each, "
                                         f"{c} cells, {t:.1f} s CPU",
                               "cpu": cpu_model(), "host_cpus": os.cpu_count()}
        log(f"[bench] cpu baseline: {v:.3e} cell-steps/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
