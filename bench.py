"""Benchmark of the gfx950 implicit PPFV step (Sayram-2D Solver::update).

Headline (default) workload: the mixed-precision batch -- 512 independent
instances with uneven grids (each dimension uniform in 32..512, per-instance D
seed and amplitude 0.1..10x) plus one 4096 x 4096 grid, advanced together by
the batched path. The default precision is SG_PC32: FP64 BiCGSTAB (operator,
vectors, residual, convergence test) preconditioned by an FP32 multigrid
V-cycle, with FP64 assembly and flux weights. --precision mixed measures the
FP32 Krylov solve inside FP64 iterative refinement, --precision fp64 the FP64
solve throughout. A step is one implicit time step of every instance; inputs
are resident in HBM before the timed region starts, and each step is one
graph launch per instance group (plus pipelined chunks when some instance
needs more iterations than the previous step).

Beside it, on the same batch:
  * the batched step in each of the three precisions (a shorter run each);
  * the plain FP64 path -- the same kernels, one instance at a time, FP64
    throughout -- timed for the speed-up;
  * accuracy over every step of the run: per instance and per step, the
    relative L2 error of the measured precision against the batched FP64
    result (bit-identical to the plain FP64 path), and of both against the
    CPU oracle on the smallest instances;
  * HIP-event timing of every kernel family over the timed steps (an eager
    run on one stream), with algorithmic bytes: the roofline of the dominant
    family, and a per-family breakdown;
  * continuous-batching serving on a bounded trace from the serving
    distribution (Poisson 200 req/s, sides log-uniform 32..2048, steps
    10..1000): p50/p99 latency with queueing, throughput, and every result
    against the FP64 result of the same request;
  * the CPU oracle (banded-LU restatement of Solver::update, one thread),
    timed in a separate process on one host core while the GPU work runs: the
    90 x 100 grid for its 1000 steps, and single steps of the batch's first
    instances (cpu_baseline).

Prints one JSON line on stdout (progress on stderr). --gpus above 1 is refused.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import subprocess
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
FAMILY_DESC = {
    "assemble": "k_assemble: operator and right-hand side, FP64",
    "rap": "set-up: k_pweights, k_rap (Galerkin products), k_line_factor",
    "line_l0": "k_line<T,5,*,false,*>: level-0 zebra line sweeps",
    "line_coarse": "k_line<T,9,*,*,*>: line sweeps of the per-level coarse levels (9-point)",
    "restrict": "k_restrict: restriction of the V-cycle right-hand sides",
    "prolong": "k_prolong: prolongation of the coarse corrections",
    "spmv_v": "k_bicg_v: v = M p-hat and (rh, v)",
    "spmv_t": "k_bicg_t: t = M s-hat, (t, s), (t, t)",
    "vec_update": "k_bicg_p, k_bicg_s, k_bicg_x: BiCGSTAB vector updates",
    "ir_residual": "k_ir_residual: FP64 residual of iterative refinement",
    "finalize": "k_finalize, k_commit: true residual, f of the accepted step",
    "coarse_cycle": "k_coarse_cycle: fused coarse V-cycle, one workgroup per instance",
    "finish": "k_finish: per-instance reduction finishes (latency-bound)",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def run_batched(probs, warmup, steps, precision, graph=True, opts=None, streams=0, timing=None):
    """Batched steps; returns (seconds, stats, batch kernel times). timing names
    kernel families to bracket with HIP events ("all" for every family): such
    a run launches eagerly, on one stream and in one instance group
    (streams=1, group_cells=-1), so that each dispatch's interval is that
    kernel alone."""
    if opts is None:
        opts = sg.make_opts(precision=precision, use_graph=1 if graph else -1, streams=streams)
    if timing:
        opts.use_graph, opts.streams, opts.group_cells = -1, 1, -1
    b = sg.Batch(opts)
    slots = [b.add(p) for p in probs]
    b.step(warmup)
    if timing:
        b.time_kernel(timing)
    t0 = time.perf_counter()
    b.step(steps)
    dt = time.perf_counter() - t0
    fam = {}
    if timing:
        for name in sg.KERNEL_IDS:
            fam[name] = b.kernel_time(name)
        fam["_bracket_ms"] = b.event_overhead_ms()
    st = [b.stats(s) for s in slots]
    b.close()
    return dt, st, fam


def run_plain_fp64(probs, warmup, steps):
    """The plain FP64 path: the same kernels, one instance at a time."""
    total = 0.0
    for k, p in enumerate(probs):
        s = sg.Solver(p, sg.make_opts(precision=sg.SG_FP64))
        s.update(warmup)
        t0 = time.perf_counter()
        s.update(steps)
        total += time.perf_counter() - t0
        s.close()
        if k % 128 == 127:
            log(f"[bench] plain FP64: {k + 1}/{len(probs)} instances")
    return total


def accuracy(probs, nsteps, precision, pname, n_oracle):
    """Per instance and per step: the measured precision against the batched
    FP64 result (bit-identical to the plain FP64 path, tests/test_gpu_parity),
    and both against the CPU oracle on the n_oracle smallest instances."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    bm = sg.Batch(sg.make_opts(precision=precision))
    b64 = sg.Batch(sg.make_opts(precision=sg.SG_FP64))
    sm = [bm.add(p) for p in probs]
    s64 = [b64.add(p) for p in probs]
    small = sorted(range(len(probs)), key=lambda k: probs[k].ncells)[:n_oracle]
    fo = {k: probs[k].f0.copy() for k in small}
    per_step = []
    worst_inst = (0.0, -1, -1)
    eo_m = eo_64 = 0.0
    minf = math.inf
    maxres = 0.0
    for n in range(nsteps):
        bm.step(1)
        b64.step(1)
        errs = np.empty(len(probs))
        for k in range(len(probs)):
            fm, f64 = bm.f(sm[k]), b64.f(s64[k])
            errs[k] = rel_l2(fm, f64)
            minf = min(minf, float(fm.min()))
            maxres = max(maxres, bm.stats(sm[k])["last_rel_residual"])
            if errs[k] > worst_inst[0]:
                worst_inst = (float(errs[k]), k, n + 1)
            if k in fo:
                fo[k] = orc.step(orc.OracleProblem(probs[k]), fo[k])
                eo_m = max(eo_m, rel_l2(fm, fo[k]))
                eo_64 = max(eo_64, rel_l2(f64, fo[k]))
        per_step.append({"step": n + 1, "max": float(errs.max()), "median": float(np.median(errs)),
                         "p99": float(np.percentile(errs, 99))})
    bm.close()
    b64.close()
    k = worst_inst[1]
    return {"steps_compared": nsteps, "precision": pname, "instances": len(probs),
            "basis": "relative L2 of every instance at every step of the run against the batched "
                     "FP64 result of the same instance and step",
            "max_rel_l2_vs_fp64": float(max(s["max"] for s in per_step)),
            "median_rel_l2_vs_fp64": float(np.median([s["median"] for s in per_step])),
            "worst_instance": {"index": k, "shape": [probs[k].nx, probs[k].ny],
                               "step": worst_inst[2], "rel_l2": worst_inst[0]},
            "per_step_vs_fp64": per_step,
            "min_f": minf, "max_true_rel_residual": maxres,
            "oracle_instances": [[probs[k].nx, probs[k].ny] for k in small],
            "max_rel_l2_vs_oracle": eo_m, "max_rel_l2_fp64_vs_oracle": eo_64}


def serving(n_req, seed, precision, pname, lo, hi, smin, smax, budget_gib):
    """Continuous batching under Poisson arrivals (200 req/s) on a bounded
    seeded trace from the serving distribution. The arrivals of the whole
    trace land within a fraction of a second, far faster than one GPU serves
    this distribution (it offers ~1e10 cell-steps/s), so the run is a
    closed-loop backlog and its completed-request rate is the measured
    sustainable rate for this mix. Every result is compared with the FP64
    result of the same request (an FP64 batch stepped to each request's step
    count; bit-identical to solving it alone)."""
    import sayram_serve as sv
    reqs = sv.poisson_trace(n_req, 200.0, lo, hi, smin, smax, seed=seed + 17)
    loop = sv.ServeLoop(hbm_budget=int(budget_gib * (1 << 30)), precision=precision)
    done, wall = loop.run(reqs, realtime=True)
    loop.close()
    s = sv.summarize(done, wall)
    ok = [r for r in done if r.error is None]
    # FP64 results of the same requests
    b = sg.Batch(sg.make_opts(precision=sg.SG_FP64))
    slots = [b.add(r.problem) for r in ok]
    targets = np.zeros(max(slots) + 1 if slots else 0, dtype=np.int64)
    for sl, r in zip(slots, ok):
        targets[sl] = r.steps
    t0 = time.perf_counter()
    b.step_until(targets, max((r.steps for r in ok), default=0))
    t64 = time.perf_counter() - t0
    errs = [rel_l2(r.result, b.f(sl)) for sl, r in zip(slots, ok)]
    b.close()
    cells = [r.problem.ncells for r in reqs]
    s.update({"trace": f"{n_req} requests, Poisson 200 req/s, sides log-uniform {lo}..{hi}, steps "
                       f"uniform {smin}..{smax}, seed {seed + 17}, precision {pname}, HBM budget "
                       f"{budget_gib} GiB",
              "trace_cells_min_max": [min(cells), max(cells)],
              "trace_cell_steps": int(sum(r.cell_steps for r in reqs)),
              "arrivals_span_s": float(max(r.arrival for r in reqs)),
              "sustainable_rate_req_per_s": s["throughput_req_per_s"],
              "sustainable_rate_basis": "measured: completed requests / wall time of this run, whose "
                                        "arrivals all land in the first fraction of a second "
                                        "(closed-loop backlog)",
              "max_rel_l2_vs_fp64": float(max(errs)) if errs else None,
              "median_rel_l2_vs_fp64": float(np.median(errs)) if errs else None,
              "fp64_same_requests_s": t64})
    return s


def cpu_worker(args):
    """Runs in its own process on one host core: the CPU oracle on the 90 x 100
    grid (full-tensor D, dt = 1e-3 day) for c1_steps steps timed whole, then
    single steps of the batch's first instances for about cpu_budget s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    out = {}
    p = sc.seeded_case(90, 100, seed=1, scale=1.0)
    op = orc.OracleProblem(p)
    t0, w0 = time.thread_time(), time.perf_counter()
    orc.run(op, p.f0, args.c1_steps)
    t1, w1 = time.thread_time(), time.perf_counter()
    out["c1"] = {"cells": p.ncells, "steps": args.c1_steps, "cpu_s": t1 - t0, "wall_s": w1 - w0}
    probs = sc.batch_problems(args.instances, args.lo, args.hi, 0, args.seed)
    cells, t_total, n = 0, 0.0, 0
    for q in probs:
        if q.ncells * q.nx * q.nx > 4e10:   # banded LU is N nx^2 work per step
            continue
        oq = orc.OracleProblem(q)
        a = time.thread_time()
        orc.step(oq, q.f0)
        t_total += time.thread_time() - a
        cells += q.ncells
        n += 1
        if t_total >= args.cpu_budget:
            break
    out["batch"] = {"instances": n, "cells": cells, "cpu_s": t_total}
    print(json.dumps(out), flush=True)


def start_cpu_baseline(args):
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", "--c1-steps", str(args.c1_steps),
           "--cpu-budget", str(args.cpu_budget), "--instances", str(args.instances), "--lo",
           str(args.lo), "--hi", str(args.hi), "--seed", str(args.seed)]
    env = dict(os.environ, OMP_NUM_THREADS="1")
    ncpu = os.cpu_count() or 1

    def pin():
        try:
            os.sched_setaffinity(0, {ncpu - 1})
        except (AttributeError, OSError):
            pass
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, env=env, preexec_fn=pin, text=True)


def family_table(fam, steps):
    out = {}
    for name, (ms, nl, by) in fam.items():
        if not nl:
            continue
        e = {"desc": FAMILY_DESC[name], "ms_per_step": ms / steps, "dispatches": nl,
             "avg_dispatch_us": ms / nl * 1e3}
        if by > 0:
            gbs = by / (ms * 1e-3) / 1e9
            e.update({"algorithmic_bytes_per_dispatch": by / nl, "achieved_gbs": gbs,
                      "frac": gbs / HBM_PEAK_GBS})
        out[name] = e
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--lo", type=int, default=32)
    ap.add_argument("--hi", type=int, default=512)
    ap.add_argument("--big", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="CPU seconds of single oracle steps on the batch's instances (0: no CPU "
                         "baseline)")
    ap.add_argument("--c1-steps", type=int, default=1000,
                    help="steps of the CPU oracle on the 90 x 100 grid, timed whole")
    ap.add_argument("--precision", choices=["pc32", "mixed", "fp64"], default="pc32",
                    help="precision of the batched step that is measured")
    ap.add_argument("--graph", choices=["on", "off"], default="on",
                    help="one captured hipGraph per step (per instance group)")
    ap.add_argument("--streams", type=int, choices=[0, 1, 2], default=0,
                    help="HIP streams of the line sweeps' tile shapes (0: library default, 1)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch parameters on the workload first (sayram_tune) and "
                         "time the step with the best set")
    ap.add_argument("--skip-precisions", action="store_true",
                    help="do not time the batched step in the other precisions")
    ap.add_argument("--oracle-checks", type=int, default=12)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-accuracy", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="requests in the serving trace (0 skips the serving measurement)")
    ap.add_argument("--serve-lo", type=int, default=32)
    ap.add_argument("--serve-hi", type=int, default=2048)
    ap.add_argument("--serve-steps", type=int, nargs=2, default=[10, 1000])
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--profile-only", action="store_true",
                    help="only the timed batched steps (for rocprofv3); with --graph off, only "
                         "the kernel-timing run (eager, one stream, HIP events per dispatch)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times: the known-size kernel that "
                         "scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--cpu-worker", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args)
        return
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

    cpu_proc = None
    if args.cpu_budget > 0 and not args.profile_only:
        cpu_proc = start_cpu_baseline(args)
    t0 = time.perf_counter()
    probs = sc.batch_problems(args.instances, args.lo, args.hi, args.big, args.seed)
    cells = sum(p.ncells for p in probs)
    log(f"[bench] {len(probs)} instances, {cells / 1e6:.2f} M cells, inputs in "
        f"{time.perf_counter() - t0:.1f}s")

    precs = {"pc32": sg.SG_PC32, "mixed": sg.SG_MIXED, "fp64": sg.SG_FP64}
    prec = precs[args.precision]
    graph = args.graph == "on"
    tuned = None
    if args.autotune:
        import sayram_tune as tu
        best, bv, trials = tu.tune(probs, prec, steps=2, warmup=3, log=log)
        tuned = {"best": best, "value": bv, "trials": len(trials)}
        graph = best["use_graph"] == 1

    def tuned_opts(use_graph):
        if not tuned:
            return None
        o = tu.opts_for(prec, dict(tuned["best"], use_graph=use_graph))
        o.streams = args.streams
        return o

    def kernel_timing_run():
        # HIP events around every dispatch of every family: eager launches on
        # one stream (same kernels, same launch sizes as the headline run)
        dte, _, fam = run_batched(probs, args.warmup, args.steps, prec, graph=False,
                                  opts=tuned_opts(-1), timing="all")
        bracket = fam.pop("_bracket_ms")
        log(f"[bench] eager launches, one stream (kernel timing): {dte / args.steps * 1e3:.2f} ms/step; "
            f"HIP-event interval around a one-workgroup kernel {bracket * 1e3:.1f} us")
        for name, (ms, nl, by) in sorted(fam.items(), key=lambda kv: -kv[1][0]):
            if nl:
                log(f"[bench]   {name:13s} {ms / args.steps:7.2f} ms/step, {nl:6d} dispatches, "
                    f"mean {ms / nl * 1e3:7.1f} us" +
                    (f", {by / (ms * 1e-3) / 1e9:6.0f} GB/s algorithmic" if by > 0 else ""))
        return dte, fam, bracket

    if args.profile_only and not graph:
        kernel_timing_run()
        return
    dt, st, _ = run_batched(probs, args.warmup, args.steps, prec, graph=graph,
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
           "launch": {"graph": graph, "streams": args.streams or 1, "autotuned": tuned,
                      "graph_launches_per_step": "one per instance group, plus pipelined chunks "
                                                 "when an instance needs more iterations than the "
                                                 "previous step's largest count"},
           "solver": {"krylov_iterations_mean": float(np.mean(iters)),
                      "krylov_iterations_max": int(max(iters)),
                      "refinements_mean": float(np.mean(refs)),
                      "min_f": float(min(s["min_f"] for s in st)),
                      "clamped_cells": int(sum(s["clamp_count"] for s in st)),
                      "rejected_steps": int(sum(s["status"] != 0 for s in st)),
                      "max_rel_residual": float(max(s["last_rel_residual"] for s in st))}}
    dte, fam, bracket = kernel_timing_run()
    out["launch"]["eager_one_stream_ms_per_step"] = dte / args.steps * 1e3
    table = family_table(fam, args.steps)
    if table:
        with_bytes = {k: v for k, v in table.items() if "achieved_gbs" in v}
        dom = max(with_bytes, key=lambda k: with_bytes[k]["ms_per_step"])
        d = with_bytes[dom]
        out["roofline"] = {"kernel": d["desc"], "family": dom,
                           "bound": "hbm", "achieved": d["achieved_gbs"], "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": d["frac"], "traffic": None,
                           "avg_launch_ms": d["avg_dispatch_us"] * 1e-3, "launches": d["dispatches"],
                           "algorithmic_bytes_per_launch": d["algorithmic_bytes_per_dispatch"],
                           "timing": "HIP events around every dispatch on its stream, over the "
                                     "timed steps of an eager run on one stream and one instance "
                                     "group so that each interval is the kernel alone (the "
                                     "headline run overlaps tile shapes and instance groups); the "
                                     "dominant family is the one with bytes and the most time",
                           "event_bracket_us": bracket * 1e3}
        tf = os.path.join(ROOT, "profiles", f"{dom}_traffic.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                t = json.load(fh)
            out["roofline"]["traffic"] = t["traffic_bytes_per_dispatch"]
            out["roofline"]["traffic_basis"] = t["basis"]
            if "rocprof" in t:
                out["roofline"]["rocprof"] = t["rocprof"]
        out["roofline_families"] = table

    if not args.skip_precisions:
        out["precisions"] = {}
        for name, pv in precs.items():
            if name == args.precision:
                out["precisions"][name] = {"value": value, "ms_per_step": dt / args.steps * 1e3,
                                           "krylov_iterations_mean": float(np.mean(iters))}
                continue
            # the same warm-up as the headline, then a shorter timed run
            n2 = min(args.steps, 10)
            d2, st2, _ = run_batched(probs, args.warmup, n2, pv)
            out["precisions"][name] = {
                "value": cells * n2 / d2, "ms_per_step": d2 / n2 * 1e3, "steps": n2,
                "krylov_iterations_mean": float(np.mean([s["last_iters"] for s in st2])),
                "refinements_mean": float(np.mean([s["last_refinements"] for s in st2]))}
            log(f"[bench] {name} batched: {d2 / n2 * 1e3:.2f} ms/step")

    if not args.skip_plain:
        # one step per instance after the headline's warm-up steps
        t_plain = run_plain_fp64(probs, args.warmup, 1)
        plain_value = cells / t_plain
        log(f"[bench] plain FP64: {t_plain * 1e3:.1f} ms/step, {plain_value:.3e} cell-steps/s")
        out["plain_fp64"] = {"value": plain_value, "unit": "cell-steps/s",
                             "ms_per_step": t_plain * 1e3,
                             "path": "same kernels, one instance at a time, FP64",
                             "sample": f"one step of every instance after {args.warmup} warm-up "
                                       "steps"}
        out["speedup_vs_plain_fp64"] = value / plain_value

    if not args.skip_accuracy:
        acc = accuracy(probs, args.warmup + args.steps, prec, args.precision, args.oracle_checks)
        out["accuracy"] = acc
        log(f"[bench] accuracy over {acc['steps_compared']} steps: max {acc['max_rel_l2_vs_fp64']:.2e}"
            f" median {acc['median_rel_l2_vs_fp64']:.2e} vs FP64, vs oracle "
            f"{acc['max_rel_l2_vs_oracle']:.2e}")

    if args.serve_requests > 0:
        out["serving"] = serving(args.serve_requests, args.seed, prec, args.precision,
                                 args.serve_lo, args.serve_hi, args.serve_steps[0],
                                 args.serve_steps[1], args.serve_budget_gib)
        log(f"[bench] serving: {out['serving']}")

    if cpu_proc is not None:
        res = json.loads(cpu_proc.communicate()[0].strip().splitlines()[-1])
        c1, bt = res["c1"], res["batch"]
        v = c1["cells"] * c1["steps"] / c1["cpu_s"]
        out["cpu_baseline"] = {
            "value": v, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"90 x 100 grid (full-tensor D, dt = 1e-3 day), {c1['steps']} steps timed "
                      f"whole: {c1['cpu_s']:.1f} s CPU on one pinned core, beside the GPU work",
            "batch_sample": {"value": bt["cells"] / bt["cpu_s"] if bt["cpu_s"] else None,
                             "unit": "cell-steps/s",
                             "sample": f"first {bt['instances']} instances of the batch, 1 step "
                                       f"each, {bt['cells']} cells, {bt['cpu_s']:.1f} s CPU"},
            "note": "the oracle's banded LU without pivoting costs N nx^2 per step, more than the "
                    "reference's SparseLU with COLAMD fill on large grids, so these are lower bounds "
                    "on the reference's CPU speed; the GPU/CPU ratio says nothing about kernel "
                    "quality (the roofline fraction does)",
            "cpu": cpu_model(), "host_cpus": os.cpu_count()}
        log(f"[bench] cpu baseline: {v:.3e} cell-steps/s (90 x 100, {c1['steps']} steps)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
