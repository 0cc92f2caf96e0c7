"""The benchmark workloads at their real sizes, as GPU tests.

  * the batched workloads: 512 instances with dims uniform in 32..512 (seed
    2026, per-instance D seed, amplitude 0.1..10x) plus one 4096 x 4096 grid,
    each instance one Solver::update (source/Solver.cc:270-290) per step, in
    FP64, in SG_PC32 (FP64 BiCGSTAB, FP32 V-cycle) and in SG_MIXED (FP32
    BiCGSTAB in FP64 refinement): every instance, every step: the step
    accepted, the true residual small, f >= 0, the FP32 forms within 1e-10 of
    FP64, and the smallest instances on the CPU oracle;
  * one 2048 x 2048 grid with tau = 1 day, 200 steps: FP64, PC32 and MIXED
    agree to 1e-10 every 10 steps, and the transposed problem gives the
    transposed solution (a size-independent property the oracle pins on
    small grids);
  * continuous-batching serving on a bounded trace drawn from the serving
    distribution (Poisson arrivals at 200 req/s, sides log-uniform 32..2048,
    steps uniform in 10..1000, capped here to fit the test time): each
    result equals its solo run bit for bit, and the small ones match the
    oracle.
"""
import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10          # relative L2 against the FP64 result / oracle, per step
# true relative residual of every accepted step: the solve stops at rtol 1e-14
# or where the residual stalls at the FP64 rounding floor (measured up to
# ~2e-11 on this batch); the per-step error budget is 1e-10
RES_TOL = 1e-10


def _oracle_cheap(p, steps):
    # banded LU costs N nx^2 per step: keep each check to a few seconds
    return p.ncells * p.nx * p.nx * steps <= 4e9


def test_batched_workload_every_instance_every_step(gpu_lib):
    sg = gpu_lib
    probs = sc.batch_problems(512, 32, 512, 4096, 2026)
    assert len(probs) == 513 and sum(p.ncells for p in probs) == 55002813
    precs = {"fp64": sg.SG_FP64, "pc32": sg.SG_PC32, "mixed": sg.SG_MIXED}
    batches = {}
    for name, pv in precs.items():
        b = sg.Batch(sg.make_opts(precision=pv))
        batches[name] = (b, [b.add(p) for p in probs])
        print(f"[batch] {name}: {len(probs)} instances resident, {b.bytes_in_use() / 2**30:.1f} GiB",
              flush=True)
    small = sorted(range(len(probs)), key=lambda k: probs[k].ncells)[:12]
    fo = {k: probs[k].f0.copy() for k in small}
    nsteps = 3
    worst = {"res": 0.0, "pc32": 0.0, "mixed": 0.0, "oracle": 0.0}
    for step in range(nsteps):
        for name, (b, _) in batches.items():
            rc = b.step_status(1)
            assert rc == 0, f"{name} step {step + 1}: {sg.lib().sg_last_error().decode()}"
        for k in small:
            fo[k] = orc.step(orc.OracleProblem(probs[k]), fo[k])
        for k in range(len(probs)):
            f = {}
            for name, (b, slots) in batches.items():
                st = b.stats(slots[k])
                assert st["status"] == 0 and st["steps"] == step + 1
                assert st["last_rel_residual"] <= RES_TOL, (name, k, st["last_rel_residual"])
                assert st["min_f"] >= 0.0 and st["last_nonfinite"] == 0
                worst["res"] = max(worst["res"], st["last_rel_residual"])
                f[name] = b.f(slots[k])
            assert f["fp64"].min() >= 0.0
            for name in ("pc32", "mixed"):
                e = rel_l2(f[name], f["fp64"])
                worst[name] = max(worst[name], e)
                assert e <= TOL, f"{name} vs fp64, instance {k} {probs[k].nx}x{probs[k].ny}, step {step + 1}: {e:.3e}"
            if k in fo:
                for name in precs:
                    e = rel_l2(f[name], fo[k])
                    worst["oracle"] = max(worst["oracle"], e)
                    assert e <= TOL, f"{name} vs oracle, instance {k}, step {step + 1}: {e:.3e}"
        print(f"[batch] step {step + 1}: worst true residual {worst['res']:.2e}, pc32 vs fp64 "
              f"{worst['pc32']:.2e}, mixed vs fp64 {worst['mixed']:.2e}, vs oracle "
              f"{worst['oracle']:.2e}", flush=True)
    for b, _ in batches.values():
        b.close()


def test_single_2048_grid_200_steps(gpu_lib):
    sg = gpu_lib
    p = sc.seeded_case(2048, 2048, seed=5, scale=1.0, inv_tau=1.0)
    sol = {name: sg.Solver(p, sg.make_opts(precision=pv))
           for name, pv in (("fp64", sg.SG_FP64), ("pc32", sg.SG_PC32), ("mixed", sg.SG_MIXED))}
    worst = 0.0
    for blk in range(20):
        for name, s in sol.items():
            for _ in range(10):
                s.update(1)
                st = s.stats()
                assert st["status"] == 0 and st["last_rel_residual"] <= RES_TOL, (name, st)
        f = {name: s.f() for name, s in sol.items()}
        for name in f:
            assert f[name].min() >= 0.0
        for name in ("pc32", "mixed"):
            e = rel_l2(f[name], f["fp64"])
            worst = max(worst, e)
            assert e <= TOL, f"{name} vs fp64 after {10 * (blk + 1)} steps: {e:.3e}"
        if blk % 5 == 4:
            print(f"[2048] {10 * (blk + 1)} steps: worst FP32-form vs FP64 {worst:.2e}", flush=True)
    for s in sol.values():
        s.close()


def test_2048_grid_transposition(gpu_lib):
    """Transposing every input transposes the solution (grid axes, G, f0,
    1/tau, D with Dxx <-> Dyy, sides XMIN <-> YMIN and XMAX <-> YMAX)."""
    sg = gpu_lib
    p = sc.seeded_case(2048, 1536, seed=15, scale=2.0, inv_tau=1.0)
    p.bc_type = [sc.ZEROFLUX] * 4
    q = sc.Problem(p.ny, p.nx, p.ye.copy(), p.xe.copy(), p.G.T.copy(), p.Dyy.T.copy(),
                   p.Dxy.T.copy(), p.Dxx.T.copy(), p.inv_tau.T.copy(), [sc.ZEROFLUX] * 4,
                   [np.zeros(p.nx + 1), np.zeros(p.nx + 1), np.zeros(p.ny + 1), np.zeros(p.ny + 1)],
                   p.f0.T.copy(), p.dt)
    a = sg.Solver(p, sg.make_opts(precision=sg.SG_FP64))
    b = sg.Solver(q, sg.make_opts(precision=sg.SG_FP64))
    a.update(10)
    b.update(10)
    assert rel_l2(b.f().T, a.f()) <= TOL
    a.close()
    b.close()


def test_serving_trace_from_the_serving_distribution(gpu_lib):
    import sayram_serve as sv
    sg = gpu_lib
    reqs = sv.poisson_trace(24, 200.0, 32, 2048, 10, 1000, seed=2031)
    for r in reqs:
        r.steps = min(r.steps, 24)   # the trace's sizes; steps capped for test time
    # a budget that holds a few of these requests at a time, so that most are
    # admitted as earlier ones retire and join a running batch
    budget = 640 << 20
    loop = sv.ServeLoop(hbm_budget=budget, precision=sg.SG_PC32)
    done, wall = loop.run(reqs, realtime=True)
    loop.close()
    assert len(done) == len(reqs) and all(r.error is None for r in done)
    first = min(r.admitted for r in done)
    assert sum(1 for r in done if r.admitted > first + 1e-9) > len(done) // 2
    assert max(max(r.problem.nx, r.problem.ny) for r in reqs) > 1024
    checked = 0
    for r in done:
        assert r.stats["steps"] == r.steps and r.result.min() >= 0.0
        solo = sg.Solver(r.problem, sg.make_opts(precision=sg.SG_PC32))
        solo.update(r.steps)
        assert np.array_equal(solo.f(), r.result), f"request {r.rid} differs from its solo run"
        solo.close()
        if _oracle_cheap(r.problem, r.steps):
            ref = orc.run(orc.OracleProblem(r.problem), r.problem.f0, r.steps)
            assert rel_l2(r.result, ref) <= TOL
            checked += 1
    assert checked >= 1
    s = sv.summarize(done, wall)
    print(f"[serve] {s}", flush=True)
    assert s["latency_p99_ms"] >= s["latency_p50_ms"] > 0
