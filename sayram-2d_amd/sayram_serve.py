"""Continuous-batching serving loop for implicit PPFV solve requests.

A request is one Sayram-2D problem (grid, D tensor, boundary data, f0) plus a
number of implicit steps. The loop keeps every admitted request resident on the
GPU -- its f, Lambda, operator hierarchy and Krylov state live in its own
device slab -- and advances all resident requests together, one implicit step
per tick, with the batched kernels (sg_batch_step_until). Requests that arrive
while others run are admitted in arrival order as soon as the HBM budget has
room (sg_instance_bytes tells the cost before anything is allocated); a request
that reaches its step count is retired at the end of that tick (its f is
copied to the host and its slab freed), which frees budget for the next ones.
A tick is as many steps as the resident request nearest to its end still
needs, so that it ends when that request does; while arrivals are still due
and the budget could take them, a tick is one step, so that each one is
admitted at the next step.

Because every reduction is per instance and in a fixed order, a request's
result does not depend on what else is resident: it is bit-identical to the
same problem solved alone.

Latency is measured from arrival to the result on the host.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

import sayram_cases as sc
import sayram_gpu as sg


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    problem: sc.Problem
    steps: int
    admitted: float = math.nan
    finished: float = math.nan
    done_steps: int = 0     # steps taken so far (after the last tick)
    result: np.ndarray | None = None
    stats: dict = field(default_factory=dict)
    error: str | None = None   # why the request was rejected or stopped

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def cell_steps(self) -> int:
        return self.problem.ncells * self.steps


def poisson_trace(n: int, rate: float, lo: int, hi: int, smin: int, smax: int, seed: int,
                  scale_range=(0.1, 10.0)):
    """Seeded trace: exponential inter-arrival times (rate req/s), grid sizes
    log-uniform in [lo, hi] per dimension, steps uniform in [smin, smax]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += float(rng.exponential(1.0 / rate))
        nx = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        ny = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        steps = int(rng.integers(smin, smax + 1))
        s = float(math.exp(rng.uniform(math.log(scale_range[0]), math.log(scale_range[1]))))
        p = sc.seeded_case(nx, ny, seed=int(seed * 7919 + k), scale=s)
        out.append(Request(k, t, p, steps))
    return out


class ServeLoop:
    def __init__(self, hbm_budget: int, precision: int = sg.SG_MIXED, opts=None):
        self.precision = precision
        self.budget = int(hbm_budget)
        self.batch = sg.Batch(opts or sg.make_opts(precision=precision), self.budget)
        self.ticks = 0
        self.busy_s = 0.0

    def fits_ever(self, req: Request) -> bool:
        return sg.Batch.instance_bytes(req.problem.nx, req.problem.ny, self.precision) <= self.budget

    def run(self, requests, realtime: bool = True, log=None):
        """Serves the requests; with realtime=False arrivals are released as
        fast as the loop runs (a closed-loop backlog). Returns every request,
        in completion order: a request that can never fit the HBM budget is
        rejected on arrival, and one whose step the library rejects (the
        solve did not converge) is stopped there; both carry `error`."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue = []                       # arrived, not admitted (FIFO)
        resident = {}                    # slot -> Request
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        i = 0
        done = []
        while i < len(pending) or queue or resident:
            tnow = now()
            while i < len(pending) and (not realtime or pending[i].arrival <= tnow):
                r = pending[i]
                i += 1
                if self.fits_ever(r):
                    queue.append(r)
                else:
                    r.error = (f"{r.problem.nx}x{r.problem.ny} needs more than the HBM budget "
                               f"of {self.budget} bytes")
                    r.finished = now()
                    done.append(r)
                    if log:
                        log(f"[serve] request {r.rid} rejected: {r.error}")
            # admit in arrival order while the budget has room
            while queue:
                slot = self.batch.try_add(queue[0].problem)
                if slot is None:
                    break
                r = queue.pop(0)
                r.admitted = now()
                resident[slot] = r
            if not resident:
                if i < len(pending):     # idle until the next arrival
                    time.sleep(max(0.0, pending[i].arrival - now()))
                continue
            # one tick: every resident request advances, none past its target;
            # k steps, k the fewest any resident still needs (one step while
            # arrivals are due and nothing waits for budget: a new arrival
            # is then admitted at the next step)
            nslot = max(resident) + 1
            targets = np.full(nslot, -1, dtype=np.int64)
            for s, r in resident.items():
                targets[s] = r.steps
            k = min(r.steps - r.done_steps for r in resident.values())
            if i < len(pending) and not queue:
                k = 1
            tb = time.perf_counter()
            rc = self.batch.step_until(targets, max(1, k))
            self.busy_s += time.perf_counter() - tb
            if rc not in (sg.SG_OK, sg.SG_ERR_NOCONV):
                raise sg.SayramError(rc, sg.lib().sg_last_error().decode())
            self.ticks += 1
            for s in list(resident):
                st = self.batch.stats(s)
                resident[s].done_steps = st["steps"]
                failed = st["status"] == sg.SG_ERR_NOCONV
                if st["steps"] >= resident[s].steps or failed:
                    r = resident.pop(s)
                    r.result = self.batch.f(s)
                    r.stats = st
                    if failed:
                        r.error = (f"step {st['steps'] + 1} rejected: the solve did not converge "
                                   f"(true residual {st['last_rel_residual']:.3e})")
                    self.batch.remove(s)
                    r.finished = now()
                    done.append(r)
                    if log:
                        log(f"[serve] request {r.rid} {r.problem.nx}x{r.problem.ny} x{r.steps} "
                            f"{'failed' if failed else 'done'}, latency {r.latency * 1e3:.1f} ms")
        wall = now()
        return done, wall

    def close(self):
        self.batch.close()


def summarize(done, wall):
    """Latency percentiles (arrival to result on the host, queueing included)
    and throughput over the requests that completed; rejected or failed ones
    are counted apart."""
    ok = [r for r in done if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.array([math.nan])
    cs = sum(r.cell_steps for r in ok)
    return {"requests": len(done), "completed": len(ok), "failed": len(done) - len(ok),
            "wall_s": wall,
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "latency_max_ms": float(lat.max() * 1e3),
            "throughput_cell_steps_per_s": cs / wall,
            "throughput_req_per_s": len(ok) / wall,
            "min_f": float(min(r.result.min() for r in ok)) if ok else math.nan}
