"""The serving loop's host logic on the CPU, with a stand-in batch that only
counts steps: a tick runs as many steps as the resident request nearest to
its end still needs, one step while arrivals are due and nothing waits for
budget, and every request ends with exactly its step count."""
import time
from types import SimpleNamespace

import numpy as np

import sayram_gpu as sg
import sayram_serve as sv


class CountingBatch:
    """sg.Batch's serving surface: a budget of `cap` instances, steps counted
    per slot, every step_until call recorded."""

    def __init__(self, cap, step_s=0.0):
        self.cap, self.step_s = cap, step_s
        self.slots = {}
        self.calls = []

    def try_add(self, prob):
        if len(self.slots) >= self.cap:
            return None
        s = next(k for k in range(self.cap) if k not in self.slots)
        self.slots[s] = 0
        return s

    def step_until(self, targets, nsteps=1):
        self.calls.append(int(nsteps))
        time.sleep(self.step_s * nsteps)
        for s in self.slots:
            t = int(targets[s]) if s < len(targets) else -1
            for _ in range(nsteps):
                if 0 <= t <= self.slots[s]:
                    break
                self.slots[s] += 1
        return sg.SG_OK

    def stats(self, s):
        return {"steps": self.slots[s], "status": sg.SG_OK, "last_rel_residual": 0.0}

    def f(self, s):
        return np.full(4, float(self.slots[s]))

    def remove(self, s):
        del self.slots[s]

    def close(self):
        pass


def make_loop(cap, too_big=(), step_s=0.0):
    loop = object.__new__(sv.ServeLoop)
    loop.precision, loop.budget = sg.SG_PC32, cap
    loop.batch = CountingBatch(cap, step_s)
    loop.ticks, loop.busy_s = 0, 0.0
    loop.fits_ever = lambda r: r.rid not in too_big
    return loop


def requests(steps, arrivals=None):
    arrivals = arrivals or [0.0] * len(steps)
    return [sv.Request(k, a, SimpleNamespace(nx=8, ny=8), n)
            for k, (n, a) in enumerate(zip(steps, arrivals))]


def test_a_tick_runs_to_the_next_retirement():
    loop = make_loop(cap=8)
    done, _ = loop.run(requests([3, 7, 5]), realtime=False)
    assert loop.batch.calls == [3, 2, 2]
    assert [r.rid for r in done] == [0, 2, 1]
    assert all(r.stats["steps"] == r.steps and r.result[0] == r.steps for r in done)


def test_admission_waits_for_budget_and_ticks_follow_it():
    loop = make_loop(cap=1)
    done, _ = loop.run(requests([4, 2, 3]), realtime=False)
    assert loop.batch.calls == [4, 2, 3]
    assert [r.rid for r in done] == [0, 1, 2]
    assert all(r.stats["steps"] == r.steps for r in done)


def test_single_steps_while_arrivals_are_due():
    # the second request arrives while the first runs (2 ms a step) and the
    # budget has room: the loop steps one at a time until it arrives, so that
    # it is admitted at the next step; then ticks run to each retirement
    loop = make_loop(cap=4, step_s=0.002)
    done, _ = loop.run(requests([50, 3], arrivals=[0.0, 0.05]), realtime=True)
    first, second = sorted(done, key=lambda r: r.rid)
    calls = loop.batch.calls
    assert first.stats["steps"] == 50 and second.stats["steps"] == 3
    assert 0.05 <= second.admitted < first.finished     # they overlapped
    assert sum(calls) == 50                              # every tick stepped the first
    n1 = calls.index(3)                                  # the tick that ran the second
    assert all(c == 1 for c in calls[:n1]) and calls[n1 + 1:] == [50 - n1 - 3]


def test_a_request_that_never_fits_is_rejected_alone():
    loop = make_loop(cap=2, too_big={1})
    done, _ = loop.run(requests([2, 5, 3]), realtime=False)
    rej = [r for r in done if getattr(r, "error", None)]
    assert [r.rid for r in rej] == [1] and rej[0].result is None
    ok = [r for r in done if not getattr(r, "error", None)]
    assert sorted(r.rid for r in ok) == [0, 2] and all(r.stats["steps"] == r.steps for r in ok)
