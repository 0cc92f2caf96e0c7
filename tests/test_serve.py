"""Continuous batching: requests admitted by HBM budget while others run, each
result identical to the same problem solved alone."""
import numpy as np
import pytest

import sayram_gpu as sg
import sayram_serve as sv


def test_poisson_trace_is_seeded_and_bounded():
    a = sv.poisson_trace(20, 50.0, 16, 128, 3, 9, seed=4)
    b = sv.poisson_trace(20, 50.0, 16, 128, 3, 9, seed=4)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(16 <= r.problem.nx <= 128 and 16 <= r.problem.ny <= 128 for r in a)
    assert all(3 <= r.steps <= 9 for r in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_continuous_batching_alone_equals_mixed(gpu_lib, precision):
    reqs = sv.poisson_trace(14, 400.0, 16, 96, 2, 9, seed=11)
    # a budget that holds only a few requests at once, so admission waits for
    # retirements and requests join a running batch
    per = max(sg.Batch.instance_bytes(r.problem.nx, r.problem.ny, precision) for r in reqs)
    loop = sv.ServeLoop(hbm_budget=3 * per, precision=precision)
    done, wall = loop.run(reqs, realtime=True)
    loop.close()
    assert len(done) == len(reqs)
    joined_late = sum(1 for r in done if r.admitted > min(x.admitted for x in done) + 1e-9)
    assert joined_late > 0
    for r in done:
        assert r.stats["steps"] == r.steps and r.result.min() >= 0.0
        solo = sg.Solver(r.problem, sg.make_opts(precision=precision))
        solo.update(r.steps)
        assert np.array_equal(solo.f(), r.result), f"request {r.rid} differs from its solo run"
        solo.close()
    s = sv.summarize(done, wall)
    assert s["latency_p99_ms"] >= s["latency_p50_ms"] > 0
