"""Autotuner of the batched implicit step's launch parameters, by measurement.

On a caller's workload (a list of problems: the production mix, or a sample of
it) the tuner times the batched step for candidate values of the parameters
that select how the step is launched, and returns the fastest set:

* ``check_every``: Krylov iterations between host convergence checks. A larger
  value issues fewer host synchronisations; a smaller one runs fewer batched
  iterations after the last instance has converged.
* ``use_graph``: Krylov chunks replayed from captured hipGraphs (1) or launched
  eagerly (-1).
* ``line_tiles``: the line lengths at which the line solves switch tile shape
  (rows: short / medium limits; columns: short / medium limits), and
  ``coarse_tiles`` the same for the coarser levels.
* ``streams``: the tile shapes of a sweep on two HIP streams (2) or one (1).
* ``group_cells``: the size from which instances are stepped as a second
  group, concurrently with the rest (-1: one group).
* ``fuse_cells``: the size below which an instance's coarse levels run in the
  fused one-workgroup cycle (-1: never; above 1024 only with an FP32 V-cycle).

``check_every``, ``use_graph``, ``streams`` and ``group_cells`` never change results: every
instance's result is bit-identical for all their values. ``line_tiles``,
``coarse_tiles`` and ``fuse_cells`` change the rounding of the line solves (not
their accuracy), so the tuned values are part of the options a run is
reproduced with. A candidate the library refuses for the precision (an
``SG_ERR_ARG``) is skipped.

Each trial builds a fresh batch, so no state is carried between trials. It runs
``warmup`` steps, then times ``steps`` steps. The search is coordinate descent
from the library defaults, with every measured point cached.
"""
from __future__ import annotations

import time

import sayram_gpu as sg

DEFAULTS = {"check_every": 1, "use_graph": 1, "row_short": 16, "row_mid": 160,
            "col_short": 16, "col_mid": 768, "crow_mid": 160, "ccol_mid": 768, "streams": 1,
            "fuse_cells": 1024, "group_cells": 1 << 21}

CANDIDATES = {
    "check_every": [1, 2, 3, 4],
    "use_graph": [1, -1],
    "row_mid": [96, 160, 256],
    "col_mid": [384, 768, 1536],
    "row_short": [16, 24, 32],
    "col_short": [16, 24, 32],
    "crow_mid": [96, 160, 256],
    "ccol_mid": [128, 256, 768],
    "streams": [1, 2],
    "fuse_cells": [-1, 256, 1024, 2048, 4096],
    "group_cells": [-1, 1 << 18, 1 << 21],
}


def opts_for(precision: int, p: dict):
    return sg.make_opts(precision=precision, check_every=p["check_every"], use_graph=p["use_graph"],
                        line_tiles=(p["row_short"], p["row_mid"], p["col_short"], p["col_mid"]),
                        coarse_tiles=(p["row_short"], p.get("crow_mid", 160), p["col_short"],
                                      p.get("ccol_mid", 768)),
                        streams=p.get("streams", 1), fuse_cells=p.get("fuse_cells", 1024),
                        group_cells=p.get("group_cells", 1 << 21))


def measure(problems, precision: int, p: dict, steps: int = 2, warmup: int = 1) -> float:
    """Cell-steps per second of the batched step with parameters p."""
    b = sg.Batch(opts_for(precision, p))
    try:
        for q in problems:
            b.add(q)
        b.step(warmup)
        t0 = time.perf_counter()
        b.step(steps)
        dt = time.perf_counter() - t0
    finally:
        b.close()
    return sum(q.ncells for q in problems) * steps / dt


def tune(problems, precision: int = sg.SG_PC32, steps: int = 2, warmup: int = 1,
         order=("check_every", "use_graph", "streams", "group_cells", "fuse_cells", "col_mid",
                "row_mid", "col_short", "row_short", "ccol_mid", "crow_mid"),
         log=None):
    """Coordinate descent from DEFAULTS. Returns (best parameters, best value,
    list of (parameters, cell-steps/s) for every measured point)."""
    cache = {}

    def value(p):
        key = tuple(sorted(p.items()))
        if key not in cache:
            try:
                cache[key] = measure(problems, precision, p, steps, warmup)
            except sg.SayramError as e:
                if e.code != sg.SG_ERR_ARG:
                    raise
                cache[key] = 0.0          # not valid for this precision
            if log:
                log(f"[tune] {p} -> {cache[key]:.4e} cell-steps/s")
        return cache[key]

    best = dict(DEFAULTS)
    best_v = value(best)
    for name in order:
        for cand in CANDIDATES[name]:
            p = dict(best, **{name: cand})
            v = value(p)
            if v > best_v:
                best, best_v = p, v
    trials = [(dict(k), v) for k, v in cache.items()]
    return best, best_v, trials
