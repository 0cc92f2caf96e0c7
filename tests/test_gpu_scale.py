"""Size-independent properties at full size, extreme aspect ratios and
nonuniform grids in the batched path.

  * D = 0 makes every flux vanish (source/Solver.cc:99-164 with Lambda = 0), so
    the step is the loss alone, f^{n+1} = f^n / (1 + dt/tau) cell by cell
    (source/Solver.cc:190-199): a closed form checked on a 4096 x 4096 grid
    stepped in a batch beside ordinary instances.
  * Lines of 512 and 1024 cells across 3 (rows, columns), batched with the full-size
    grid, against the oracle every step.
  * Geometrically stretched edges in both directions (the reference's Mesh
    takes any increasing edges, source/Mesh.cc:36-65), batched, against the
    oracle every step.
"""
import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10


def stretched(p, rx, ry):
    """The problem on edges whose widths grow geometrically by rx (x) and ry
    (y) from cell to cell, over the same span; the cell arrays are kept."""
    def edges(e, r):
        n = len(e) - 1
        w = r ** np.arange(n)
        return e[0] + (e[-1] - e[0]) * np.concatenate([[0.0], np.cumsum(w) / w.sum()])
    p.xe = np.ascontiguousarray(edges(p.xe, rx))
    p.ye = np.ascontiguousarray(edges(p.ye, ry))
    if p.bc_type[2] == sc.DIRICHLET:
        p.bc_val[2] = np.sin(p.xe) * p.bc_val[2].max()
    if p.bc_type[3] == sc.DIRICHLET:
        p.bc_val[3] = np.sin(p.xe) * p.bc_val[3].max()
    return p


def test_loss_only_closed_form_at_full_size_and_thin_strips(gpu_lib):
    sg = gpu_lib
    big = sc.seeded_case(4096, 4096, seed=41, scale=1.0)
    for a in ("Dxx", "Dxy", "Dyy"):
        setattr(big, a, np.zeros_like(getattr(big, a)))
    rng = np.random.default_rng(5)
    big.inv_tau = np.ascontiguousarray(rng.uniform(0.0, 50.0, (4096, 4096)))
    strips = [sc.seeded_case(512, 3, seed=42, scale=2.0), sc.seeded_case(3, 1024, seed=43, scale=0.5)]
    b = sg.Batch(sg.make_opts(precision=sg.SG_PC32))
    sb = b.add(big)
    ss = [b.add(p) for p in strips]
    refs = [p.f0.copy() for p in strips]
    f = big.f0.copy()
    for step in range(3):
        assert b.step_status(1) == 0
        f = f / (1.0 + big.dt * big.inv_tau)
        g = b.f(sb)
        assert rel_l2(g, f) <= 1e-13, f"4096^2 loss-only step {step + 1}: {rel_l2(g, f):.3e}"
        for k, (p, sl) in enumerate(zip(strips, ss)):
            refs[k] = orc.step(orc.OracleProblem(p), refs[k])
            e = rel_l2(b.f(sl), refs[k])
            assert e <= TOL, f"{p.nx}x{p.ny} step {step + 1}: {e:.3e}"
    b.close()


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_stretched_grids_batched_against_oracle(gpu_lib, precision):
    sg = gpu_lib
    probs = [stretched(sc.seeded_case(60, 44, seed=51, scale=2.0), 1.03, 0.97),
             stretched(sc.seeded_case(37, 90, seed=52, scale=0.3), 0.95, 1.04),
             stretched(sc.seeded_case(128, 20, seed=53, scale=6.0, inv_tau=2.0), 1.01, 1.1)]
    probs[1].bc_type = [sc.ZEROFLUX] * 4
    b = sg.Batch(sg.make_opts(precision=precision))
    slots = [b.add(p) for p in probs]
    refs = [p.f0.copy() for p in probs]
    for step in range(4):
        b.step(1)
        for k, (p, sl) in enumerate(zip(probs, slots)):
            refs[k] = orc.step(orc.OracleProblem(p), refs[k])
            e = rel_l2(b.f(sl), refs[k])
            assert e <= TOL, f"instance {k} step {step + 1}: {e:.3e}"
            assert b.f(sl).min() >= 0.0
    b.close()


def test_every_coarse_tile_shape_in_one_batch_against_oracle(gpu_lib):
    """The line groups of a coarse level's colour share one launch across
    tile shapes, and which shapes occur depends on line length and V-cycle
    precision. These grids give the first coarse level (more than 1024
    cells, so swept by the per-level launches) rows of 200 and 250 (4 x 256
    tiles), 20 and 60 (16 x 64) and 15 cells (64 x 16), and columns of 12
    cells (64 x 16), of 20, 125 and 300 cells (32 x 32 tiles in FP32, 16 x 64
    in FP64) and of 1000 cells (1024-thread 16 x 256 tiles in FP32). Batched
    together, in every precision, every instance stays on the oracle every
    step."""
    sg = gpu_lib
    shapes = [(400, 40), (40, 600), (30, 2000), (120, 250), (500, 24)]
    probs = [sc.seeded_case(nx, ny, seed=61 + k, scale=s)
             for k, ((nx, ny), s) in enumerate(zip(shapes, (2.0, 0.5, 1.0, 4.0, 1.5)))]
    refs = [[p.f0.copy()] for p in probs]
    for k, p in enumerate(probs):
        for _ in range(3):
            refs[k].append(orc.step(orc.OracleProblem(p), refs[k][-1]))
    for precision in (0, 1, 2):
        b = sg.Batch(sg.make_opts(precision=precision))
        slots = [b.add(p) for p in probs]
        for step in range(3):
            assert b.step_status(1) == 0
            for k, s in enumerate(slots):
                e = rel_l2(b.f(s), refs[k][step + 1])
                assert e <= TOL, f"precision {precision} {shapes[k]} step {step + 1}: {e:.3e}"
        b.close()
