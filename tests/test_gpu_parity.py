"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Assembly is compared bit for bit (same operations, same order, no fused
multiply-add on either side); the solution is compared per step at the FP64
tolerance 1e-10 relative L2 over whole trajectories, for the plain FP64 path,
the mixed-precision path and the batched path.
"""
import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10   # relative L2 per step against the FP64 oracle

SHAPES = [(1, 1), (1, 7), (7, 1), (2, 2), (5, 3), (16, 12), (33, 17), (40, 64)]


def variant(p, kind):
    """Boundary/loss variants of a seeded problem."""
    if kind == "default":      # XMIN Dirichlet 0, XMAX zero flux, YMIN/YMAX Dirichlet
        return p
    if kind == "all_dirichlet":
        p.bc_type = [sc.DIRICHLET] * 4
        p.bc_val[1] = np.sin(p.xe[-1]) * np.exp(-(np.exp(p.ye) - 0.2) / 0.1)
        return p
    if kind == "all_zeroflux":
        p.bc_type = [sc.ZEROFLUX] * 4
        return p
    if kind == "lc_loss":      # loss-cone variant: XMIN zero flux, 1/tau > 0 near the loss cone
        p.bc_type = [sc.ZEROFLUX, sc.ZEROFLUX, sc.DIRICHLET, sc.DIRICHLET]
        it = np.zeros((p.nx, p.ny))
        it[: max(1, p.nx // 4), :] = 3.0
        p.inv_tau = it
        return p
    raise ValueError(kind)


KINDS = ["default", "all_dirichlet", "all_zeroflux", "lc_loss"]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", KINDS)
def test_assembly_bitwise(gpu_lib, shape, kind):
    p = variant(sc.seeded_case(*shape, seed=11, scale=3.0), kind)
    s = gpu_lib.Solver(p)
    bands, R = s.debug_assemble()
    s.close()
    ob, oR = orc.assemble(orc.OracleProblem(p), p.f0)
    for k in range(5):
        assert np.array_equal(bands[k], ob[k]), (
            f"band {k}: max rel diff {np.max(np.abs(bands[k]-ob[k]))/max(np.max(np.abs(ob[k])),1e-300):.3e}")
    assert np.array_equal(R, oR), f"R: max diff {np.max(np.abs(R-oR)):.3e}"


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_steps_vs_oracle(gpu_lib, shape, kind, precision):
    p = variant(sc.seeded_case(*shape, seed=7, scale=5.0), kind)
    op = orc.OracleProblem(p)
    s = gpu_lib.Solver(p, gpu_lib.make_opts(precision=precision))
    f_ref = p.f0.copy()
    worst = 0.0
    for n in range(8):
        f_ref = orc.step(op, f_ref)
        s.update(1)
        f = s.f()
        e = rel_l2(f, f_ref)
        worst = max(worst, e)
        assert e <= TOL, f"step {n+1}: rel L2 {e:.3e}"
        assert f.min() >= 0.0
    st = s.stats()
    assert st["steps"] == 8 and st["status"] == 0
    assert abs(s.t() - 8 * p.dt) < 1e-15
    s.close()


def test_batch_equals_solo(gpu_lib):
    probs = [sc.seeded_case(nx, ny, seed=30 + k, scale=s) for k, (nx, ny, s) in
             enumerate([(37, 21, 0.3), (8, 90, 4.0), (64, 64, 1.0), (3, 5, 9.0)])]
    for prec in (0, 1, 2):
        b = gpu_lib.Batch(gpu_lib.make_opts(precision=prec))
        slots = [b.add(p) for p in probs]
        b.step(3)
        for p, sl in zip(probs, slots):
            s = gpu_lib.Solver(p, gpu_lib.make_opts(precision=prec))
            s.update(3)
            assert np.array_equal(b.f(sl), s.f()), "batched result differs from the solo run"
            s.close()
        b.close()


def test_graph_replay_equals_eager(gpu_lib):
    """Krylov chunks replayed from captured hipGraphs give bit-identical states
    to eager launches, step after step (each step replays the graphs many
    times, with the host's convergence checks and the step's other kernels
    between the replays), in every precision, also after the batch changes."""
    probs = [sc.seeded_case(nx, ny, seed=60 + k, scale=s) for k, (nx, ny, s) in
             enumerate([(45, 31, 0.5), (12, 70, 3.0), (64, 48, 1.0)])]
    for prec in (0, 1, 2):
        bg = gpu_lib.Batch(gpu_lib.make_opts(precision=prec, use_graph=1))
        be = gpu_lib.Batch(gpu_lib.make_opts(precision=prec, use_graph=-1))
        sg_ = [bg.add(p) for p in probs[:2]]
        se_ = [be.add(p) for p in probs[:2]]
        for step in range(6):
            if step == 3:          # a new instance: maps rebuilt, graphs recaptured
                sg_.append(bg.add(probs[2]))
                se_.append(be.add(probs[2]))
            bg.step(1)
            be.step(1)
            for a, b in zip(sg_, se_):
                assert np.array_equal(bg.f(a), be.f(b)), f"precision {prec}, step {step}"
        bg.close()
        be.close()


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_fused_coarse_cycle_matches_per_level_cycle(gpu_lib, precision):
    """The fused coarse cycle (one launch per instance from its first level of
    at most fuse_cells cells) solves the same coarse problems as the per-level
    launches: trajectories agree to rounding, and both stay on the oracle.
    Shapes cover fusion from level 0 (small grids), from a coarse level, and
    a batch mixing the two."""
    probs = [sc.seeded_case(nx, ny, seed=90 + k, scale=s) for k, (nx, ny, s) in
             enumerate([(24, 20, 2.0), (70, 45, 0.5), (33, 130, 5.0)])]
    out = {}
    # FP32 V-cycles also keep levels of up to 2048 and 4096 cells in LDS
    # (other LDS layouts of the fused kernel)
    fcs = (0, -1, 300) if precision == 0 else (0, -1, 300, 2048, 4096)
    for fc in fcs:
        b = gpu_lib.Batch(gpu_lib.make_opts(precision=precision, fuse_cells=fc))
        slots = [b.add(p) for p in probs]
        b.step(4)
        out[fc] = [b.f(s) for s in slots]
        b.close()
    for k, p in enumerate(probs):
        ref = orc.run(orc.OracleProblem(p), p.f0, 4)
        for fc in fcs:
            if fc != -1:
                assert rel_l2(out[fc][k], out[-1][k]) <= 1e-11
        for fc in out:
            assert rel_l2(out[fc][k], ref) <= TOL


@pytest.mark.parametrize("csm", [4, 8])
def test_coarse_smoother_schedules_stay_on_the_oracle(gpu_lib, csm):
    """Coarse levels smoothed with x-lines only or y-lines only: a different
    preconditioner, the same solution to the solve tolerance."""
    probs = [sc.seeded_case(nx, ny, seed=95 + k, scale=s) for k, (nx, ny, s) in
             enumerate([(70, 45, 0.5), (33, 130, 2.0), (120, 96, 1.0)])]
    b = gpu_lib.Batch(gpu_lib.make_opts(precision=gpu_lib.SG_PC32, coarse_smoother=csm))
    slots = [b.add(p) for p in probs]
    b.step(3)
    for p, sl in zip(probs, slots):
        assert rel_l2(b.f(sl), orc.run(orc.OracleProblem(p), p.f0, 3)) <= TOL
    b.close()


@pytest.mark.parametrize("precision", ["pc32", "fp64"])
@pytest.mark.parametrize("graph", [1, -1], ids=["graph", "eager"])
def test_two_streams_equal_one_stream_bitwise(gpu_lib, graph, precision):
    """With streams = 2 the tile shapes of a sweep run on two streams, and the
    per-step line factorisations run on the second stream beside the chain of
    Galerkin products; they write disjoint data, so every result equals the
    one-stream schedule (the default) bit for bit. Line lengths 12, 100, 300 and 900 give
    sweeps with three tile shapes in both directions (FP64 V-cycles use other
    shapes than FP32 ones)."""
    prec = {"pc32": gpu_lib.SG_PC32, "fp64": gpu_lib.SG_FP64}[precision]
    probs = [sc.seeded_case(nx, ny, seed=120 + k) for k, (nx, ny) in
             enumerate([(12, 300), (100, 12), (300, 100), (900, 40), (60, 900)])]
    out = {}
    for ns in (1, 2):
        b = gpu_lib.Batch(gpu_lib.make_opts(precision=prec, streams=ns, use_graph=graph))
        slots = [b.add(p) for p in probs]
        b.step(3)
        out[ns] = [b.f(s) for s in slots]
        b.close()
    for k in range(len(probs)):
        assert np.array_equal(out[1][k], out[2][k])


def test_two_groups_equal_one_group_bitwise(gpu_lib):
    """A batch steps its instances of at least group_cells cells as a second
    group, concurrently with the rest. Grouping changes no result: slots,
    per-instance targets, removal and re-use of a slot and the statistics
    behave as with one group, and every state is equal bit for bit."""
    shapes = [(40, 30), (120, 90), (64, 64), (200, 150), (33, 17)]
    probs = [sc.seeded_case(nx, ny, seed=140 + k) for k, (nx, ny) in enumerate(shapes)]
    extra = sc.seeded_case(150, 130, seed=150)
    out = {}
    for gc in (-1, 5000):
        b = gpu_lib.Batch(gpu_lib.make_opts(precision=gpu_lib.SG_PC32, group_cells=gc))
        slots = [b.add(p) for p in probs]
        b.step(2)
        b.remove(slots[1])
        s_new = b.add(extra)
        assert s_new == slots[1]
        targets = [4, 3, 5, 3, 4]
        b.step_until(targets, 3)
        res = [b.f(s) for s in slots]
        st = [b.stats(s) for s in slots]
        out[gc] = (res, [(x["steps"], x["last_iters"], x["total_iters"]) for x in st])
        b.close()
    for k in range(len(probs)):
        assert np.array_equal(out[-1][0][k], out[5000][0][k])
    assert out[-1][1] == out[5000][1]
    assert [x[0] for x in out[5000][1]] == [4, 3, 5, 3, 4]


@pytest.mark.parametrize("smoother", [0, 12, 13, 15, 5])
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_smoother_schedules_stay_on_the_oracle(gpu_lib, smoother, precision):
    """The default (0: post-smoothing alone on the finest level, an x-line
    pre-sweep before it on the coarser ones), and on every level:
    post-smoothing alone (12), an x-line pre-sweep before it (13: the fused
    cycle then pre-smooths too, and only the even rows' residual is
    restricted), pre-sweeps in both directions (15: every residual formed) and
    x-lines only (5): different preconditioners, the same solution to the
    solve tolerance, in every precision."""
    probs = [sc.seeded_case(nx, ny, seed=160 + k, scale=s) for k, (nx, ny, s) in
             enumerate([(70, 45, 0.5), (33, 130, 2.0), (120, 96, 1.0), (9, 300, 3.0)])]
    b = gpu_lib.Batch(gpu_lib.make_opts(precision=precision, smoother=smoother))
    slots = [b.add(p) for p in probs]
    b.step(3)
    for p, sl in zip(probs, slots):
        assert rel_l2(b.f(sl), orc.run(orc.OracleProblem(p), p.f0, 3)) <= TOL
        assert b.stats(sl)["status"] == 0
    b.close()
