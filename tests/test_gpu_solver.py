"""GPU behaviour beyond per-step parity: golden vectors, long trajectories,
full-size properties (where the CPU oracle cannot follow), the error paths of
the C ABI, resume, determinism and the batch lifecycle."""
import glob
import time
import os

import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
from conftest import rel_l2
from test_oracle import load_golden

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(g)[:-4] for g in GOLDEN])
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_gpu_matches_golden(gpu_lib, path, precision):
    p, z = load_golden(path)
    s = gpu_lib.Solver(p, gpu_lib.make_opts(precision=precision))
    bands, R = s.debug_assemble()
    scale = np.max(np.abs(z["bands0"]), axis=0)
    assert np.all(np.abs(bands - z["bands0"]) <= 1e-11 * scale[None, :])
    for k in range(z["f_steps"].shape[0]):
        s.update(1)
        assert rel_l2(s.f(), z["f_steps"][k]) <= 1e-12, f"step {k + 1}"
    s.close()


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_long_trajectory_c1_grid(gpu_lib, precision):
    """The reference's CPU-sized case: 90 (alpha0 5..90 deg) x 100 (ln E,
    0.2..5 MeV), full tensor D, dt = 1e-3 day; 300 steps against the oracle,
    checked every 10 steps at 1e-10 relative L2."""
    p = sc.seeded_case(90, 100, seed=1, scale=1.0)
    op = orc.OracleProblem(p)
    s = gpu_lib.Solver(p, gpu_lib.make_opts(precision=precision))
    f = p.f0.copy()
    worst = 0.0
    for blk in range(30):
        f = orc.run(op, f, 10)
        s.update(10)
        g = s.f()
        worst = max(worst, rel_l2(g, f))
        assert worst <= 1e-10, f"after {10 * (blk + 1)} steps: {worst:.3e}"
        assert g.min() >= 0.0
    s.close()


def test_c2_size_properties(gpu_lib):
    """2048 x 2048, tau = 1 day: true residual at the FP64 floor, f >= 0, and
    the mixed and PC32 paths within 1e-10 of the FP64 path after 10 steps."""
    p = sc.seeded_case(2048, 2048, seed=5, scale=1.0, inv_tau=1.0)
    sol = [gpu_lib.Solver(p, gpu_lib.make_opts(precision=prec)) for prec in (0, 1, 2)]
    for _ in range(10):
        for s in sol:
            s.update(1)
            assert s.stats()["last_rel_residual"] <= 1e-12
    fa, fb, fc = (s.f() for s in sol)
    assert fa.min() >= 0.0 and fb.min() >= 0.0 and fc.min() >= 0.0
    assert rel_l2(fb, fa) <= 1e-10
    assert rel_l2(fc, fa) <= 1e-10
    for s in sol:
        s.close()


def test_conservation_at_scale(gpu_lib):
    """Zero-flux sides, no loss: sum G dx dy f is conserved at 1024^2."""
    p = sc.seeded_case(1024, 1024, seed=6, scale=3.0)
    p.bc_type = [sc.ZEROFLUX] * 4
    w = p.G * np.outer(np.diff(p.xe), np.diff(p.ye))
    m0 = np.sum(w * p.f0)
    for prec in (0, 1, 2):
        s = gpu_lib.Solver(p, gpu_lib.make_opts(precision=prec))
        s.update(5)
        assert abs(np.sum(w * s.f()) - m0) <= 1e-11 * m0
        s.close()


@pytest.mark.parametrize("warm", [-1, 0], ids=["cold", "warm"])
def test_resume_and_determinism(gpu_lib, warm):
    """Resume = upload f and istep. With the solve started at f (warm_start
    -1) the resumed run continues bit for bit; with the default warm start the
    first resumed step has no increment to extrapolate, so it starts elsewhere
    and agrees to the solve tolerance. Either way a run is bit-reproducible."""
    p = sc.seeded_case(70, 45, seed=7, scale=2.0)
    opts = lambda: gpu_lib.make_opts(warm_start=warm)
    a = gpu_lib.Solver(p, opts())
    a.update(6)
    b = gpu_lib.Solver(p, opts())
    b.update(3)
    f3 = b.f()
    b.close()
    c = gpu_lib.Solver(p, opts())
    c.set_f(f3, 3)
    c.update(3)
    if warm < 0:
        assert np.array_equal(c.f(), a.f())
    else:
        assert rel_l2(c.f(), a.f()) <= 1e-12
    assert abs(c.t() - 6 * p.dt) < 1e-15
    d = gpu_lib.Solver(p, opts())
    d.update(6)
    assert np.array_equal(d.f(), a.f())   # bit-reproducible
    for s in (a, c, d):
        s.close()


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_warm_start_same_solution_fewer_iterations(gpu_lib, precision):
    """Starting each solve at f + (last increment) changes the start, not the
    system: both trajectories stay on the oracle, and on smooth trajectories
    the warm start needs fewer iterations in total over a batch. (One
    instance alone may not show it: with loose inner solves, `SG_MIXED`
    counts whole refinements' worth of iterations, and the 96 x 80 case
    below takes 53 warm against 48 cold while the batch saves 7%.)"""
    sg = gpu_lib
    shapes = [(96, 80), (120, 64), (72, 140), (150, 110), (64, 64), (130, 90), (48, 160),
              (110, 120)]
    probs = [sc.seeded_case(nx, ny, seed=21 + k, scale=s) for k, ((nx, ny), s) in
             enumerate(zip(shapes, (1.0, 0.5, 2.0, 4.0, 0.3, 1.5, 6.0, 0.8)))]
    refs = [orc.run(orc.OracleProblem(p), p.f0, 12) for p in probs]
    its = {}
    for warm in (-1, 0):
        b = sg.Batch(sg.make_opts(precision=precision, warm_start=warm))
        slots = [b.add(p) for p in probs]
        b.step(4)
        t0 = [b.stats(s)["total_iters"] for s in slots]
        b.step(8)
        its[warm] = sum(b.stats(s)["total_iters"] - a for s, a in zip(slots, t0))
        for k, (s, ref) in enumerate(zip(slots, refs)):
            assert rel_l2(b.f(s), ref) <= 1e-10, f"{shapes[k]} warm {warm}"
        b.close()
    assert its[0] < its[-1], f"warm {its[0]} against cold {its[-1]} iterations"


def test_error_paths(gpu_lib):
    sg = gpu_lib
    p = sc.seeded_case(20, 10, seed=8)
    q = sc.seeded_case(20, 10, seed=8)
    q.xe = q.xe.copy()
    q.xe[5] = q.xe[4]              # Grid2D::validate (source/Grid2D.h:44-67)
    with pytest.raises(sg.SayramError) as e:
        sg.Solver(q)
    assert e.value.code == sg.SG_ERR_ARG and "strictly increasing" in str(e.value)
    q = sc.seeded_case(20, 10, seed=8)
    q.Dxx = q.Dxx.copy()
    q.Dxx[3, 3] = np.nan
    with pytest.raises(sg.SayramError) as e:
        sg.Solver(q)
    assert e.value.code == sg.SG_ERR_ARG
    q = sc.seeded_case(20, 10, seed=8)
    q.bc_val[2] = None              # Dirichlet side without values (source/Solver.cc:392-394)
    with pytest.raises(sg.SayramError) as e:
        sg.Solver(q)
    assert e.value.code == sg.SG_ERR_BC
    # a step the solve cannot finish is rejected: the status reaches the
    # caller, f and the step count stay as they were, and the next call with a
    # reachable tolerance proceeds from there
    s = sg.Solver(p, sg.make_opts(max_iters=1, rtol=1e-16))
    rc = s.step_status(1)
    st = s.stats()
    assert rc == sg.SG_ERR_NOCONV and st["status"] == sg.SG_ERR_NOCONV
    assert "rejected" in sg.lib().sg_last_error().decode()
    assert st["steps"] == 0 and s.t() == 0.0 and st["last_nonfinite"] == 0
    assert np.array_equal(s.f(), p.f0)
    with pytest.raises(sg.SayramError) as e:
        s.update(1)
    assert e.value.code == sg.SG_ERR_NOCONV
    s.close()
    with pytest.raises(sg.SayramError) as e:
        sg.Batch(sg.make_opts(precision=3))
    assert e.value.code == sg.SG_ERR_ARG
    b = sg.Batch(sg.make_opts(), hbm_budget=sg.Batch.instance_bytes(20, 10) + 1024)
    assert b.add(p) == 0
    assert b.try_add(p) is None          # over the HBM budget: nothing allocated
    b.close()


def test_batch_lifecycle_and_step_until(gpu_lib):
    sg = gpu_lib
    probs = [sc.seeded_case(30 + 7 * k, 20 + 5 * k, seed=40 + k, scale=1.0 + k) for k in range(4)]
    b = sg.Batch()
    slots = [b.add(p) for p in probs]
    b.remove(slots[1])
    assert b.size() == 3
    s_new = b.add(probs[1])           # reuses the free slot
    assert s_new == slots[1]
    targets = np.array([2, 5, 3, 1], dtype=np.int64)
    assert b.step_until(targets, 6) == 0
    for sl, t, p in zip(slots, targets, probs):
        assert b.stats(sl)["steps"] == t
        ref = sg.Solver(p)
        ref.update(int(t))
        assert np.array_equal(b.f(sl), ref.f())
        ref.close()
    b.close()


def test_clamp_option(gpu_lib):
    """f = max(x, 0) by default, f = x with no_clamp (the reference's copy-back,
    source/Solver.cc:279-284): one step from the same state solves the same
    system, so the two differ exactly by the clamp, and the clamp count is the
    number of negative entries of x."""
    sg = gpu_lib
    p = sc.seeded_case(48, 40, seed=12, scale=4.0)
    a = sg.Solver(p)
    b = sg.Solver(p, sg.make_opts(no_clamp=1))
    a.update(1)
    b.update(1)
    fb = b.f()
    assert np.array_equal(a.f(), np.maximum(fb, 0.0))
    sa, sb = a.stats(), b.stats()
    assert sa["last_clamps"] == int(np.sum(fb < 0)) and sb["last_clamps"] == 0
    assert sa["last_min_x"] == sb["last_min_x"] == fb.min()
    assert sa["min_f"] == max(fb.min(), 0.0)
    ref = orc.run(orc.OracleProblem(p), p.f0, 1)
    assert rel_l2(fb, ref) <= 1e-10
    a.close()
    b.close()


def test_kernel_timing_changes_no_result(gpu_lib):
    """HIP-event timing of every kernel family brackets each dispatch of an
    eager run: results equal the untimed run bit for bit, every family that
    ran reports dispatches, the timed intervals fit within the wall time, and
    the bracket overhead is a positive interval."""
    sg = gpu_lib
    probs = [sc.seeded_case(nx, ny, seed=70 + k) for k, (nx, ny) in
             enumerate([(40, 30), (130, 90), (64, 200)])]
    out = []
    for timed in (False, True):
        b = sg.Batch(sg.make_opts(precision=sg.SG_PC32, use_graph=-1))
        slots = [b.add(p) for p in probs]
        b.step(1)
        if timed:
            b.time_kernel("all")
        t0 = time.perf_counter()
        b.step(2)
        wall_ms = (time.perf_counter() - t0) * 1e3
        out.append([b.f(s) for s in slots])
        if timed:
            fam = {k: b.kernel_time(k) for k in sg.KERNEL_IDS}
            for k in ("assemble", "rap", "line_l0", "spmv_v", "spmv_t", "vec_update",
                      "finalize"):
                assert fam[k][1] > 0 and fam[k][0] > 0, k
            assert fam["assemble"][1] == 2           # one dispatch per step
            assert fam["ir_residual"][1] == 0        # PC32 does no refinement
            tot = b.kernel_time()
            assert abs(tot[0] - sum(v[0] for v in fam.values())) <= 1e-9 * tot[0] + 1e-12
            assert tot[1] == sum(v[1] for v in fam.values())
            assert tot[0] <= wall_ms
            ov = b.event_overhead_ms()
            assert 0.0 < ov < 1.0
        b.close()
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)


def test_batch_resume_one_slot(gpu_lib):
    """sg_batch_set_f resumes one instance of a batch from a saved state and
    step count: it continues bit-identically to the uninterrupted run, and the
    other instances are untouched. A state of the wrong shape is refused."""
    sg = gpu_lib
    probs = [sc.seeded_case(nx, ny, seed=80 + k, scale=s) for k, (nx, ny, s) in
             enumerate([(50, 35, 1.0), (24, 60, 3.0), (41, 41, 0.5)])]
    cold = sg.make_opts(warm_start=-1)     # solves start at f: resumes are bitwise
    ref = sg.Batch(cold)
    rs = [ref.add(p) for p in probs]
    ref.step(2)
    saved = ref.f(rs[1])
    ref.step(3)
    b = sg.Batch(cold)
    bs = [b.add(p) for p in probs]
    b.step(5)
    b.set_f(bs[1], probs[1].f0, 0)      # restart slot 1 from f0 ...
    b.step_until([5, 2, 5], 5)          # ... and run it back to step 2
    assert b.stats(bs[1])["steps"] == 2
    assert np.array_equal(b.f(bs[1]), saved)
    b.set_f(bs[1], saved, 2)            # resume from the saved state
    b.step_until([5, 5, 5], 5)
    for r, s in zip(rs, bs):
        assert np.array_equal(b.f(s), ref.f(r))
        assert b.stats(s)["steps"] == 5
    with pytest.raises(ValueError):
        b.set_f(bs[0], np.zeros((3, 3)), 0)
    with pytest.raises(sg.SayramError):
        b.set_f(99, np.zeros((3, 3)), 0)
    b.close()
    ref.close()


def test_fp64_vcycle_as_strong_as_fp32(gpu_lib):
    """The FP64 and FP32 V-cycles run the same schedule on the same hierarchy,
    one in higher precision: per instance and step the FP64 solve needs at
    most one iteration more than the PC32 one (with every coarse line swept
    they agree to about that). A sweep that skipped some line groups of a
    level, in one precision only, would still converge, only in more
    iterations: this is what shows it. The shapes cover the 9-point levels'
    rows and columns in every tile shape both precisions use (lines of at
    most 16 cells, medium and long)."""
    sg = gpu_lib
    shapes = [(200, 150), (300, 90), (96, 400), (128, 128), (520, 40), (36, 610), (900, 700)]
    probs = [sc.seeded_case(nx, ny, seed=170 + k, scale=s)
             for k, ((nx, ny), s) in enumerate(zip(shapes, (0.5, 2.0, 4.0, 1.0, 6.0, 0.3, 1.5)))]
    its = {}
    for prec in (sg.SG_PC32, sg.SG_FP64):
        b = sg.Batch(sg.make_opts(precision=prec))
        slots = [b.add(p) for p in probs]
        its[prec] = []
        for _ in range(4):
            b.step(1)
            its[prec].append([b.stats(s)["last_iters"] for s in slots])
        b.close()
    for step, (a, c) in enumerate(zip(its[sg.SG_PC32], its[sg.SG_FP64])):
        for k, (i32, i64) in enumerate(zip(a, c)):
            assert i64 <= i32 + 1, f"{shapes[k]} step {step + 1}: FP64 {i64} vs PC32 {i32} iterations"
