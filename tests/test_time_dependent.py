"""Time-dependent Equations (Equation::update(t), source/Equation.h:57;
source/Solver.cc:286-289): after every solve the reference calls
eq.update(t), rebuilds Lambda = G*D (update_Lambda, :57-65) and the vertex
table with the Dirichlet values at the new t (:385-422), so step n uses the
inputs of t_n = n dt.

CPU: pins of the oracle's per-step schedule (orc.run_schedule). GPU: the
library refreshed through sg_set_coeffs / sg_set_bc_values, the batched
sg_batch_set_coeffs / sg_batch_set_bc_values, and the C++ GpuSolver (which
calls Equation::update and hands over what changed), every step against the
oracle at 1e-10 relative L2.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
from conftest import ROOT, rel_l2

TOL = 1e-10
DRIVER = os.path.join(ROOT, "sayram-2d_amd", "sayram_gpu_run")


def _base(nx=23, ny=19, seed=21):
    p = sc.seeded_case(nx, ny, seed=seed, scale=2.0)
    p.inv_tau = np.full((nx, ny), 0.5)
    return p


def _schedule(p, nsteps):
    return [sc.problem_at(p, n * p.dt) for n in range(nsteps)]


# --------------------------------------------------------------------------
# oracle pins (CPU)
# --------------------------------------------------------------------------
def test_schedule_decay_uses_start_of_step_loss():
    """D = 0 and 1/tau(t) = tau0 (1 + r t): f^{n+1} = f^n / (1 + dt/tau(t_n)).
    Taking 1/tau at t_{n+1} instead changes every step by ~r dt (4%)."""
    p = _base(6, 5)
    for a in ("Dxx", "Dxy", "Dyy"):
        setattr(p, a, np.zeros_like(getattr(p, a)))
    p.inv_tau = np.linspace(0.5, 4.0, p.nx * p.ny).reshape(p.nx, p.ny)
    q = sc.time_dependent(p, rtau=40.0)
    out = orc.run_schedule(_schedule(q, 5), p.f0)
    f = p.f0.copy()
    for n in range(5):
        f = f / (1 + p.dt * (p.inv_tau * (1.0 + 40.0 * (n * p.dt))))
        np.testing.assert_allclose(out[n], f, rtol=1e-13)
        wrong = out[n - 1] if n else p.f0
        wrong = wrong / (1 + p.dt * (p.inv_tau * (1.0 + 40.0 * ((n + 1) * p.dt))))
        assert rel_l2(out[n], wrong) > 1e-6


def test_schedule_piecewise_constant_equals_restart():
    """A schedule that switches inputs after 3 steps equals two plain runs
    chained, bit for bit."""
    a = _base(9, 8, seed=3)
    b = sc.problem_at(sc.time_dependent(a, rG=5.0, rD=30.0, rbc=10.0), 0.01)
    out = orc.run_schedule([a] * 3 + [b] * 2, a.f0)
    ref = orc.run(orc.OracleProblem(b), orc.run(orc.OracleProblem(a), a.f0, 3), 2)
    assert np.array_equal(out[-1], ref)
    assert np.array_equal(out[2], orc.run(orc.OracleProblem(a), a.f0, 3))


def test_schedule_dirichlet_values_at_start_of_step():
    """f = c everywhere with Dirichlet vertex values c on every side is a steady
    state (no loss). With values c (1 + r t), step 0 (t = 0) keeps f = c to
    rounding, and step 1 (t = dt) moves it: the boundary data of a step are
    those of its start time."""
    p = _base(8, 7, seed=4)
    p.inv_tau = np.zeros((p.nx, p.ny))
    c = 0.75
    p.bc_type = [sc.DIRICHLET] * 4
    p.bc_val = [np.full(p.ny + 1, c), np.full(p.ny + 1, c), np.full(p.nx + 1, c), np.full(p.nx + 1, c)]
    p.f0 = np.full((p.nx, p.ny), c)
    q = sc.time_dependent(p, rbc=200.0)
    out = orc.run_schedule(_schedule(q, 2), p.f0)
    assert np.max(np.abs(out[0] - c)) <= 1e-12 * c
    assert np.max(np.abs(out[1] - c)) > 1e-5 * c


def test_problem_at_is_the_base_problem_at_t0_and_linear_in_t():
    p = sc.time_dependent(_base(5, 4), rG=2.0, rD=-3.0, rtau=4.0, rbc=5.0)
    q0 = sc.problem_at(p, 0.0)
    for a in ("G", "Dxx", "Dxy", "Dyy", "inv_tau"):
        assert np.array_equal(getattr(q0, a), getattr(p, a))
    q = sc.problem_at(p, 0.1)
    np.testing.assert_allclose(q.G, p.G * 1.2, rtol=1e-15)
    np.testing.assert_allclose(q.Dxy, p.Dxy * 0.7, rtol=1e-15)
    np.testing.assert_allclose(q.bc_val[2], p.bc_val[2] * 1.5, rtol=1e-15)
    assert p.bc_type[1] == sc.ZEROFLUX and not np.any(q.bc_val[1])   # XMAX: zero flux


# --------------------------------------------------------------------------
# GPU
# --------------------------------------------------------------------------
def _refresh(solver_like, q, slot=None):
    args = dict(G=q.G, Dxx=q.Dxx, Dxy=q.Dxy, Dyy=q.Dyy, inv_tau=q.inv_tau)
    if slot is None:
        solver_like.set_coeffs(**args)
    else:
        solver_like.set_coeffs(slot, **args)
    for side in range(4):
        if q.bc_type[side] == sc.DIRICHLET:
            if slot is None:
                solver_like.set_bc_values(side, q.bc_val[side])
            else:
                solver_like.set_bc_values(slot, side, q.bc_val[side])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_solver_refresh_matches_oracle_every_step(gpu_lib, precision):
    p = sc.time_dependent(_base(), rG=3.0, rD=-20.0, rtau=40.0, rbc=25.0)
    n = 8
    sched = _schedule(p, n + 1)
    refs = orc.run_schedule(sched[:n], p.f0)
    s = gpu_lib.Solver(sched[0], gpu_lib.make_opts(precision=precision))
    for k in range(n):
        s.update(1)
        _refresh(s, sched[k + 1])      # Equation::update(t_{k+1}) after the step
        e = rel_l2(s.f(), refs[k])
        assert e <= TOL, f"step {k + 1}: {e:.3e}"
        assert s.stats()["status"] == 0 and s.f().min() >= 0.0
    # the refreshed inputs matter: a solver that keeps t = 0 inputs drifts away
    c = gpu_lib.Solver(sched[0], gpu_lib.make_opts(precision=precision))
    c.update(n)
    assert rel_l2(c.f(), refs[-1]) > 1e-6
    s.close()
    c.close()


@pytest.mark.gpu
def test_partial_refresh_equals_full_refresh(gpu_lib):
    """NULL arrays keep their device values: refreshing only D equals
    refreshing everything when only D changes."""
    p = sc.time_dependent(_base(31, 17, seed=5), rD=50.0)
    sched = _schedule(p, 5)
    a = gpu_lib.Solver(sched[0])
    b = gpu_lib.Solver(sched[0])
    for k in range(4):
        a.update(1)
        b.update(1)
        q = sched[k + 1]
        a.set_coeffs(G=q.G, Dxx=q.Dxx, Dxy=q.Dxy, Dyy=q.Dyy, inv_tau=q.inv_tau)
        b.set_coeffs(Dxx=q.Dxx, Dxy=q.Dxy, Dyy=q.Dyy)
    assert np.array_equal(a.f(), b.f())
    refs = orc.run_schedule(sched[:4], p.f0)
    assert rel_l2(a.f(), refs[-1]) <= TOL
    a.close()
    b.close()


@pytest.mark.gpu
def test_rejected_refresh_changes_nothing(gpu_lib):
    """Non-finite coefficients are rejected before anything is replaced, and
    the error paths of the refresh calls return SG_ERR_ARG."""
    sg = gpu_lib
    p = _base(20, 14, seed=6)
    a, b = sg.Solver(p), sg.Solver(p)
    bad = p.Dxy.copy()
    bad[4, 5] = np.inf
    with pytest.raises(sg.SayramError) as e:
        b.set_coeffs(Dxx=p.Dxx * 2.0, Dxy=bad)
    assert e.value.code == sg.SG_ERR_ARG and "non-finite" in str(e.value)
    with pytest.raises(sg.SayramError) as e:
        b.set_bc_values(1, np.zeros(p.ny + 1))   # XMAX is zero flux
    assert e.value.code == sg.SG_ERR_ARG
    v = p.bc_val[2].copy()
    v[0] = np.nan
    with pytest.raises(sg.SayramError) as e:
        b.set_bc_values(2, v)
    assert e.value.code == sg.SG_ERR_BC
    a.update(3)
    b.update(3)
    assert np.array_equal(a.f(), b.f())
    bt = sg.Batch()
    with pytest.raises(sg.SayramError) as e:
        bt.set_coeffs(5, G=p.G)                  # no such slot
    assert e.value.code == sg.SG_ERR_ARG
    for h in (a, b, bt):
        h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 2])
def test_batch_refresh_matches_oracle_every_step(gpu_lib, precision):
    """Two time-dependent instances with different rates and a constant one
    in one batch (sg_batch_set_coeffs, sg_batch_set_bc_values)."""
    ps = [sc.time_dependent(_base(23, 19, seed=7), rD=30.0, rbc=-8.0),
          sc.time_dependent(_base(40, 12, seed=8), rG=-4.0, rtau=25.0, rbc=12.0),
          _base(17, 33, seed=9)]
    n = 6
    scheds = [_schedule(p, n + 1) for p in ps]
    refs = [orc.run_schedule(s[:n], p.f0) for s, p in zip(scheds, ps)]
    b = gpu_lib.Batch(gpu_lib.make_opts(precision=precision))
    slots = [b.add(s[0]) for s in scheds]
    for k in range(n):
        b.step(1)
        for m in range(2):
            _refresh(b, scheds[m][k + 1], slot=slots[m])
        for m, sl in enumerate(slots):
            e = rel_l2(b.f(sl), refs[m][k])
            assert e <= TOL, f"instance {m}, step {k + 1}: {e:.3e}"
    b.close()


INI = """[basic]
run_id = td
nalpha0 = {nx}
nE = {ny}
alpha0min = 5
alpha0max = 90
Emin = 0.2
Emax = 5
T = 0.0105
nsteps = 11

[diagnostics]
nplots = 5

[diffusion_coefficients]
dID = none
"""


@pytest.mark.gpu
def test_driver_time_dependent_equation(tmp_path):
    """The C++ drop-in (GpuSolver through the driver) with an Equation whose G,
    D, 1/tau and Dirichlet values change in update(t): every snapshot and the
    final state against the oracle's schedule."""
    nx, ny = 24, 20
    (tmp_path / "p.ini").write_text(INI.format(nx=nx, ny=ny))
    base = sc.seeded_case(nx, ny, seed=78, scale=2.0, dt=0.0105 / 10)
    base.inv_tau = np.full((nx, ny), 0.25)
    p = sc.time_dependent(base, rG=2.0, rD=-15.0, rtau=30.0, rbc=20.0)
    sc.write_problem(p, str(tmp_path / "prob.bin"))
    r = subprocess.run([DRIVER, "p.ini", "prob.bin"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    refs = orc.run_schedule(_schedule(p, 10), p.f0)
    f = np.fromfile(tmp_path / "output" / "td" / "td_final.bin", dtype=np.float64).reshape(nx, ny)
    assert rel_l2(f, refs[-1]) <= TOL
    h5 = tmp_path / "output" / "td" / "td_data.h5"
    if h5.exists():
        from conftest import H5DUMP, read_h5
        if H5DUMP:
            for k in range(1, 6):
                assert rel_l2(read_h5(str(h5), f"/f/{k}"), refs[2 * k - 1]) <= TOL
