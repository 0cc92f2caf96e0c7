"""The reference's Albert & Young (2005) chorus cases (source/Cases/
Albert_Young.{h,cc} and Albert_Young_LC.{h,cc}) through the driver.

* The case inputs the driver builds (host/albert_young.h, from the HDF5 table
  D/AlbertYoung_chorus.h5) equal the Python restatement
  (sayram_cases.albert_young_case) bit for bit.
* The interpolation rules (locate / calWeight / interp2D) and the closed forms
  of G and the loss-cone rate are pinned independently.
* The p.ini runs (80 x 80, T = 1 day, 500 steps, 10 snapshots), the
  reference's own configuration, stay within 1e-10 of the oracle at every
  snapshot on the GPU.

tests/golden/AlbertYoung_chorus.h5 is the reference's diffusion table
D/AlbertYoung_chorus.h5 (data, 91 alpha0 x 49 E), copied unchanged.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
from conftest import ROOT, read_h5, rel_l2

DRIVER = os.path.join(ROOT, "sayram-2d_amd", "sayram_gpu_run")
TABLE = os.path.join(ROOT, "tests", "golden", "AlbertYoung_chorus.h5")

# the values of the reference's p.ini and p_AlbertYoungLC.ini
INI = """[basic]
run_id = {run}
nalpha0 = {nx}
nE = {ny}
alpha0min = {a0min}
alpha0max = 90
Emin = 0.2
Emax = 5
T = 1.0
nsteps = {nsteps}

[diagnostics]
nplots = {nplots}

[diffusion_coefficients]
dID = AlbertYoung_chorus
"""


def table():
    return {k: read_h5(TABLE, "/" + k) for k in ("alpha0", "E", "Daa", "Dap", "Dpp")}


def setup(tmp_path, lc, nx=80, ny=80, nsteps=500, nplots=10):
    run = "AlbertYoungLC" if lc else "AlbertYoung"
    (tmp_path / "p.ini").write_text(INI.format(run=run, nx=nx, ny=ny, a0min=0 if lc else 5,
                                               nsteps=nsteps, nplots=nplots))
    (tmp_path / "D").mkdir()
    shutil.copy(TABLE, tmp_path / "D" / "AlbertYoung_chorus.h5")
    dt = 1.0 / ((nsteps // nplots) * nplots)
    return run, sc.albert_young_case(nx, ny, 0 if lc else 5, 90, 0.2, 5, dt, table(), lc=lc)


def read_problem(path, nx, ny):
    raw = open(path, "rb").read()
    assert raw[:7] == b"SGPROB1"
    off = 8
    n = np.frombuffer(raw, np.int32, 2, off); off += 8
    assert tuple(n) == (nx, ny)
    arrs = []
    for _ in range(6):
        arrs.append(np.frombuffer(raw, np.float64, nx * ny, off).reshape(nx, ny)); off += 8 * nx * ny
    bt = np.frombuffer(raw, np.int32, 4, off); off += 16
    vals = []
    for s in range(4):
        k = ny + 1 if s < 2 else nx + 1
        vals.append(np.frombuffer(raw, np.float64, k, off)); off += 8 * k
    return arrs, list(bt), vals


def driver_with_hdf5():
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    r = subprocess.run(["ldd", DRIVER], capture_output=True, text=True)
    if "libhdf5" not in r.stdout:
        pytest.skip("driver built without HDF5")


@pytest.mark.parametrize("lc", [False, True], ids=["albert_young", "albert_young_lc"])
def test_driver_inputs_equal_python_restatement(tmp_path, lc):
    driver_with_hdf5()
    _, p = setup(tmp_path, lc)
    args = [DRIVER, "--dump-case", "case.bin"] + (["--case", "albert_young_lc"] if lc else []) + ["p.ini"]
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    arrs, bt, vals = read_problem(tmp_path / "case.bin", p.nx, p.ny)
    for name, a, b in zip(("G", "Dxx", "Dxy", "Dyy", "inv_tau", "f0"), arrs,
                          (p.G, p.Dxx, p.Dxy, p.Dyy, p.inv_tau, p.f0)):
        assert np.array_equal(a, b), f"{name}: max rel diff {np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))}"
    assert bt == list(p.bc_type)
    for s in range(4):
        want = p.bc_val[s] if p.bc_val[s] is not None else np.zeros_like(vals[s])
        assert np.array_equal(vals[s], want), f"side {s}"


def test_interpolation_rules():
    # nodes are reproduced exactly
    raw = [[float(7 * i + 3 * j) * 0.1 + 1.0 for j in range(6)] for i in range(5)]
    for i in range(4):
        for j in range(5):
            i0, wi = sc._cal_weight(float(i), 4)
            j0, wj = sc._cal_weight(float(j), 5)
            assert (i0, wi, j0, wj) == (i, 1.0, j, 1.0)
            assert sc._interp2d(raw, i0, j0, wi, wj) == raw[i][j]
    # a linear table is reproduced inside each cell (bilinear exactness)
    lin = lambda x, y: 0.25 + 1.5 * x - 0.75 * y
    raw = [[lin(i, j) for j in range(6)] for i in range(5)]
    for px, py in ((0.3, 0.7), (2.5, 4.9), (3.99, 0.01)):
        i0, wi = sc._cal_weight(px, 4)
        j0, wj = sc._cal_weight(py, 5)
        assert abs(sc._interp2d(raw, i0, j0, wi, wj) - lin(px, py)) <= 1e-14
    # at or past the last node: clamped to the last interval, weight 0
    assert sc._cal_weight(4.0, 4) == (3, 0.0)
    assert sc._cal_weight(7.2, 4) == (3, 0.0)
    # below the table: the reference's size_t index wraps out of range and is
    # clamped the same way (the last row is used)
    assert sc._cal_weight(-0.3, 4) == (3, 0.0)


def test_shipped_grids_lie_inside_the_table():
    t = table()
    tx = t["alpha0"] * sc.GPI / 180
    ty = t["E"]
    for a0min in (5, 0):
        p = sc.albert_young_case(80, 80, a0min, 90, 0.2, 5, 2e-3, t)
        x = 0.5 * (p.xe[:-1] + p.xe[1:])
        y = 0.5 * (p.ye[:-1] + p.ye[1:])
        px = (x - tx[0]) / (tx[-1] - tx[0]) * (len(tx) - 1)
        py = (y - math.log(ty[0])) / (math.log(ty[-1]) - math.log(ty[0])) * (len(ty) - 1)
        assert px.min() >= 0 and px.max() < len(tx) - 1
        assert py.min() >= 0 and py.max() < len(ty) - 1


def test_closed_forms_of_G_f0_and_loss_cone():
    t = table()
    p = sc.albert_young_case(8, 6, 5, 90, 0.2, 5, 2e-3, t)
    q = sc.albert_young_case(40, 6, 0, 90, 0.2, 5, 2e-3, t, lc=True)
    E0 = sc.E0
    for prob, a0min in ((p, 5), (q, 0)):
        x = 0.5 * (prob.xe[:-1] + prob.xe[1:])
        E = np.exp(0.5 * (prob.ye[:-1] + prob.ye[1:]))
        X, EE = np.meshgrid(x, E, indexing="ij")
        mom = np.sqrt(EE * (EE + 2 * E0))
        # G = p^2 T(a) sin cos / (dlnE/dp), dlnE/dp = p / (E (E + E0))
        G = mom * (1.30 - 0.56 * np.sin(X)) * np.sin(X) * np.cos(X) * EE * (EE + E0)
        assert np.max(np.abs(prob.G - G) / np.abs(G)) <= 1e-13
    # loss cone of a dipole at L = 4.5: sin^4 a_lc = 1 / (L^5 (4L - 3))
    alc = math.asin((4.5 ** 5 * (4 * 4.5 - 3)) ** -0.25)
    x = 0.5 * (q.xe[:-1] + q.xe[1:])
    inside = x < alc
    assert inside.any() and not inside.all()
    assert np.all(q.inv_tau[inside] > 0) and np.all(q.inv_tau[~inside] == 0)
    assert np.all(p.inv_tau == 0)
    # f0 vanishes on the 5-degree loss-cone edge up to gEPS (Albert_Young.h:39)
    Emid = math.exp(0.5 * (p.ye[0] + p.ye[1]))
    mom = math.sqrt(Emid * (Emid + 2 * E0))
    a = 0.5 * (p.xe[0] + p.xe[1])
    want = math.exp(-(Emid - 0.2) / 0.1) * (math.sin(a) - math.sin(5 * math.pi / 180)) / mom ** 2
    assert abs(p.f0[0, 0] - sc.EPS - want) <= 1e-15 * abs(want) + 1e-30


@pytest.mark.gpu
@pytest.mark.parametrize("lc", [False, True], ids=["albert_young", "albert_young_lc"])
def test_p_ini_run_on_gpu_matches_oracle(tmp_path, lc):
    driver_with_hdf5()
    run, p = setup(tmp_path, lc)
    args = [DRIVER] + (["--case", "albert_young_lc"] if lc else []) + ["p.ini"]
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    h5 = str(tmp_path / "output" / run / f"{run}_data.h5")
    op = orc.OracleProblem(p)
    f = p.f0.copy()
    snap0 = read_h5(h5, "/f/0")
    assert rel_l2(snap0, f) <= 1e-15
    for k in range(1, 11):
        f = orc.run(op, f, 50)
        g = read_h5(h5, f"/f/{k}")
        e = rel_l2(g, f)
        assert e <= 1e-10, f"snapshot {k} (step {50 * k}): rel L2 {e:.3e}"
        assert g.min() >= 0.0
    t = read_h5(h5, "/t")
    assert np.allclose(t, np.linspace(0.0, 1.0, 11), rtol=0, atol=1e-15)
