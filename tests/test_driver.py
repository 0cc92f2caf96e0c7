"""The C++ drop-in (host/GpuSolver.h) through the driver that mirrors
source/main.cc: p.ini parsing with the reference's nsteps rounding, the time
loop, and the HDF5 snapshot layout of source/main.cc:57-89."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
from conftest import ROOT, rel_l2

DRIVER = os.path.join(ROOT, "sayram-2d_amd", "sayram_gpu_run")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

INI = """[basic]
run_id = {run}
# a comment line without '=' (ignored, as mINI does)
nalpha0 = {nx}
nE = {ny}
alpha0min = 5
alpha0max = 90
Emin = 0.2
Emax = 5
; a ';' comment
T = 0.0105
nsteps = 11

[diagnostics]
nplots = 5

[diffusion_coefficients]
dID = none
"""


def _setup(tmp_path, nx, ny):
    # nsteps = 11 rounds down to 10 = 2 * nplots; dt = T / 10 = 1.05e-3
    (tmp_path / "p.ini").write_text(INI.format(run="drv", nx=nx, ny=ny))
    p = sc.seeded_case(nx, ny, seed=77, scale=2.0, dt=0.0105 / 10)
    sc.write_problem(p, str(tmp_path / "prob.bin"))
    return p


def test_driver_without_gpu_reports_error(tmp_path):
    import sayram_gpu as sg
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    if sg.lib().sg_device_count() > 0:
        pytest.skip("a GPU is visible")
    _setup(tmp_path, 6, 5)
    r = subprocess.run([DRIVER, "p.ini", "prob.bin"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "no GPU visible" in r.stderr


@pytest.mark.gpu
def test_driver_matches_oracle_and_writes_reference_layout(tmp_path):
    nx, ny = 24, 20
    p = _setup(tmp_path, nx, ny)
    r = subprocess.run([DRIVER, "p.ini", "prob.bin"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "CPU time used" in r.stdout
    out = tmp_path / "output" / "drv"
    f = np.fromfile(out / "drv_final.bin", dtype=np.float64).reshape(nx, ny)
    ref = orc.run(orc.OracleProblem(p), p.f0, 10)
    assert rel_l2(f, ref) <= 1e-10
    h5 = out / "drv_data.h5"
    if not h5.exists():
        pytest.skip("driver built without HDF5")
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    hdr = subprocess.run([H5DUMP, "-H", str(h5)], capture_output=True, text=True).stdout
    dims = dict(re.findall(r'DATASET "([^"]+)" \{.*?DATASPACE\s+SIMPLE \{ \( ([^)]*) \)', hdr, re.S))
    assert dims["alpha0"].strip() == str(nx) and dims["logEN"].strip() == str(ny)
    assert dims["t"].strip() == "6"
    groups = re.findall(r'GROUP "(\d+)"|DATASET "(\d)"', hdr)
    snaps = sorted(int(b) for a, b in groups if b)
    assert snaps == [0, 1, 2, 3, 4, 5]
    d = subprocess.run([H5DUMP, "-d", "/f/5", "-m", "%.17g", "-y", "-w", "0", str(h5)], capture_output=True,
                       text=True).stdout
    body = d.split("DATA {", 1)[1].split("}", 1)[0]
    vals = np.array([float(x) for x in body.replace(",", " ").split()])
    assert vals.shape == (nx * ny,)
    assert rel_l2(vals.reshape(nx, ny), ref) <= 1e-10


def test_driver_config_side_effects_without_gpu(tmp_path):
    """Parameters' side effects and errors (source/Parameters.cc:7-23, 59-60):
    no argument reads ./p.ini, the ini file is copied to
    ./output/<run_id>/<run_id>.ini, and nsteps < nplots is an error instead of
    0 steps with dt = T / 0. These run before the GPU is touched."""
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    _setup(tmp_path, 6, 5)
    r = subprocess.run([DRIVER, "--case", "albert_young_lc"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    # ./p.ini was read (it names dID = none, whose table does not exist) and
    # copied before the Equation was built
    assert r.returncode == 2 and "D/none.h5" in r.stderr
    assert (tmp_path / "output" / "drv" / "drv.ini").read_text() == (tmp_path / "p.ini").read_text()
    bad = INI.format(run="few", nx=6, ny=5).replace("nsteps = 11", "nsteps = 3")
    (tmp_path / "few.ini").write_text(bad)
    r = subprocess.run([DRIVER, "few.ini", "prob.bin"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "less than nplots" in r.stderr
    r = subprocess.run([DRIVER, "missing.ini"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "cannot open missing.ini" in r.stderr


@pytest.mark.gpu
def test_driver_copies_ini_and_reads_default_p_ini(tmp_path):
    _setup(tmp_path, 12, 10)
    (tmp_path / "p.ini").write_text(INI.format(run="dflt", nx=12, ny=10))
    sc.write_problem(sc.seeded_case(12, 10, seed=77, scale=2.0, dt=0.0105 / 10),
                     str(tmp_path / "prob.bin"))
    r = subprocess.run([DRIVER, "p.ini", "prob.bin"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "output" / "dflt" / "dflt.ini").read_text() == (tmp_path / "p.ini").read_text()


@pytest.mark.gpu
def test_driver_reports_a_rejected_step(tmp_path):
    """A step the library rejects reaches the driver as an exception from
    GpuSolver::update (the reference throws std::runtime_error on errors)."""
    _setup(tmp_path, 24, 20)
    r = subprocess.run([DRIVER, "--max-iters", "1", "--rtol", "1e-16", "p.ini", "prob.bin"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "rejected" in r.stderr
