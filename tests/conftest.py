"""Shared test setup: the `gpu` marker, import paths, and small helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("sayram-2d_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="session")
def gpu_lib():
    import sayram_gpu as sg
    L = sg.lib()
    if L.sg_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return sg


H5DUMP = None
for _c in ("h5dump", "/opt/conda/bin/h5dump"):
    import shutil as _sh
    _p = _sh.which(_c) or (_c if os.path.exists(_c) else None)
    if _p:
        H5DUMP = _p
        break


def read_h5(path, name):
    """A dataset of an HDF5 file as float64, through h5dump at 17 significant
    digits (round-trips every double); shape from the dataspace."""
    import re
    import subprocess
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    out = subprocess.run([H5DUMP, "-d", name, "-m", "%.17g", "-y", "-w", "0", path],
                         capture_output=True, text=True, check=True).stdout
    dims = [int(v) for v in re.search(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \)", out).group(1).split(",")]
    body = out.split("DATA {", 1)[1].split("}", 1)[0]
    vals = np.array([float(v) for v in body.replace(",", " ").split()], dtype=np.float64)
    return vals.reshape(dims)
