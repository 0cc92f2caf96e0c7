"""CPU checks of the C ABI library and the host-side logic (no GPU needed).

The library must load, export every function include/sayram_gpu.h declares,
and fail loudly (SG_ERR_HIP, no CPU fallback) when no GPU is visible.
"""
import os
import re

import numpy as np
import pytest

import sayram_cases as sc
import sayram_gpu as sg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sayram_gpu.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_solver_surface():
    names = header_functions()
    for must in ("sg_create", "sg_step", "sg_get_f", "sg_destroy", "sg_batch_create",
                 "sg_batch_add", "sg_batch_step", "sg_batch_remove"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = sg.lib()
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_python_front_end_binds_every_symbol():
    bound = {s[0] for s in sg.SYMBOLS}
    assert set(header_functions()) <= bound


def test_abi_version():
    assert sg.lib().sg_abi_version() == 4


def test_no_gpu_fails_loudly():
    L = sg.lib()
    if L.sg_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(sg.SayramError) as e:
        sg.Solver(sc.seeded_case(4, 3))
    assert e.value.code == sg.SG_ERR_HIP and "no CPU fallback" in str(e.value)
    with pytest.raises(sg.SayramError) as e:
        sg.Batch()
    assert e.value.code == sg.SG_ERR_HIP


def test_option_errors_come_before_any_device_call():
    for opts in (sg.make_opts(precision=3), sg.make_opts(fuse_cells=1025), sg.make_opts(streams=3),
                 sg.make_opts(warm_start=2), sg.make_opts(warm_start=-2)):
        with pytest.raises(sg.SayramError) as e:
            sg.Batch(opts)
        assert e.value.code == sg.SG_ERR_ARG
    with pytest.raises(sg.SayramError) as e:
        sg.Batch(sg.make_opts(fuse_cells=2048))          # FP64 V-cycles: at most 1024
    assert "fuse_cells" in str(e.value)
    for opts in (sg.make_opts(precision=sg.SG_PC32, fuse_cells=4097),
                 sg.make_opts(coarse_smoother=3), sg.make_opts(coarse_smoother=6)):
        with pytest.raises(sg.SayramError) as e:
            sg.Batch(opts)
        assert e.value.code == sg.SG_ERR_ARG
    if sg.lib().sg_device_count() == 0:
        # valid options get as far as the missing GPU
        for opts in (sg.make_opts(precision=sg.SG_PC32, fuse_cells=4096),
                     sg.make_opts(coarse_smoother=4)):
            with pytest.raises(sg.SayramError) as e:
                sg.Batch(opts)
            assert e.value.code == sg.SG_ERR_HIP


def test_instance_bytes_accounting():
    b64 = [sg.Batch.instance_bytes(n, n, sg.SG_FP64) for n in (32, 64, 512, 4096)]
    assert all(x < y for x, y in zip(b64, b64[1:]))
    assert sg.Batch.instance_bytes(4096, 4096, sg.SG_MIXED) < b64[-1]
    # FP64 Krylov vectors with an FP32 hierarchy sit between the two
    pc = sg.Batch.instance_bytes(4096, 4096, sg.SG_PC32)
    assert sg.Batch.instance_bytes(4096, 4096, sg.SG_MIXED) < pc < b64[-1]
    assert sg.Batch.instance_bytes(0, 5, sg.SG_FP64) == 0
    assert sg.Batch.instance_bytes(8, 8, 3) == 0          # no such precision
    # bounded per-cell footprint (memory sized for many resident instances)
    assert b64[-1] / (4096 * 4096) < 300


def test_seeded_case_is_deterministic_and_well_formed():
    a = sc.seeded_case(37, 21, seed=5, scale=2.0)
    b = sc.seeded_case(37, 21, seed=5, scale=2.0)
    for k in ("G", "Dxx", "Dxy", "Dyy", "f0"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    assert a.G.shape == (37, 21) and a.xe.shape == (38,) and a.ye.shape == (22,)
    assert np.all(np.diff(a.xe) > 0) and np.all(np.diff(a.ye) > 0)
    assert np.all(a.G > 0) and np.all(a.f0 >= 0)
    day = 86400.0 * 2.0
    assert a.Dxx.min() >= 1e-6 * day * (1 - 1e-12) and a.Dxx.max() <= 1e-2 * day * (1 + 1e-12)
    assert np.all(np.abs(a.Dxy) <= 0.9 * np.sqrt(a.Dxx * a.Dyy) * (1 + 1e-12))
    # XMIN Dirichlet 0 (loss cone), XMAX zero flux, YMIN/YMAX Dirichlet f0
    assert a.bc_type == [sc.DIRICHLET, sc.ZEROFLUX, sc.DIRICHLET, sc.DIRICHLET]
    assert np.all(a.bc_val[0] == 0.0)


def test_uniform_edges_follow_make_uniform():
    """xe[i] = min + i*(max-min)/n (source/main.cc:25-34), degrees to radians as
    a*pi/180 (source/Parameters.h:35-36)."""
    p = sc.seeded_case(90, 100)
    d = (90 * sc.GPI / 180 - 5 * sc.GPI / 180) / 90.0
    assert p.xe[7] == 5 * sc.GPI / 180 + d * 7.0
    assert p.ye[0] == np.log(0.2) and abs(p.ye[-1] - np.log(5.0)) < 1e-15


def test_batch_specs():
    s = sc.batch_specs(512, 32, 512, 2026)
    assert len(s) == 512
    assert all(32 <= x["nx"] <= 512 and 32 <= x["ny"] <= 512 for x in s)
    assert all(0.1 <= x["scale"] <= 10.0 for x in s)
    assert s == sc.batch_specs(512, 32, 512, 2026)


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of sg_opts and sg_step_stats have the C layout: a
    program compiled against include/sayram_gpu.h prints sizes and offsets."""
    import ctypes
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = {"sg_opts": [f for f, _ in sg.SgOpts._fields_],
              "sg_step_stats": [f for f, _ in sg.SgStats._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sayram_gpu.h"', 'int main(void) {']
    for t, fs in fields.items():
        lines.append(f'printf("{t} %zu\\n", sizeof({t}));')
        for f in fs:
            lines.append(f'printf("{t}.{f} %zu\\n", offsetof({t}, {f}));')
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                          check=True).stdout.split("\n") if l)
    for t, cls in (("sg_opts", sg.SgOpts), ("sg_step_stats", sg.SgStats)):
        assert int(out[t]) == ctypes.sizeof(cls), t
        for f in fields[t]:
            assert int(out[f"{t}.{f}"]) == getattr(cls, f).offset, f"{t}.{f}"
