"""The launch-parameter autotuner (sayram-2d_amd/sayram_tune.py) and the
invariances it relies on."""
import numpy as np
import pytest

import oracle as orc
import sayram_cases as sc
import sayram_gpu as sg
import sayram_tune as tu
from conftest import rel_l2


def test_tuner_defaults_are_the_library_defaults():
    o = tu.opts_for(sg.SG_PC32, tu.DEFAULTS)
    assert list(o.line_tiles) == [16, 160, 16, 768] and list(o.coarse_tiles) == [16, 160, 16, 768]
    assert o.check_every == 1 and o.use_graph == 1 and o.streams == 1 and o.fuse_cells == 1024
    assert o.group_cells == 1 << 21
    for name, cands in tu.CANDIDATES.items():
        assert tu.DEFAULTS[name] in cands, name


@pytest.mark.gpu
def test_check_every_never_changes_results(gpu_lib):
    probs = [sc.seeded_case(nx, ny, seed=70 + k, scale=s) for k, (nx, ny, s) in
             enumerate([(40, 33, 0.5), (17, 64, 4.0)])]
    out = []
    for ce in (1, 3):
        b = gpu_lib.Batch(gpu_lib.make_opts(precision=sg.SG_PC32, check_every=ce))
        slots = [b.add(p) for p in probs]
        b.step(4)
        out.append([b.f(s) for s in slots])
        b.close()
    for a, c in zip(*out):
        assert np.array_equal(a, c)


@pytest.mark.gpu
def test_tile_thresholds_change_rounding_only(gpu_lib):
    """Every tile shape solves the same lines exactly; other thresholds give the
    same trajectory up to rounding, and both stay on the oracle."""
    p = sc.seeded_case(60, 90, seed=71, scale=2.0)
    fs = []
    # defaults (16x64 rows, 32x32 columns here), 4x256 rows with 16x256 columns,
    # 64x16 everywhere
    for tiles in ((0, 0, 0, 0), (16, 40, 16, 64), (400, 400, 500, 500)):
        s = gpu_lib.Solver(p, gpu_lib.make_opts(precision=sg.SG_PC32, line_tiles=tiles))
        s.update(3)
        fs.append(s.f())
        s.close()
    ref = orc.run(orc.OracleProblem(p), p.f0, 3)
    for f in fs:
        assert rel_l2(f, fs[0]) <= 1e-12
        assert rel_l2(f, ref) <= 1e-10


@pytest.mark.gpu
def test_tuner_runs_and_returns_a_measured_point(gpu_lib):
    probs = [sc.seeded_case(48, 40, seed=72), sc.seeded_case(30, 90, seed=73)]
    best, v, trials = tu.tune(probs, steps=1, warmup=1,
                              order=("check_every", "col_mid"))
    assert v > 0 and len(trials) >= 4
    assert any(t == best and abs(tv - v) < 1e-9 * v for t, tv in trials)
