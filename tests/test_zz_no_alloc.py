"""The no-allocation read-out sg_batch_get_f_static against sg_batch_get_f:
equal bit for bit on two calls in a row (the second on a smaller instance,
whose read-out reuses the static scratch the first one filled further), in a
batch of both instance groups, and over-size instances rejected with a clear
error."""
import numpy as np
import pytest

import sayram_cases as sc

pytestmark = pytest.mark.gpu


def test_static_readout_equals_allocating_readout(gpu_lib):
    sg = gpu_lib
    probs = [sc.seeded_case(1500, 1400, seed=201, scale=2.0),   # second instance group
             sc.seeded_case(300, 250, seed=202, scale=0.5),
             sc.seeded_case(37, 21, seed=203, scale=4.0)]
    b = sg.Batch(sg.make_opts(precision=sg.SG_PC32))
    slots = [b.add(p) for p in probs]
    b.step(2)
    for sl in slots:            # larger first, then smaller: two calls in a row each
        a = b.f(sl)
        s = b.f_static(sl)
        assert np.array_equal(a, s)
        assert np.array_equal(b.f_static(sl), a)
    b.step(1)
    assert np.array_equal(b.f_static(slots[2]), b.f(slots[2]))
    assert np.array_equal(b.f_static(slots[0]), b.f(slots[0]))
    b.close()


def test_static_readout_rejects_oversize_instances(gpu_lib):
    sg = gpu_lib
    b = sg.Batch(sg.make_opts(precision=sg.SG_PC32))
    big = b.add(sc.seeded_case(2049, 2049, seed=204))      # 4.2 Mi cells > 4 Mi
    small = b.add(sc.seeded_case(64, 48, seed=205))
    with pytest.raises(sg.SayramError) as e:
        b.f_static(big)
    assert e.value.code == sg.SG_ERR_ARG and "static read-out scratch" in str(e.value)
    out = np.empty(16)
    rc = sg.lib().sg_batch_get_f_static(b.h, 7, out.ctypes.data_as(sg._dp))   # no such slot
    assert rc == sg.SG_ERR_ARG and b"slot" in sg.lib().sg_last_error()
    assert np.array_equal(b.f_static(small), b.f(small))
    b.close()
