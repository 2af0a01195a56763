"""fvgpu_ensemble_update_noalloc (staging in the library's fixed device
array, no allocation) against fvgpu_ensemble_update (staging in a buffer
the handle allocates) on two identical handles: bit-identical arrays after
two calls in a row, the second smaller (fewer members, a 2-D array), and a
clear error beyond the 64-member maximum."""
import ctypes as C

import numpy as np
import pytest

import myfvcomcode_amd as F
from myfvcomcode_amd import enkf

pytestmark = pytest.mark.gpu


def test_noalloc_update_matches_allocating_update(built):
    parts = [F.Model.build(0, nx=21, ny=7, kb=6, seed=300 + j) for j in range(5)]
    u, offs = F.Model.union(parts)
    print("case: 5 channel members 21x7 KB=6; update t1 (5 members), then el (3 members)", flush=True)
    dev_a, dev_b = F.DeviceModel(u), F.DeviceModel(u)
    dev_a.run(2, graph=False)
    dev_b.run(2, graph=False)
    for name in ("t1", "el", "ua"):
        assert np.array_equal(dev_a.get(name), dev_b.get(name))
    rng = np.random.default_rng(17)
    calls = [(5, "t1"), (3, "el")]
    for nens, name in calls:
        ens = enkf.Ensemble(u, offs, nens)
        X5 = np.eye(nens) + 0.25 * rng.standard_normal((nens, nens))
        before = dev_a.get(name)
        ens.update(dev_a, name, X5)
        ens.update(dev_b, name, X5, noalloc=True)
        a, b = dev_a.get(name), dev_b.get(name)
        assert not np.array_equal(a, before), name
        assert np.array_equal(a, b), name
        print("  %s with %d members: identical" % (name, nens), flush=True)
    # the untouched arrays stay identical too
    for name in ("t1", "el", "s1", "u"):
        assert np.array_equal(dev_a.get(name), dev_b.get(name)), name


def test_noalloc_rejects_more_than_64_members(built):
    parts = [F.Model.build(0, nx=21, ny=7, kb=6, seed=310 + j) for j in range(2)]
    u, offs = F.Model.union(parts)
    dev = F.DeviceModel(u)
    L = F.gpu_lib()
    off = np.zeros(65, np.int64)
    x5 = np.eye(65)
    rc = L.fvgpu_ensemble_update_noalloc(dev._h, b"el", 65, off.ctypes.data, off.ctypes.data, 1, 1, x5.ctypes.data)
    assert rc != 0
    L.fvgpu_last_error.restype = C.c_char_p
    assert b"64" in L.fvgpu_last_error()
