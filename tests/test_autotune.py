"""The autotuner (fvgpu_autotune) only changes workgroup sizes: the model
state is restored after tuning and the tuned step produces bit-identical
results to the untuned one."""
import numpy as np
import pytest

import myfvcomcode_amd as F

pytestmark = pytest.mark.gpu


def test_autotune_keeps_results_bit_identical(built):
    m = F.Model.build(1, nx=61, ny=31, kb=9, n_seamounts=2)
    print("case: jittered mesh 61x31 KB=9, 3 + autotune + 3 steps vs 6 steps", flush=True)
    a = F.DeviceModel(m)
    a.run(3, graph=False)
    rep = a.autotune(1)
    assert rep and all(size in (64, 128, 256) for size, _ in rep.values())
    a.run(3, graph=True)
    b = F.DeviceModel(m)
    b.run(6, graph=True)
    for name in ("el", "ua", "va", "u", "v", "w", "t1", "s1", "q2", "km", "kh", "wts"):
        assert np.array_equal(a.get(name), b.get(name)), name
    assert a.diagnostics().iint == b.diagnostics().iint == 6


def test_concurrency_off_is_bit_identical(built):
    """The step's independent chains on one stream (concurrency 0) give the
    same bits as with every pair on two streams (3), eager and captured."""
    m = F.Model.build(1, nx=61, ny=31, kb=9, n_seamounts=2)
    print("case: jittered mesh 61x31 KB=9, 4 steps on one stream vs two", flush=True)
    a = F.DeviceModel(m)
    a.set_concurrency(0)
    a.run(2, graph=False)
    a.run(2, graph=True)
    b = F.DeviceModel(m)
    b.set_concurrency(3)
    b.run(2, graph=False)
    b.run(2, graph=True)
    for name in ("el", "ua", "va", "u", "v", "w", "t1", "s1", "q2", "q2l", "l", "km", "kh", "kq"):
        assert np.array_equal(a.get(name), b.get(name)), name


def test_profiling_times_one_in_n_launches(built):
    """Timing events around one in `every` launches of a family: the count
    is the sample, the average a device time, the results unchanged."""
    m = F.Model.build(0)
    print("case: 81x21 channel, 2 eager steps timing 1 in 4 ext_momentum launches", flush=True)
    a = F.DeviceModel(m)
    a.set_profiling(True, kernel="ext_momentum", every=4)
    a.run(2, graph=False)
    a.set_profiling(False)
    stats = a.kernel_stats()
    us, cnt = stats["ext_momentum"]
    per_step = 4 * m.hdr.p.isplit
    assert cnt == 2 * per_step // 4
    assert 0.0 < us < 1e5
    assert set(stats) == {"ext_momentum"}
    b = F.DeviceModel(m)
    b.run(2, graph=False)
    for name in ("el", "ua", "va", "u", "t1"):
        assert np.array_equal(a.get(name), b.get(name)), name
    with pytest.raises(RuntimeError):
        a.set_profiling(True, kernel="ext_momentum", every=0)
