"""Snapshot and restart: a run resumed from a file written after N steps
continues bit for bit as the uninterrupted run, and the file carries
FVCOM's restart variables in FVCOM's layout."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import myfvcomcode_amd as F
from myfvcomcode_amd import restart


def _oracle_pair(tmp_path, steps_before=2, steps_after=2):
    from oracle import Oracle
    m = F.Model.build(0, nx=21, ny=7, kb=6)
    a = Oracle(m)
    a.run(steps_before)
    path = str(tmp_path / "state.nc")
    restart.save(path, m, a, a.iint)
    a.run(steps_after)
    m2 = F.Model.build(0, nx=21, ny=7, kb=6)
    iint, arrays = restart.load(path, m2)
    b = Oracle(m2)
    restart.put(b, arrays)
    b.run(steps_after)
    return m, a, b, path, iint


def test_oracle_resumes_bit_for_bit(built, tmp_path):
    m, a, b, path, iint = _oracle_pair(tmp_path)
    assert iint == 2 and a.iint == b.iint == 4
    bad = [n for n in restart.STATE if not np.array_equal(a.get(n), b.get(n))]
    assert not bad, bad


def test_file_layout_matches_fvcom(built, tmp_path):
    from oracle import Oracle
    m = F.Model.build(0, nx=21, ny=7, kb=6)
    o = Oracle(m)
    o.run(1)
    path = str(tmp_path / "s.nc")
    restart.save(path, m, o, o.iint)
    f = netcdf_file(path, "r", mmap=False)
    try:
        v = f.variables
        assert f.dimensions["node"] == m.m and f.dimensions["nele"] == m.n
        assert f.dimensions["siglay"] == m.kb - 1 and f.dimensions["siglev"] == m.kb
        assert v["temp"].dimensions == ("time", "siglay", "node")
        assert v["omega"].dimensions == ("time", "siglev", "node")
        assert v["u"].dimensions == ("time", "siglay", "nele")
        np.testing.assert_array_equal(v["zeta"][0], o.get("el")[1:])
        np.testing.assert_array_equal(v["temp"][0], o.get("t1").reshape(m.kb, -1)[:m.kb - 1, 1:])
        np.testing.assert_array_equal(v["nv"][:], m.field("nv").reshape(3, -1)[:, 1:])
        assert int(v["iint"][0]) == 1
    finally:
        f.close()


def test_load_rejects_other_mesh(built, tmp_path):
    from oracle import Oracle
    m = F.Model.build(0, nx=21, ny=7, kb=6)
    o = Oracle(m)
    path = str(tmp_path / "s.nc")
    restart.save(path, m, o, 0)
    with pytest.raises(ValueError):
        restart.load(path, F.Model.build(0, nx=23, ny=7, kb=6))


@pytest.mark.gpu
def test_device_resumes_bit_for_bit(built, tmp_path):
    """The device resumed from its own snapshot continues bit for bit, and a
    snapshot of the oracle resumes on the device as the oracle continues."""
    from oracle import Oracle
    m = F.Model.build(1, nx=61, ny=31, kb=9, n_seamounts=2)
    print("case: jittered mesh 61x31 KB=9, 3 steps, snapshot, 3 more (eager and graph)", flush=True)
    a = F.DeviceModel(m)
    a.run(3, graph=True)
    path = str(tmp_path / "d.nc")
    restart.save(path, m, a, a.diagnostics().iint)
    a.run(3, graph=True)
    m2 = F.Model.build(1, nx=61, ny=31, kb=9, n_seamounts=2)
    iint, arrays = restart.load(path, m2)
    b = F.DeviceModel(m2)
    restart.put(b, arrays)
    b.run(1, graph=False)
    b.run(2, graph=True)
    assert b.diagnostics().iint == a.diagnostics().iint == 6
    bad = [n for n in restart.STATE if not np.array_equal(a.get(n), b.get(n))]
    assert not bad, bad
