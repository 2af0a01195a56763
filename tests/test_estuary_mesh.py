"""The reference's own Estuary example mesh (Examples/Estuary/tstinp, 845
nodes, 1570 cells, 25 open-boundary nodes; tests/golden/estuary_mesh.npz)
stepped with a depth floor of 2 m (its deck runs wetting and drying, which
this build does not have) and a 0.3 m M2 tide on its open boundary: the
oracle steps it stably; the HIP path reproduces the oracle bit for bit,
called from Python and from a Fortran host that hands the mesh over
through FVGPU_SETUP_MESH."""
import os
import subprocess

import numpy as np
import pytest

import myfvcomcode_amd as F
from helpers import ELEM2, ELEM3, NODE2, NODE3, compare

HERE = os.path.dirname(os.path.abspath(__file__))
KB, FLOOR, M2, AMP = 6, 2.0, 44712.0, 0.3


def _mesh():
    z = np.load(os.path.join(HERE, "golden", "estuary_mesh.npz"))
    return z["x"], z["y"], z["nv"], np.maximum(z["h"], FLOOR), z["obc"]


def estuary_model():
    x, y, nv, h, obc = _mesh()
    m = F.Model.from_mesh(x, y, nv, h, KB, obc, F.Model.case_defaults(0).p)
    m.set_tide([M2], [AMP], [0.0])
    m.cold_start()
    return m


def test_estuary_oracle_runs(built):
    from oracle import Oracle
    m = estuary_model()
    o = Oracle(m)
    o.run(20)
    el, ua = o.get("el"), o.get("ua")
    assert np.all(np.isfinite(el)) and np.all(np.isfinite(o.get("t1")))
    assert np.abs(ua).max() > 1e-3 and not o.depth_fail
    obc = m.field("i_obc_n")
    assert np.abs(el[obc]).max() > 0.01            # the tide has come in


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_estuary_gpu_parity(built, graph):
    from oracle import Oracle
    m = estuary_model()
    dev, orc = F.DeviceModel(m), Oracle(m)
    n = 10
    print("case: reference Estuary mesh 845 nodes KB=%d, %d steps, graph=%s" % (KB, n, graph), flush=True)
    dev.run(n, graph=graph)
    orc.run(n)
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0)
    assert not bad, bad


@pytest.mark.gpu
def test_fortran_host_mesh_matches_oracle(built, tmp_path):
    """A Fortran host reads the Estuary mesh, hands it to FVGPU_SETUP_MESH
    with the tide, runs INTERNAL_STEP and writes EL, UA and T1: they equal
    the oracle's on the same mesh bit for bit, and its REPORT line equals
    the oracle's budgets."""
    from oracle import Oracle
    exe = os.path.join(os.path.dirname(F.__file__), "lib", "fvcom_mesh")
    if not os.path.exists(exe):
        pytest.skip("the Fortran mesh driver is not built (no amdflang)")
    x, y, nv, h, obc = _mesh()
    mf, out = tmp_path / "mesh.txt", tmp_path / "state.txt"
    with open(mf, "w") as f:
        f.write("%d %d %d %d %d\n" % (len(x), len(nv), KB, len(obc), 1))
        for a, b, c in zip(x, y, h):
            f.write("%r %r %r\n" % (float(a), float(b), float(c)))
        for t in nv:
            f.write("%d %d %d\n" % tuple(int(v) for v in t))
        for j in obc:
            f.write("%d\n" % int(j))
        f.write("%r %r %r\n" % (M2, AMP, 0.0))
    n = 8
    print("case: Fortran host, Estuary mesh through FVGPU_SETUP_MESH, %d steps" % n, flush=True)
    r = subprocess.run([exe, str(mf), str(n), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = np.array([float(v) for v in open(out).read().split()])
    m = estuary_model()
    m1, n1 = m.m + 1, m.n + 1
    el, ua, t1 = vals[:m1], vals[m1:m1 + n1], vals[m1 + n1:]
    orc = Oracle(m)
    orc.run(n)
    assert np.array_equal(el, orc.get("el"))
    assert np.array_equal(ua, orc.get("ua"))
    assert np.array_equal(t1, orc.get("t1"))
    line = [l for l in r.stdout.splitlines() if "IINT=" in l][0].replace("=", " ").split()
    vol, salt, _ = orc.budgets()
    assert int(line[1]) == n
    assert abs(float(line[3]) - vol) <= 1e-15 * vol and abs(float(line[5]) - salt) <= 1e-15 * salt
