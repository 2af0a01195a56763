"""Reverse Cuthill-McKee numbering (fv_rcm_order): a mesh whose node and
element numbers were shuffled, as a grid file may number them, gets back a
narrow band; the time step on the renumbered mesh is the same physics (the
oracle's results agree under the permutation to round-off, the sums being
taken in another order), and the device reproduces the oracle bit for bit on
it."""
import numpy as np
import pytest

import myfvcomcode_amd as F


def _channel(nx=41, ny=11):
    c = F.Model.build(0, nx=nx, ny=ny, kb=6, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    return x, y, nv, c.field("h")[1:].copy(), c.field("i_obc_n").copy()


def _shuffled(seed=3, **kw):
    x, y, nv, h, obc = _channel(**kw)
    rng = np.random.default_rng(seed)
    pn = rng.permutation(len(x)).astype(np.int32) + 1
    pe = rng.permutation(len(nv)).astype(np.int32) + 1
    return F.renumber_mesh(x, y, nv, h, obc, pn, pe)


def _bandwidth(nv):
    return int(np.max(nv.max(axis=1) - nv.min(axis=1)))


def _model(mesh):
    x, y, nv, h, obc = mesh
    m = F.Model.from_mesh(x, y, nv, h, 6, obc, F.Model.case_defaults(0).p)
    m.set_tide([44712.0], [0.5], [0.0])
    m.cold_start()
    return m


def test_rcm_is_a_permutation_and_narrows_the_band(built):
    x, y, nv, h, obc = _channel()
    sx, sy, snv, sh, sobc = _shuffled()
    node_new, elem_new = F.rcm_order(len(sx), snv)
    assert sorted(node_new) == list(range(1, len(sx) + 1))
    assert sorted(elem_new) == list(range(1, len(snv) + 1))
    rx, ry, rnv, rh, robc = F.renumber_mesh(sx, sy, snv, sh, sobc, node_new, elem_new)
    # elements follow the nodes: element numbers grow with their nodes
    lo = rnv.min(axis=1)
    assert np.all(np.diff(lo) >= 0)
    print("bandwidth: generator %d, shuffled %d, RCM %d" % (_bandwidth(nv), _bandwidth(snv), _bandwidth(rnv)))
    assert _bandwidth(snv) > 10 * _bandwidth(nv)
    assert _bandwidth(rnv) <= 1.5 * _bandwidth(nv)
    # the same triangles, the same open boundary in the same order
    old_of_new = np.empty(len(sx), np.int64)
    old_of_new[node_new - 1] = np.arange(len(sx))
    assert np.array_equal(rx, sx[old_of_new]) and np.array_equal(rh, sh[old_of_new])
    assert np.array_equal(old_of_new[robc - 1] + 1, sobc)


def test_rcm_mesh_steps_the_same_physics(built):
    """Oracle on the shuffled mesh and on its RCM renumbering: after 5
    internal steps every field agrees under the permutation to 1e-9
    relative (only the summation order of the scatter loops differs)."""
    from oracle import Oracle
    sm = _shuffled()
    node_new, elem_new = F.rcm_order(len(sm[0]), sm[2])
    rm = F.renumber_mesh(*sm, node_new, elem_new)
    a, b = _model(sm), _model(rm)
    oa, ob = Oracle(a), Oracle(b)
    oa.run(5)
    ob.run(5)
    kb = a.kb
    for name in ("el", "ua", "u", "t1", "s1", "q2"):
        va, vb = oa.get(name), ob.get(name)
        node = va.size in (a.m + 1, (a.m + 1) * kb)
        perm = node_new if node else elem_new
        if va.size in (a.m + 1, a.n + 1):
            mapped = va[1:][np.argsort(perm)]
            other = vb[1:]
        else:
            n1 = va.size // kb
            mapped = va.reshape(kb, n1)[:-1, 1:][:, np.argsort(perm)]
            other = vb.reshape(kb, n1)[:-1, 1:]
        scale = max(np.abs(other).max(), 1e-30)
        assert np.abs(mapped - other).max() <= 1e-9 * scale, name
    assert np.abs(ob.get("ua")).max() > 1e-3


@pytest.mark.gpu
def test_rcm_mesh_gpu_parity(built):
    from oracle import Oracle
    from helpers import ELEM2, ELEM3, NODE2, NODE3, compare
    sm = _shuffled()
    node_new, elem_new = F.rcm_order(len(sm[0]), sm[2])
    for mesh, what in ((sm, "shuffled"), (F.renumber_mesh(*sm, node_new, elem_new), "RCM")):
        m = _model(mesh)
        print("case: 41x11 channel, %s numbering, 6 steps (graph)" % what, flush=True)
        dev, orc = F.DeviceModel(m), Oracle(m)
        dev.run(6, graph=True)
        orc.run(6)
        worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0)
        assert not bad, (what, bad)
        del dev
