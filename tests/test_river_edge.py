"""Edge rivers (RIVER_INFLOW_LOCATION = 'edge') on the oracle, pinned by
what the set-up and the physics fix, not by the oracle's own formulas:

  set-up        SET_RIVER_BNDRY_METRICS (mod_force.F:3515-3568): the river's
                nodes are the two ends of its cell's one solid edge, their
                shares RDISQ are their CV areas' fractions (sum 1), QAREA is
                the edge length times the mean depth and ANGLEQ points from
                the edge into the cell; cells that are not on the solid
                boundary, or have two solid edges, are refused
  volume        a closed basin gains DTE*QDIS per external step, whichever
                kind the river is (extel_edge.F:160-173)
  momentum      from rest, the first step moves the river's cell along
                ANGLEQ (advave_edge_gcn.F:983-990, the wall projection's
                river-mouth rule bcond_gcn.F:362-372)
  T/S values    'specified': both nodes of the edge carry TDIS/SDIS after
                the step (bcond_ts.F:70-80); 'calculated': they are raised
                to at least TDIS/SDIS (adjust_ts.F:158-166)

GPU parity with edge rivers, alone and beside a node river, is at the end
of this file.
"""
import math

import numpy as np
import pytest

import myfvcomcode_amd as F
from oracle import Oracle

from helpers import ELEM2, ELEM3, NODE2, NODE3, compare
from test_oracle import closed_basin


def _one_edge_cells(m):
    """Cells with ISBCE = 1 and exactly one solid edge (NBE = 0)."""
    n1 = m.n + 1
    nbe = m.field("nbe").reshape(3, n1)
    isbce = m.field("isbce")
    zeros = (nbe[:, 1:] == 0).sum(axis=0)
    return [i for i in range(1, n1) if isbce[i] == 1 and zeros[i - 1] == 1]


def _west_cell(m):
    """The one-solid-edge cell nearest the middle of the channel's west end."""
    xc, yc = m.field("xc"), m.field("yc")
    cells = _one_edge_cells(m)
    ymid = 0.5 * (yc[1:].min() + yc[1:].max())
    return min(cells, key=lambda i: (xc[i] - xc[1:].min()) ** 2 + (yc[i] - ymid) ** 2)


def test_edge_river_setup(built):
    m = F.Model.build(0, river=0)
    cell = _west_cell(m)
    m.set_river_edge(cell, 50.0, 12.0, 1.0)
    assert m.hdr.numqbc == 1
    assert int(m.field("icellq")[0]) == cell and int(m.field("inodeq")[0]) == 0
    n1 = m.n + 1
    nq = m.field("n_icellq")
    a, b = int(nq[0]), int(nq[1])
    nv = m.field("nv").reshape(3, n1)[:, cell]
    nbe = m.field("nbe").reshape(3, n1)[:, cell]
    # the edge opposite the node whose NBE is 0 joins the other two nodes
    opposite = int(nv[list(nbe).index(0)])
    assert {a, b} == set(int(x) for x in nv) - {opposite}
    isonb = m.field("isonb")
    assert isonb[a] == 1 and isonb[b] == 1
    r = m.field("rdisq")
    art1 = m.field("art1")
    assert r[0] == pytest.approx(art1[a] / (art1[a] + art1[b]), rel=1e-15)
    assert r[0] + r[1] == pytest.approx(1.0, rel=1e-15)
    vx, vy, h = m.field("vx"), m.field("vy"), m.field("h")
    length = math.hypot(vx[a] - vx[b], vy[a] - vy[b])
    assert m.field("qarea")[0] == pytest.approx(length * 0.5 * (h[a] + h[b]), rel=1e-14)
    ang = m.field("angleq")[0]
    mx, my = 0.5 * (vx[a] + vx[b]), 0.5 * (vy[a] + vy[b])
    to_cell = np.array([m.field("xc")[cell] - mx, m.field("yc")[cell] - my])
    d = np.array([math.cos(ang), math.sin(ang)])
    assert d @ to_cell > 0.0                                      # into the cell
    assert abs(d @ np.array([vx[a] - vx[b], vy[a] - vy[b]])) < 1e-6 * length   # normal to the edge


def test_edge_river_refuses_bad_cells(built):
    m = F.Model.build(0, river=0)
    isbce = m.field("isbce")
    interior = int(np.flatnonzero(isbce[1:] == 0)[0]) + 1
    with pytest.raises(RuntimeError):
        m.set_river_edge(interior, 10.0)
    n1 = m.n + 1
    zeros = (m.field("nbe").reshape(3, n1)[:, 1:] == 0).sum(axis=0)
    corner = [i for i in range(1, n1) if isbce[i] == 1 and zeros[i - 1] >= 2]
    if corner:
        with pytest.raises(RuntimeError):
            m.set_river_edge(corner[0], 10.0)
    assert m.hdr.numqbc == 0


@pytest.mark.parametrize("kind", ["edge", "node"])
def test_volume_budget_equals_river_inflow(built, kind):
    q = 80.0
    m, nx, ny = closed_basin()
    if kind == "edge":
        m.set_river_edge(_west_cell(m), q, 15.0, 0.0)
    else:
        m.set_river((ny // 2) * nx + 1, q, 15.0, 0.0)
    m.cold_start()
    art1 = m.field("art1")[1:]
    o = Oracle(m)
    v0 = (art1 * o.get("d")[1:]).sum()
    p = m.hdr.p
    expect = 0.0
    for iint in range(1, 5):
        o.internal_step()
        expect += p.isplit * p.dte * q * math.tanh(iint / p.iramp)
        v = (art1 * o.get("d")[1:]).sum()
        assert v - v0 == pytest.approx(expect, rel=1e-9)


def test_edge_river_pushes_its_cell_along_angleq(built):
    m, nx, ny = closed_basin(p_obc_elevation_on=0)
    cell = _west_cell(m)
    m.set_river_edge(cell, 200.0, 15.0, 30.0)
    m.cold_start()
    o = Oracle(m)
    o.external_step(1)
    ua, va = o.get("ua"), o.get("va")
    ang = m.field("angleq")[0]
    along = ua[cell] * math.cos(ang) + va[cell] * math.sin(ang)
    assert along > 0.0
    assert along > 10.0 * abs(-ua[cell] * math.sin(ang) + va[cell] * math.cos(ang))


@pytest.mark.parametrize("specified", [1, 0])
def test_edge_river_tracer_values(built, specified):
    m, nx, ny = closed_basin(river_ts_specified=specified)
    cell = _west_cell(m)
    tdis, sdis = 25.0, 2.0
    m.set_river_edge(cell, 60.0, tdis, sdis)
    m.cold_start()
    o = Oracle(m)
    o.internal_step()
    nq = m.field("n_icellq")
    t1, s1 = o.get("t1").reshape(m.kb, -1), o.get("s1").reshape(m.kb, -1)
    for j in (int(nq[0]), int(nq[1])):
        if specified:
            assert np.all(t1[:-1, j] == tdis) and np.all(s1[:-1, j] == sdis)
        else:
            assert np.all(t1[:-1, j] >= tdis) and np.all(s1[:-1, j] >= sdis)
            assert np.all(t1[:-1, j] == tdis)       # the basin is at 15, below TDIS


# ---------------------------------------------------------------- GPU parity
@pytest.mark.gpu
@pytest.mark.parametrize("case", ["edge", "edge+node", "edge calculated"])
def test_edge_river_gpu_parity(built, case):
    kw = {"river": 0} if case != "edge+node" else {}
    if case == "edge calculated":
        kw["p_river_ts_specified"] = 0
    m = F.Model.build(0, **kw)
    m.set_river_edge(_west_cell(m), 120.0, 18.0, 0.5)
    if case == "edge+node":
        cells = _one_edge_cells(m)
        yc = m.field("yc")
        m.set_river_edge(max(cells, key=lambda i: yc[i]), 40.0, 10.0, 3.0)   # a second, on the north wall
    m.cold_start()
    dev, orc = F.DeviceModel(m), Oracle(m)
    n = 8
    print("case: 81x21 estuary channel, %s river(s), %d steps" % (case, n), flush=True)
    dev.run(n, graph=True)
    orc.run(n)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    for k, v in sorted(rep.items(), key=lambda x: -x[1])[:8]:
        print("  %-8s %.3e" % (k, v))
    d = dev.diagnostics()
    assert d.nonfinite == 0 and d.depth_fail == 0
    assert not bad, bad
