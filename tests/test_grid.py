"""Grid topology and metrics (libfvgrid), pinned by what geometry fixes.

The metrics restate GRID_METRICS (FVCOM_source/grid_metrics.F:55-134); the
reference ships no metric fixtures, so each array is checked against an
invariant a wrong index, sign or transposed operand would break.
"""
import numpy as np
import pytest

import myfvcomcode_amd as F


def tiny_mesh():
    """2x2 cells of a unit square split into 8 triangles (9 nodes)."""
    x = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2], float) * 100.0
    y = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2], float) * 100.0
    ccw = []
    for j in range(2):
        for i in range(2):
            a = j * 3 + i + 1
            b, d = a + 1, a + 3
            c = d + 1
            ccw += [(a, b, c), (a, c, d)]
    nv = np.array([[t[0], t[2], t[1]] for t in ccw], np.int32)  # clockwise
    h = np.full(9, 10.0)
    return x, y, nv, h


@pytest.fixture(scope="module")
def cfg0(built):
    return F.Model.build(0)


@pytest.fixture(scope="module")
def jit(built):
    return F.Model.build(1, nx=41, ny=21, kb=6)


def test_counts_config0(cfg0):
    m = cfg0
    assert (m.m, m.n) == (1701, 3200)
    assert m.hdr.ne == 4900                        # Euler: NE = N + M - 1
    assert m.hdr.ncv == 3 * m.n                    # tge.F:330
    assert m.hdr.iobcn == 19 and m.hdr.numqbc == 1


def test_tiny_mesh_topology(built):
    x, y, nv, h = tiny_mesh()
    m = F.Model.from_mesh(x, y, nv, h, 5)
    assert m.hdr.ne == 16 and m.hdr.ncv == 24
    isonb = m.field("isonb")[1:]
    assert list(isonb) == [1, 1, 1, 1, 0, 1, 1, 1, 1]   # only the centre node is interior
    ntsn = m.field("ntsn")[1:]
    ntve = m.field("ntve")[1:]
    assert ntve[4] == 6 and ntsn[4] == 7          # interior ring closes: NTSN = NTVE + 1
    assert ntsn[0] == ntve[0] + 3                 # boundary ring: self, NTVE+1 nodes, self
    off = m.field("nbsn_off")
    ring = m.field("nbsn")[off[5]:off[6]]
    assert ring[0] == ring[-1] and sorted(set(ring)) == [1, 2, 4, 6, 8, 9]  # diagonals 1-5-9
    # area of every triangle is 100*100/2
    assert np.allclose(m.field("art")[1:], 5000.0)


def test_area_partitions(cfg0, jit):
    for m in (cfg0, jit):
        art = m.field("art")[1:]
        art1 = m.field("art1")[1:]
        assert abs(art1.sum() - art.sum()) <= 1e-9 * art.sum()
        # ART2 = sum of ART around the node (cell_area.F:330-333)
        off = m.field("nbve_off")
        nbve = m.field("nbve")
        art_all = m.field("art")
        art2 = m.field("art2")
        for i in range(1, m.m + 1, 97):
            assert art2[i] == pytest.approx(art_all[nbve[off[i]:off[i + 1]]].sum(), rel=1e-14)
    assert cfg0.field("art")[1:].sum() == pytest.approx(40000.0 * 10000.0, rel=1e-12)


def test_lsq_gradient_exact_for_linear_fields(jit):
    """A1U/A2U give the exact gradient of a linear field at interior
    elements (least squares is exact on linear data, shape_coef_gcn.F:104-122)."""
    m = jit
    n1 = m.n + 1
    a1u = m.field("a1u").reshape(4, n1)
    a2u = m.field("a2u").reshape(4, n1)
    nbe = m.field("nbe").reshape(3, n1)
    xc, yc = m.field("xc"), m.field("yc")
    isbce = m.field("isbce")
    f = 3.0 + 2.5e-4 * xc - 7.0e-4 * yc
    idx = np.where(isbce[1:] == 0)[0] + 1
    vals = np.stack([f[idx], f[nbe[0, idx]], f[nbe[1, idx]], f[nbe[2, idx]]])
    gx = (a1u[:, idx] * vals).sum(0)
    gy = (a2u[:, idx] * vals).sum(0)
    assert np.allclose(gx, 2.5e-4, rtol=1e-9, atol=1e-13)
    assert np.allclose(gy, -7.0e-4, rtol=1e-9, atol=1e-13)
    # boundary elements fall back to first order (A1U = 0)
    bidx = np.where(isbce[1:] == 1)[0] + 1
    assert np.all(a1u[:, bidx] == 0.0)


def test_awx_awy_exact_gradient(jit):
    """AWX/AWY (WREAL's element gradient of nodal fields) are exact for
    linear fields on every element."""
    m = jit
    n1 = m.n + 1
    nv = m.field("nv").reshape(3, n1)
    awx = m.field("awx").reshape(3, n1)
    awy = m.field("awy").reshape(3, n1)
    vx, vy = m.field("vx"), m.field("vy")
    f = 1.0 + 4e-4 * vx + 9e-4 * vy
    gx = (awx[:, 1:] * f[nv[:, 1:]]).sum(0)
    gy = (awy[:, 1:] * f[nv[:, 1:]]).sum(0)
    assert np.allclose(gx, 4e-4, rtol=1e-9)
    assert np.allclose(gy, 9e-4, rtol=1e-9)


def test_control_volume_polygons_close(jit):
    """Sum of signed CV sub-edge vectors around an interior node is zero, and
    the ring vectors DLTXTRIE/DLTYTRIE close (tge.F:611-667, edge_len.F)."""
    m = jit
    ncv = m.hdr.ncv
    niec = m.field("niec").reshape(2, ncv + 1)
    dx, dy = m.field("dltxe"), m.field("dltye")
    sx = np.zeros(m.m + 1)
    sy = np.zeros(m.m + 1)
    np.add.at(sx, niec[0, 1:], -dx[1:])
    np.add.at(sx, niec[1, 1:], dx[1:])
    np.add.at(sy, niec[0, 1:], -dy[1:])
    np.add.at(sy, niec[1, 1:], dy[1:])
    inner = np.where(m.field("isonb") == 0)[0]
    inner = inner[inner > 0]
    scale = 200.0
    assert np.abs(sx[inner]).max() < 1e-9 * scale
    assert np.abs(sy[inner]).max() < 1e-9 * scale
    off = m.field("nbsn_off")
    tx, ty = m.field("dltxtrie"), m.field("dltytrie")
    for i in inner[::37]:
        assert abs(tx[off[i]:off[i + 1]].sum()) < 1e-9 * scale
        assert abs(ty[off[i]:off[i + 1]].sum()) < 1e-9 * scale


def test_cv_areas_from_divergence(jit):
    """Divergence theorem on each interior node's control volume: the flux
    of the field (x, 0) through its sub-edges equals ART1 (tge/cell_area)."""
    m = jit
    ncv = m.hdr.ncv
    niec = m.field("niec").reshape(2, ncv + 1)
    dy = m.field("dltye")
    flux = np.zeros(m.m + 1)
    # outward flux of (1,0)*x ... use f = (x, 0): flux through sub-edge = x_mid * (-DLTYE)
    # evaluated exactly for linear fields with the midpoint rule
    ntrg = m.field("ntrg")
    xc, xijc = m.field("xc"), m.field("xijc")
    # sub-edge endpoints: element centroid and edge midpoint -> midpoint x
    # recover the edge midpoint from DLTXE: xij = xc(ntrg) + dltxe
    xm = xc[ntrg[1:]] + 0.5 * m.field("dltxe")[1:]
    np.add.at(flux, niec[0, 1:], xm * dy[1:])
    np.add.at(flux, niec[1, 1:], -xm * dy[1:])
    inner = np.where(m.field("isonb") == 0)[0]
    inner = inner[inner > 0]
    art1 = m.field("art1")
    assert np.allclose(np.abs(flux[inner]), art1[inner], rtol=1e-9)


def test_sigma_levels(cfg0):
    m = cfg0
    z = m.field3("z")[:, 1]
    dz = m.field3("dz")[:, 1]
    assert z[0] == 0.0 and z[-1] == -1.0
    assert abs(dz[:-1].sum() - 1.0) < 1e-12
    assert np.all(dz[:-1] > 0)


def test_obc_setup(cfg0):
    m = cfg0
    obc = m.field("i_obc_n")
    nxt = m.field("next_obc")
    vx = m.field("vx")
    # east boundary: the neighbour inward from each OBC node lies to the west
    assert np.all(vx[nxt] < vx[obc])
    dn = m.field("dltn_obc")
    assert np.allclose(dn, 500.0)
    ff = m.field("fluxf_obc").reshape(2, -1)
    assert np.allclose(ff.sum(0), 1.0)


def test_estuary_mesh_metrics(built):
    """The reference's Estuary example mesh (845 nodes, irregular rings)."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "estuary_mesh.npz"))
    m = F.Model.from_mesh(z["x"], z["y"], z["nv"], np.maximum(z["h"], 2.0), 11, z["obc"])
    assert (m.m, m.n) == (845, 1570)
    assert m.hdr.ne == m.n + m.m - 1
    art, art1 = m.field("art")[1:], m.field("art1")[1:]
    assert abs(art.sum() - art1.sum()) < 1e-9 * art.sum()
    assert m.field("ntsn").max() <= 13


def test_wheel_mesh_reaches_the_widest_lists(built):
    """The high-degree test mesh really has a node with 12 triangles (ring of
    13 entries, 24 control-volume sub-edges), so the device parity test on it
    runs the RMAX = 16 and SMAX = 24 kernel variants; the oracle steps it."""
    from helpers import wheel_mesh
    from oracle import Oracle
    m = wheel_mesh()
    assert m.field("ntsn")[1:].max() == 13 and m.field("ntve")[1:].max() == 12
    niec = m.field("niec").reshape(2, -1)[:, 1:]
    cnt = np.bincount(niec.ravel(), minlength=m.m + 1)
    assert cnt.max() == 24
    o = Oracle(m)
    o.run(5)
    assert np.abs(o.get("ua")).max() > 1e-3 and np.all(np.isfinite(o.get("t1")))
