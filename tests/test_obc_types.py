"""Open-boundary types of the external mode (TYPE_OBC, mod_obcs.F:30-39) on
the oracle, pinned without the oracle's own formulas where the physics
fixes the answer:

  IUCP          odd types linearise the cells next to the node (SETUP_OBC,
                mod_obcs.F:959-962, 993-996), even types leave them
                nonlinear; a type outside 1..8 is refused
  ASL_CLP       types 3/4 hold the boundary elevation at exactly zero
                (mod_obcs.F:599-617)
  GWI, BKI      types 5/6 and 7/8 tie the boundary node to its interior
                neighbour's new elevation in every RK stage (mod_obcs.F:
                624-699); after a whole external step the last stage's
                relation holds between the new EL and ELRK, the elevation
                the step started from
  ORE           types 9/10 (mod_obcs.F:706-737): the boundary value is a
                convex combination of its own elevation a stage back and
                the neighbour's stage input; after a step these are ELM2 of
                the node and ELM1 of NEXT_OBC
  radiation     a gravity wave that reaches a radiating boundary (5, 7, 9)
                leaves the basin, while a clamped boundary (3) reflects it:
                the energy left in the basin after the wave's round trip
                tells them apart (Blumberg and Kantha 1985)

The GPU's parity with the oracle on mixed types is in
tests/test_gpu_parity.py.
"""
import numpy as np
import pytest

import myfvcomcode_amd as F
from oracle import Oracle

GRAV = 9.81


def _channel(**kw):
    return F.Model.build(0, **kw)


def _adjacent_cells(m):
    nob = m.hdr.iobcn
    nadj = m.field("nadjc_obc")
    adj = m.field("adjc_obc").reshape(2, nob)
    return [[int(adj[j, k]) for j in range(nadj[k])] for k in range(nob)]


def test_set_obc_type_sets_iucp_by_parity(built):
    m = _channel()
    nob = m.hdr.iobcn
    assert nob > 8
    # the built model: every node type 1, its cells linear
    assert np.all(m.field("type_obc") == 1)
    cells = _adjacent_cells(m)
    iucp = m.field("iucp")
    lin = sorted({c for cs in cells for c in cs})
    assert np.all(iucp[lin] == 0)
    assert int((iucp[1:] == 0).sum()) == len(lin)
    # all even: nothing linear
    m.set_obc_type(np.full(nob, 2))
    assert np.all(m.field("iucp")[1:] == 1)
    # mixed: exactly the cells of odd-type nodes
    types = np.array([(k % 10) + 1 for k in range(nob)], np.int32)
    m.set_obc_type(types)
    want = np.ones(m.n + 1, np.int32)
    for k in range(nob):
        if types[k] % 2 == 1:
            want[cells[k]] = 0
    assert np.array_equal(m.field("iucp")[1:], want[1:])
    assert np.array_equal(m.field("type_obc"), types)


@pytest.mark.parametrize("bad", [0, 11])
def test_set_obc_type_refuses_unbuilt_types(built, bad):
    m = _channel()
    t = np.ones(m.hdr.iobcn, np.int32)
    t[3] = bad
    with pytest.raises(RuntimeError):
        m.set_obc_type(t)
    with pytest.raises(RuntimeError):
        m.set_obc_type(t[:-1] * 0 + 1)   # wrong length


def test_clamped_boundary_holds_zero(built):
    m = _channel()
    m.set_obc_type(np.full(m.hdr.iobcn, 3))
    o = Oracle(m)
    for _ in range(5):
        o.external_step(1)
    el = o.get("el")
    obc = m.field("i_obc_n")
    assert np.all(el[obc] == 0.0)
    assert np.abs(el).max() > 1e-4      # the interior moves (river inflow)


def _bump_channel(ty, dist):
    """The channel without tide, river, Coriolis or sponge, at rest under a
    Gaussian elevation bump of 0.1 m centred ``dist`` m inside the open
    end, with every OBC node of type ``ty``."""
    m = F.Model.build(0, river=0, coriolis=0.0, sponge_radius=0.0, p_obc_elevation_on=0, p_iramp=0)
    vx = m.field("vx")
    x0 = vx[1:].max() - dist
    el = m.field("el")
    el[1:] = 0.1 * np.exp(-((vx[1:] - x0) / 2000.0) ** 2)
    m.cold_start()
    m.set_obc_type(np.full(m.hdr.iobcn, ty))
    return m


@pytest.mark.parametrize("ty", [5, 6, 7, 8])
def test_radiation_relation_after_an_external_step(built, ty):
    m = _bump_channel(ty, 3000.0)
    o = Oracle(m)
    for _ in range(40):
        o.external_step(1)
    elrk = o.get("el").copy()
    o.external_step(1)
    el = o.get("el")
    obc, nxt = m.field("i_obc_n"), m.field("next_obc")
    h, dltn, grav = m.field("h"), m.field("dltn_obc"), m.field("grav_n")
    dte = m.hdr.p.dte
    # stage 4 (ALPHA_RK = 1): ELF(I1) = (CC*ELF(I2) + ELRK(I1)*r)/(CC+1)
    cc = np.sqrt(grav[obc] * h[obc]) * dte / dltn
    r = 1.0 if ty <= 6 else 1.0 - dte / 10800.0
    want = (cc * el[nxt] + elrk[obc] * r) / (cc + 1.0)
    np.testing.assert_allclose(el[obc], want, rtol=1e-13, atol=1e-15)
    # a plausible slip -- the neighbour's elevation from the step's start --
    # would not satisfy it
    slip = (cc * elrk[nxt] + elrk[obc] * r) / (cc + 1.0)
    assert np.abs(el[obc] - slip).max() > 1e-9


def _wave_energy(ty, steps):
    """The energy of the depth-mean flow left in the basin after ``steps``
    external steps from a bump 10 km inside the open end."""
    m = _bump_channel(ty, 10000.0)
    o = Oracle(m)

    def energy():
        e, d1, ua, va = o.get("el"), o.get("d1"), o.get("ua"), o.get("va")
        pot = 0.5 * GRAV * np.sum(m.field("art1")[1:] * e[1:] ** 2)
        kin = 0.5 * np.sum(m.field("art")[1:] * d1[1:] * (ua[1:] ** 2 + va[1:] ** 2))
        return pot + kin

    e0 = energy()
    for _ in range(steps):
        o.external_step(1)
    return e0, energy()


def test_radiating_boundary_lets_a_wave_leave(built):
    # ~15 m/s at the open end: the east-going half of the bump reaches the
    # boundary after ~700 s and, if reflected, is back inside by 1500 s
    steps = 320                                   # 1600 s at DTE = 5 s
    e0, e_clp = _wave_energy(3, steps)
    _, e_gwi = _wave_energy(5, steps)
    _, e_bki = _wave_energy(7, steps)
    _, e_ore = _wave_energy(9, steps)
    print("energy left: start %.4e  clamped %.4e  GWI %.4e  BKI %.4e  ORE %.4e" % (e0, e_clp, e_gwi, e_bki,
                                                                                e_ore))
    assert e_clp > 0.8 * e0          # reflected (little friction on this time scale)
    assert e_gwi < 0.75 * e_clp      # the east-going half (about half the energy) has left
    assert e_bki < 0.75 * e_clp
    assert e_ore < 0.75 * e_clp


@pytest.mark.parametrize("ty", [9, 10])
def test_orlanski_value_is_a_convex_combination(built, ty):
    m = _bump_channel(ty, 3000.0)
    assert np.all(m.field("next_obc2") > 0)
    o = Oracle(m)
    obc, nxt = m.field("i_obc_n"), m.field("next_obc")
    isonb = m.field("isonb")
    assert np.all(isonb[m.field("next_obc2")] != 2)
    moved = 0
    for _ in range(60):
        o.external_step(1)
        el, elm1, elm2 = o.get("el"), o.get("elm1"), o.get("elm2")
        a, b = elm2[obc], elm1[nxt]        # ELM1(I1) and EL(I2) of the last stage
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        tol = 1e-15 * np.maximum(1.0, np.abs(hi))
        assert np.all(el[obc] >= lo - tol) and np.all(el[obc] <= hi + tol)
        moved += int(np.any(el[obc] != 0.0))
    assert moved > 10                      # the boundary does move as the bump arrives


def test_tsobc_orlanski_carries_the_perturbation(built):
    """TYPE_TSOBC 4 (Orlanski, mod_obcs.F:1258-1285): T1M1/T1M2 are T1 one
    and two external RK stages back, set at every stage while T1 stays
    fixed, so by BCOND_TS both equal T1; the phase speed's numerator is then
    zero, MU = 0, and an outflow node keeps its own perturbation about the
    2-D upstream update of its column mean (inflow nodes keep T1)."""
    m = F.Model.build(0, nx=21, ny=7, kb=8, river=0, p_type_tsobc=4)
    kb, m1 = m.kb, m.m + 1
    o = Oracle(m)
    o.run(3)
    t1_before = o.get("t1").copy()
    s1_before = o.get("s1").copy()
    o.internal_step()
    for a, b in (("t1m1", t1_before), ("t1m2", t1_before), ("s1m1", s1_before), ("s1m2", s1_before)):
        assert np.array_equal(o.get(a), b), a
    # one BCOND_TS call on a state with distinct perturbations at the OBC nodes
    obc = m.field("i_obc_n")
    DZ = m.field("dz").reshape(kb, m1)
    art1 = m.field("art1")
    dti = m.hdr.p.dte * m.hdr.p.isplit
    rng = np.random.default_rng(11)
    t1 = o.get("t1").reshape(kb, m1).copy()
    for j in obc:
        t1[:-1, j] = t1[:-1, j] + 0.05 * rng.standard_normal(kb - 1)
    nob = m.hdr.iobcn
    uo = np.abs(rng.standard_normal(nob)) * 50.0
    uo[0] = -1.0
    xf = np.zeros((kb, nob))
    xf[:-1] = rng.standard_normal((kb - 1, nob)) * 3.0
    for nm, a in (("t1", t1.ravel()), ("t1m1", t1.ravel()), ("t1m2", t1.ravel()), ("uard_obcn", uo),
                  ("xflux_obc", xf.ravel())):
        o.put(nm, a)
    d, dt = o.get("d"), o.get("dt")
    o.routine("bcond_ts", 1)
    out = o.get("tf1").reshape(kb, m1)
    nbsn, off, ntsn = m.field("nbsn"), m.field("nbsn_off"), m.field("ntsn")
    checked = 0
    for q, j in enumerate(obc):
        if uo[q] <= 0:
            assert np.array_equal(out[:-1, j], t1[:-1, j])
            continue
        t2d = np.sum(t1[:-1, j] * DZ[:-1, j])
        mean = (t2d * dt[j] - (np.sum(xf[:-1, q]) + t2d * uo[q]) * dti / art1[j]) / d[j]
        exp = mean + (t1[:-1, j] - t2d)
        # the clip to the ring and vertical neighbours' range (fct-like) may act
        ring = nbsn[off[j]:off[j] + ntsn[j]]
        lo, hi = t1[:-1, ring].min(axis=1), t1[:-1, ring].max(axis=1)
        inside = (exp > lo) & (exp < hi)
        assert np.max(np.abs(out[:-1, j][inside] - exp[inside]), initial=0.0) < 1e-12
        checked += int(inside.sum())
    assert checked > 0
