"""Pinning of the FP64 oracle (oracle/fvcom_oracle.c) by closed forms and
invariants.  The reference ships no outputs or fixtures for this path
(its FS/testing programs cover NetCDF, MPI and interpolation only), so each
routine is checked against what the mathematics fixes:

  EXTEL_EDGE        volume budget = river inflow (flux divergence telescopes)
  EXTEL/EXTUV/PG    lake at rest stays exactly at rest; seiche period 2L/sqrt(gH)
  BCOND_ASL         OBC elevation = amp*cos(w t - phi)*RAMP after a full RK step
  ADV_T/ADV_S/FCT   constant tracer stays constant; min/max principle
  ADV_S/VDIF_TS     salt content conserved in a closed basin
  DENS2             published sigma-t values of the UNESCO-type polynomial
  fv_detmath        deterministic tanh/cos/log within a few ulp of libm

Single routines (VISCOF_H, WREAL, BCOND_GCN, VDIF_Q, BCOND_TS, ADJUST2D3D,
BOTTOM_ROUGHNESS, BAROPG, VDIF_UV) are pinned one at a time in
test_oracle_pins.py.
"""
import ctypes as C
import math

import numpy as np
import pytest

import myfvcomcode_amd as F
from oracle import Oracle, lib as oracle_lib


def closed_basin(nx=21, ny=7, kb=6, h=None, **pkw):
    """The 81x21 estuary channel mesh without open-boundary nodes."""
    c = F.Model.build(0, nx=nx, ny=ny, kb=kb, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    hh = c.field("h")[1:].copy() if h is None else np.full(c.m, float(h))
    p = F.Model.case_defaults(0).p
    for k, v in pkw.items():
        setattr(p, k, v)
    m = F.Model.from_mesh(x, y, nv, hh, kb, [], p)
    m.field3("t1")[:-1, 1:] = 15.0
    m.field3("s1")[:-1, 1:] = 30.0
    m.cold_start()
    return m, nx, ny


def test_lake_at_rest_stays_at_rest(built):
    """No tide, no river, no wind, uniform T/S: every velocity stays exactly 0."""
    m = F.Model.build(0, nx=21, ny=7, kb=6, river=0, p_obc_elevation_on=0)
    s1 = m.field("s1")
    s1[:] = np.where(s1 != 0, 30.0, 0.0)
    m.cold_start()
    o = Oracle(m)
    o.internal_step()
    # the first step is exact: every flux, pressure gradient and forcing is 0
    for name in ("ua", "va", "u", "v", "w", "el"):
        assert np.all(o.get(name) == 0.0), name
    # later steps: the implicit vertical diffusion of a uniform column returns
    # it to within an ulp, so density and velocities stay at round-off level
    o.run(5)
    for name in ("ua", "va", "u", "v"):
        assert np.abs(o.get(name)).max() < 1e-14, name
    assert np.abs(o.get("el")).max() < 1e-14


def test_constant_tracer_stays_constant(built):
    """Uniform T with tidal flow: T1 stays 15 to round-off."""
    m = F.Model.build(0, nx=21, ny=7, kb=6, river=0)
    o = Oracle(m)
    o.run(20)
    t1 = o.get("t1").reshape(m.kb, -1)[:-1, 1:]
    assert np.abs(t1 - 15.0).max() < 1e-11
    assert np.abs(o.get("ua")).max() > 1e-4   # there is flow


def test_volume_budget_equals_river_inflow(built):
    """Closed basin with a node river: sum(ART1*D) grows by DTE*QDIS per
    external step with QDIS = Q*RAMP(IINT); the CV flux divergence
    telescopes to zero (extel_edge.F:86-112,160-180)."""
    q = 80.0
    m, nx, ny = closed_basin()
    m.set_river((ny // 2) * nx + 1, q, 15.0, 0.0)
    m.cold_start()
    art1 = m.field("art1")[1:]
    o = Oracle(m)
    v0 = (art1 * o.get("d")[1:]).sum()
    p = m.hdr.p
    expect = 0.0
    for iint in range(1, 6):
        o.internal_step()
        expect += p.isplit * p.dte * q * math.tanh(iint / p.iramp)
        v = (art1 * o.get("d")[1:]).sum()
        assert v - v0 == pytest.approx(expect, rel=1e-9)


def test_closed_basin_volume_constant(built):
    m, _, _ = closed_basin()
    vx = m.field("vx")
    m.field("el")[1:] = 0.1 * (vx[1:] / vx.max() - 0.5)
    m.cold_start()
    o = Oracle(m)
    art1 = m.field("art1")[1:]
    v0 = (art1 * o.get("d")[1:]).sum()
    o.run(6)
    v = (art1 * o.get("d")[1:]).sum()
    assert abs(v - v0) <= 1e-13 * v0
    assert np.abs(o.get("ua")).max() > 1e-3


def test_obc_elevation_closed_form(built):
    """After one external step the OBC nodes carry the spectral tide at
    ExtTime times RAMP (BCOND_ASL + RK blend with ALPHA_RK(4)=1)."""
    m = F.Model.build(0, nx=21, ny=7, kb=6, river=0)
    o = Oracle(m)
    o.external_step(1)
    el = o.get("el")
    obc = m.field("i_obc_n")
    dte, iramp = m.hdr.p.dte, m.hdr.p.iramp
    t = 1 * dte                       # ExtTime after IEXT=1 of IINT=1
    ramp = math.tanh((0 + 1 / m.hdr.p.isplit) / iramp)
    period = m.field("period")[0]
    amp = m.field("apt")[0]
    expect = amp * math.cos(t * 2 * math.pi / period) * ramp
    assert np.allclose(el[obc], expect, rtol=1e-13, atol=1e-16)


def test_salt_conserved_closed_basin(built):
    """Closed basin, no FCT clipping, no bottom flux: sum(ART1*D*DZ*S1) is
    conserved to round-off (adv_s.F update + VDIF_TS column conservation)."""
    m, _, _ = closed_basin(scalar_positivity=0, noflux_bot=0)
    vx = m.field("vx")
    m.field("el")[1:] = 0.1 * (vx[1:] / vx.max() - 0.5)
    m.field3("s1")[:-1, 1:] = 30.0 * vx[1:] / vx.max()
    m.cold_start()
    o = Oracle(m)
    _, s0, h0 = o.budgets()
    o.run(10)
    _, s10, h10 = o.budgets()
    assert np.abs(o.get("ua")).max() > 1e-3
    assert abs(s10 - s0) <= 1e-12 * abs(s0)
    assert abs(h10 - h0) <= 1e-12 * abs(h0)


def test_seiche_period(built):
    """Rectangular basin of uniform depth with a cosine tilt: the west-end
    elevation first crosses zero a quarter period after release,
    T/4 = L/(2 sqrt(gH)) (EXTEL + EXTUV + barotropic PG)."""
    m, nx, ny = closed_basin(nx=41, ny=5, kb=4, h=10.0, iramp=0, cbcmin=0.0)
    m.field("cc_z0b")[1:] = 1e-12      # negligible bottom drag
    m.field("cor")[:] = 0.0
    vx = m.field("vx")
    L = vx.max()
    m.field("el")[1:] = 0.05 * np.cos(np.pi * vx[1:] / L)
    m.cold_start()
    o = Oracle(m)
    west = int(np.argmin(vx[1:])) + 1
    quarter = L / (2.0 * math.sqrt(9.81 * 10.0))
    dti = m.hdr.p.dte * m.hdr.p.isplit
    prev = o.get("el")[west]
    t_cross = None
    for s in range(1, int(3 * quarter / dti) + 2):
        o.internal_step()
        cur = o.get("el")[west]
        if prev > 0 and cur <= 0:
            t_cross = (s - 1 + prev / (prev - cur)) * dti
            break
        prev = cur
    assert t_cross is not None
    assert abs(t_cross - quarter) < 0.03 * quarter


def test_dens2_published_values(built):
    """DENS2 (eqs_of_state.F:290-314) gives sigma-t/1000; compare with the
    published UNESCO sigma-t values at atmospheric pressure."""
    m = F.Model.build(0, nx=5, ny=5, kb=4, river=0)
    s1 = m.field3("s1")
    t1 = m.field3("t1")
    # UNESCO (1981) densities at zero pressure: 1025.97, 1023.34, 999.97, 1028.11
    pts = [(35.0, 15.0, 25.97), (35.0, 25.0, 23.34), (0.0, 4.0, -0.03), (35.0, 0.0, 28.11)]
    for (s, t, sig) in pts:
        s1[:-1, 1:] = s
        t1[:-1, 1:] = t
        m.cold_start()
        rho = m.field3("rho1")[0, 1]
        assert rho * 1000.0 == pytest.approx(sig, abs=0.08), (s, t)


def _f(name):
    fn = getattr(oracle_lib(), name)
    fn.restype = C.c_double
    fn.argtypes = [C.c_double]
    return fn


@pytest.mark.parametrize("name,ref,lo,hi,maxulp", [
    ("orc_fv_cos", math.cos, -200.0, 2e5, 3),
    ("orc_fv_tanh", math.tanh, 0.0, 25.0, 4),
    ("orc_fv_log", math.log, 1e-8, 1e8, 3),
    ("orc_fv_expm1", math.expm1, -40.0, 40.0, 3),
])
def test_detmath_accuracy(built, name, ref, lo, hi, maxulp):
    fn = _f(name)
    rng = np.random.default_rng(7)
    xs = rng.uniform(lo, hi, 4000)
    if name == "orc_fv_log":
        xs = np.exp(rng.uniform(math.log(lo), math.log(hi), 4000))
    worst = 0.0
    for x in xs:
        a, b = fn(float(x)), ref(float(x))
        if b != 0:
            worst = max(worst, abs(a - b) / math.ulp(b))
    assert worst <= maxulp


@pytest.mark.parametrize("kb", [4, 64])
def test_oracle_layer_count_extremes(built, kb):
    """The fewest levels the grid accepts (KB = 4) and the most the device
    takes (KB = 64): the oracle steps both (its column work arrays hold
    levels 0..KB; an AddressSanitizer build of it runs these cases clean)."""
    m = F.Model.build(0, nx=21, ny=7, kb=kb, river=0)
    o = Oracle(m)
    o.run(3)
    for name in ("q2", "q2l", "km", "t1", "s1", "u", "el"):
        assert np.all(np.isfinite(o.get(name))), name
    assert np.abs(o.get("ua")).max() > 1e-4


def test_oracle_rejects_kb_beyond_limit(built):
    m = F.Model.build(0, nx=21, ny=7, kb=65)
    with pytest.raises(ValueError):
        Oracle(m)
