"""Parity of the HIP path with the FP64 oracle (oracle/fvcom_oracle.c) on
the same seeded inputs, through the C-ABI of include/fvgpu.h.

Tolerance: zero -- bit-identical results.  The GPU evaluates every flux
with the oracle's operation order, sums in the reference's scatter order
through sorted gather lists, and both sides use the same deterministic
tanh/cos/log/pow (include/fv_detmath.h) with floating-point contraction
off.  (The bound asked of the FP64 path is relative L2 1e-10 after N
steps; these tests hold it to 0.)
"""
import sys

import numpy as np
import pytest

from helpers import ELEM2, ELEM3, NODE2, NODE3, compare

pytestmark = pytest.mark.gpu


def _pair(config=0, **kw):
    import myfvcomcode_amd as F
    from oracle import Oracle
    m = F.Model.build(config, **kw)
    return m, F.DeviceModel(m), Oracle(m)


def test_external_step_parity(built):
    m, dev, orc = _pair(0)
    print("case: 81x21 estuary channel, one EXTERNAL_STEP", flush=True)
    for _ in range(3):
        dev.external_step(1)
        orc.external_step(1)
    rep = {}
    worst, bad = compare(dev, orc, m, ["el", "d", "ua", "va", "el1", "d1"], 0.0, rep)
    print(rep)
    assert not bad, bad


def test_internal_step_parity_one(built):
    m, dev, orc = _pair(0)
    print("case: 81x21 estuary channel, one INTERNAL_STEP", flush=True)
    dev.internal_step()
    orc.internal_step()
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    for k, v in sorted(rep.items(), key=lambda x: -x[1]):
        print("  %-8s %.3e" % (k, v))
    assert not bad, bad


def test_internal_step_parity_many(built):
    m, dev, orc = _pair(0)
    n = 30
    print("case: 81x21 estuary channel, %d INTERNAL_STEPs (graph)" % n, flush=True)
    dev.run(n, graph=True)
    orc.run(n)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    for k, v in sorted(rep.items(), key=lambda x: -x[1]):
        print("  %-8s %.3e" % (k, v))
    d = dev.diagnostics()
    assert d.nonfinite == 0 and d.depth_fail == 0
    assert d.iint == n
    assert not bad, bad


@pytest.mark.parametrize("graph", [False, True])
def test_open_boundary_types_parity(built, graph):
    """Every TYPE_OBC (1..10: elevation specified, clamped, gravity-wave,
    Blumberg-Kantha and Orlanski radiation, each linear and nonlinear)
    along one open boundary of the channel."""
    import myfvcomcode_amd as F
    from oracle import Oracle
    m = F.Model.build(0)
    types = [(k % 10) + 1 for k in range(m.hdr.iobcn)]
    m.set_obc_type(types)
    dev, orc = F.DeviceModel(m), Oracle(m)
    n = 10
    print("case: 81x21 estuary channel, TYPE_OBC 1..10 along the boundary, %d steps, graph=%s" % (n, graph),
          flush=True)
    dev.run(n, graph=graph)
    orc.run(n)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    for k, v in sorted(rep.items(), key=lambda x: -x[1])[:8]:
        print("  %-8s %.3e" % (k, v))
    d = dev.diagnostics()
    assert d.nonfinite == 0 and d.depth_fail == 0
    assert not bad, bad


def test_tsobc_orlanski_parity(built):
    """TYPE_TSOBC 4 (Orlanski radiation of the T/S perturbation) beside the
    ten elevation types: the device takes its T1M1/T1M2 as T1 (their value
    by BCOND_TS), the oracle keeps and shifts the arrays at every stage."""
    import myfvcomcode_amd as F
    from oracle import Oracle
    m = F.Model.build(0, p_type_tsobc=4)
    m.set_obc_type([(k % 10) + 1 for k in range(m.hdr.iobcn)])
    dev, orc = F.DeviceModel(m), Oracle(m)
    n = 10
    print("case: 81x21 estuary channel, TYPE_TSOBC 4, TYPE_OBC 1..10, %d steps" % n, flush=True)
    dev.run(n, graph=True)
    orc.run(n)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    for k, v in sorted(rep.items(), key=lambda x: -x[1])[:8]:
        print("  %-8s %.3e" % (k, v))
    assert not bad, bad


@pytest.mark.parametrize("graph", [False, True])
def test_jittered_wind_case_parity(built, graph):
    """Config-1 generator at 101x51 nodes, KB=11: jitter, seamounts, wind,
    M2/S2/K1; eager and graph-replayed steps both match the oracle."""
    m, dev, orc = _pair(1, nx=101, ny=51, kb=11)
    n = 12
    print("case: jittered seamount mesh 101x51 KB=11, %d steps, graph=%s" % (n, graph), flush=True)
    dev.run(n, graph=graph)
    orc.run(n)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    for k, v in sorted(rep.items(), key=lambda x: -x[1])[:8]:
        print("  %-8s %.3e" % (k, v))
    assert not bad, bad


def test_fortran_driver_matches_library(built):
    """The Fortran host (ISO_C_BINDING module + driver) advances the estuary
    channel exactly as the library called directly: its REPORT lines of
    volume, salt and minimum depth (printed to 17 significant digits) equal
    fvgpu_diagnostics of the same run."""
    import os
    import subprocess
    import myfvcomcode_amd as F
    exe = os.path.join(os.path.dirname(F.__file__), "lib", "fvcom_gpu")
    if not os.path.exists(exe):
        pytest.skip("the Fortran driver is not built (no amdflang)")
    print("case: 81x21 estuary channel, 6 INTERNAL_STEPs through the Fortran driver", flush=True)
    r = subprocess.run([exe, "0", "6", "3"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    rows = {}
    for line in r.stdout.splitlines():
        if "IINT=" in line:
            t = line.replace("=", " ").split()
            rows[int(t[1])] = (float(t[3]), float(t[5]), float(t[7]))
    assert sorted(rows) == [3, 6]
    dev = F.DeviceModel(F.Model.build(0))
    for it in (3, 6):
        dev.run(3, graph=False)
        d = dev.diagnostics()
        assert d.iint == it
        for got, want in zip(rows[it], (d.volume, d.salt, d.min_depth)):
            assert abs(got - want) <= 1e-16 * abs(want), (it, got, want)


@pytest.mark.parametrize("graph", [False, True])
def test_state_written_between_steps(built, graph):
    """A 2-D state written from the host between internal steps (UA, VA and
    EL perturbed, as a restart or data assimilation would) is what the next
    step's external mode uses: the device, which keeps derived per-sub-edge
    terms of the 2-D state across RK stages, still matches the oracle bit
    for bit."""
    m, dev, orc = _pair(0)
    print("case: 81x21 estuary channel, 2 steps, UA/VA/EL rewritten, 2 more (graph=%s)" % graph, flush=True)
    dev.run(2, graph=graph)
    orc.run(2)
    rng = np.random.default_rng(11)
    for name, scale in (("ua", 0.05), ("va", 0.05), ("el", 0.02)):
        x = orc.get(name)
        x = x + scale * rng.standard_normal(x.shape)
        x[0] = orc.get(name)[0]
        orc.put(name, x)
        dev.put(name, x)
    dev.run(2, graph=graph)
    orc.run(2)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    assert not bad, bad


def test_forcing_changed_between_steps(built):
    """Wind speed changed between internal steps through fvgpu_set_forcing
    advances exactly as the oracle reading the changed wind: the new value
    is the wind at the end of the coming step, the previous one its start,
    and each external step takes the wind interpolated to its ExtTime."""
    import myfvcomcode_amd as F
    from oracle import Oracle
    m = F.Model.build(1, nx=61, ny=31, kb=9, n_seamounts=2)
    assert m.hdr.p.wind_on
    dev, orc = F.DeviceModel(m), Oracle(m)
    print("case: jittered mesh 61x31 KB=9 with wind, 2 steps, wind changed, 2 more", flush=True)
    dev.run(2, graph=True)
    orc.run(2)
    uu, vv = m.field("uuwind"), m.field("vvwind")
    uu0, vv0 = m.field("uuwind0"), m.field("vvwind0")
    rng = np.random.default_rng(5)
    for step in range(3):
        uu0[:] = uu                                                # the oracle reads the model's arrays
        vv0[:] = vv
        uu[1:] = 1.5 * uu[1:] + rng.standard_normal(uu.size - 1)
        vv[1:] = 0.5 * vv[1:] - 3.0
        dev.set_forcing("wind", np.concatenate([uu, vv]))
        dev.run(1, graph=step > 0)
        orc.run(1)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    assert not bad, bad


def test_river_forcing_changed_between_steps(built):
    """River discharge, temperature and salinity changed between internal
    steps through fvgpu_set_forcing advance exactly as in the oracle."""
    import myfvcomcode_amd as F
    from oracle import Oracle
    m = F.Model.build(0)
    q, t, s = m.field("qdis0"), m.field("tdis"), m.field("sdis")
    assert q.size >= 1
    dev, orc = F.DeviceModel(m), Oracle(m)
    print("case: 81x21 estuary channel, 2 steps, river changed, 3 more", flush=True)
    dev.run(2, graph=False)
    orc.run(2)
    q[:] = 3.0 * q
    t[:] = t + 4.0
    s[:] = s + 1.5
    for name, v in (("qdis", q), ("tdis", t), ("sdis", s)):
        dev.set_forcing(name, v)
    dev.run(3, graph=False)
    orc.run(3)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    assert not bad, bad
    with pytest.raises(RuntimeError):
        dev.set_forcing("wind", np.zeros(2 * (m.n + 1)))    # the channel has no wind


def test_high_degree_node_parity(built):
    """A node with 12 triangles around it (the reference's limit): the
    widest register lists (rings of 16, sub-edge lists of 24) against the
    oracle, eager and as one graph per step."""
    import myfvcomcode_amd as F
    from oracle import Oracle
    from helpers import wheel_mesh
    m = wheel_mesh()
    for graph in (False, True):
        print("case: wheel mesh (node of degree 12), 10 INTERNAL_STEPs, graph=%s" % graph, flush=True)
        dev, orc = F.DeviceModel(m), Oracle(m)
        dev.run(10, graph=graph)
        orc.run(10)
        rep = {}
        worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
        assert not bad, bad
        del dev


@pytest.mark.parametrize("kb", [4, 9, 17, 21, 33, 41, 49, 64])
def test_layer_count_extremes_parity(built, kb):
    """Every register-array bound of the column kernels against the oracle:
    KB = 4 (two interior Q levels), 9, 17, 21 (the ensemble member's 20
    layers), 33, 41 (the 1M-element case's 40 layers), 49, and KB = 64 (the
    device's limit: the largest column register arrays)."""
    m, dev, orc = _pair(0, nx=21, ny=7, kb=kb)
    print("case: 21x7 estuary channel, KB=%d, 3 INTERNAL_STEPs (graph)" % kb, flush=True)
    dev.run(3, graph=True)
    orc.run(3)
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0)
    assert not bad, bad


@pytest.mark.parametrize("kb", [21, 41])
def test_jittered_mesh_real_layer_counts(built, kb):
    """The seamount generator (jitter, seamounts, wind, M2/S2/K1) at
    101x51 nodes with the layer counts the benchmark and the ensemble member
    run: KB = 41 (40 layers, the KBM1 = 40 column kernels) and KB = 21."""
    m, dev, orc = _pair(1, nx=101, ny=51, kb=kb)
    n = 4
    print("case: jittered seamount mesh 101x51 KB=%d, %d steps (graph)" % (kb, n), flush=True)
    dev.run(n, graph=True)
    orc.run(n)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    for k, v in sorted(rep.items(), key=lambda x: -x[1])[:6]:
        print("  %-8s %.3e" % (k, v))
    assert not bad, bad


@pytest.mark.parametrize("graph", [False, True])
def test_calculated_river_parity(built, graph):
    """RIVER_TS_SETTING 'calculated': the river's T/S enter through the
    advective flux and ADJUST_TS raises the river node to them after the
    tracer passes (adjust_ts.F); a warm river into the channel."""
    import myfvcomcode_amd as F
    from oracle import Oracle
    m = F.Model.build(0, river_t=24.0, river_s=3.0)
    m.hdr.p.river_ts_specified = 0
    dev, orc = F.DeviceModel(m), Oracle(m)
    n = 6
    print("case: 81x21 estuary channel, calculated river T/S, %d steps, graph=%s" % (n, graph), flush=True)
    dev.run(n, graph=graph)
    orc.run(n)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    assert not bad, bad
    j = int(m.field("inodeq")[0])
    t1 = dev.get("t1").reshape(m.kb, -1)
    assert np.all(t1[:-1, j] >= 24.0)


_SWITCHES = [
    (0, {"p_horizontal_closure": 0}),          # constant horizontal viscosity (FM1 branch of every edge routine)
    (0, {"p_vertical_closure": 0}),            # no MY2.5: KM = KH = UMOL, no ADV_Q/VDIF_Q
    (0, {"p_scalar_positivity": 0}),           # FCT off in ADV_T/ADV_S and ADV_Q
    (0, {"p_noflux_bot": 0}),
    (0, {"p_obc_elevation_on": 0}),            # no tide at the open boundary
    (0, {"p_horizontal_closure": 0, "p_vertical_closure": 0, "p_scalar_positivity": 0}),
    (1, {"p_temperature_active": 0}),          # T frozen, S advected (one tracer pass)
    (1, {"p_salinity_active": 0}),
    (1, {"p_barotropic": 1}),                  # no BAROPG density term, no DENS2
    (1, {"p_type_tsobc": 1}),                  # T/S at outflow OBC nodes: upstream, no radiation
    (1, {"p_type_tsobc": 2}),                  # radiation without relaxation
    (1, {"p_obc_on": 0}),                      # BCOND_TS leaves the OBC nodes' T/S to advection
    (0, {"p_hprnu": 2.5, "p_vprnu": 0.7, "p_umol": 3e-4}),   # Prandtl numbers off 1: KH != KM
    (0, {"p_hprnu": 2.5, "p_vprnu": 0.7, "p_umol": 3e-4, "p_vertical_closure": 0}),
    (0, {"p_cbcmin": 0.01}),                   # the drag coefficient's floor
]


@pytest.mark.parametrize("config,switches", _SWITCHES,
                         ids=["-".join("%s=%g" % (k[2:], v) for k, v in s.items()) for _, s in _SWITCHES])
def test_physics_switches_parity(built, config, switches):
    """The step's run-time switches (the namelist's HORIZONTAL_MIXING_TYPE,
    VERTICAL_MIXING_TYPE, SCALAR_POSITIVITY_CONTROL, NOFLUX_BOT_CONDITION,
    OBC_ELEVATION_FORCING_ON, TEMPERATURE_ACTIVE, SALINITY_ACTIVE,
    BAROTROPIC, TYPE_TSOBC, OBC_ON) and parameters away from their defaults (HPRNU,
    VPRNU, UMOL, CBCMIN): each changes the state against the default on
    these cases within six steps, and the device follows the oracle bit for
    bit."""
    import myfvcomcode_amd as F
    from oracle import Oracle
    kw = {} if config == 0 else {"nx": 41, "ny": 21, "kb": 9, "n_seamounts": 2}
    m = F.Model.build(config, **kw, **switches)
    dev, orc = F.DeviceModel(m), Oracle(m)
    n = 6
    print("case: config %d %s, %s, %d steps (graph)" % (config, kw or "81x21 channel", switches, n), flush=True)
    dev.run(n, graph=True)
    orc.run(n)
    rep = {}
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    d = dev.diagnostics()
    assert d.nonfinite == 0 and d.depth_fail == 0
    assert not bad, bad


def test_nonfinite_state_is_reported(built):
    """A NaN written into UA at one element: the next internal step spreads
    it and fvgpu_diagnostics reports a non-finite state, as the oracle's
    state shows; the run itself does not stop (the reference checks after
    the step, SHUTDOWN_CHECK)."""
    m, dev, orc = _pair(0, nx=21, ny=7, kb=6)
    print("case: 21x7 channel, NaN put into UA, 1 INTERNAL_STEP", flush=True)
    assert dev.diagnostics().nonfinite == 0
    ua = dev.get("ua")
    ua[m.n // 2] = np.nan
    dev.put("ua", ua)
    orc.put("ua", ua)
    dev.run(1, graph=False)
    orc.run(1)
    d = dev.diagnostics()
    assert d.nonfinite == 1
    assert not np.all(np.isfinite(orc.get("ua")))
    # the NaN reached the same entries on both sides
    for name in ("ua", "el", "u", "t1"):
        a, b = dev.get(name), orc.get(name)
        assert np.array_equal(np.isnan(a), np.isnan(b)), name


def test_depth_below_min_depth_is_reported(built):
    """An elevation that leaves a node's depth below MIN_DEPTH: DEPTH_CHECK
    (depth_check.F) flags it in the external step, on the device as in the
    oracle."""
    m, dev, orc = _pair(0, nx=21, ny=7, kb=6)
    print("case: 21x7 channel, EL set to leave D < MIN_DEPTH at one node, 1 INTERNAL_STEP", flush=True)
    assert dev.diagnostics().depth_fail == 0
    h = m.field("h")
    j = m.m // 2
    el = dev.get("el")
    el[j] = -h[j] + 0.2 * m.hdr.p.min_depth
    dev.put("el", el)
    orc.put("el", el)
    dev.run(1, graph=False)
    orc.run(1)
    assert dev.diagnostics().depth_fail == 1
    assert orc.depth_fail
