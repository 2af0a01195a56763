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
    advances exactly as the oracle reading the changed wind."""
    import myfvcomcode_amd as F
    from oracle import Oracle
    m = F.Model.build(1, nx=61, ny=31, kb=9, n_seamounts=2)
    assert m.hdr.p.wind_on
    dev, orc = F.DeviceModel(m), Oracle(m)
    print("case: jittered mesh 61x31 KB=9 with wind, 2 steps, wind changed, 2 more", flush=True)
    dev.run(2, graph=True)
    orc.run(2)
    uu, vv = m.field("uuwind"), m.field("vvwind")
    rng = np.random.default_rng(5)
    uu[1:] = 1.5 * uu[1:] + rng.standard_normal(uu.size - 1)    # the oracle reads the model's arrays
    vv[1:] = 0.5 * vv[1:] - 3.0
    dev.set_forcing("wind", np.concatenate([uu, vv]))
    dev.run(2, graph=True)
    orc.run(2)
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


@pytest.mark.parametrize("kb", [4, 64])
def test_layer_count_extremes_parity(built, kb):
    """KB = 4 (two interior Q levels) and KB = 64 (the device's limit: the
    largest column register arrays) against the oracle."""
    m, dev, orc = _pair(0, nx=21, ny=7, kb=kb)
    print("case: 21x7 estuary channel, KB=%d, 3 INTERNAL_STEPs (graph)" % kb, flush=True)
    dev.run(3, graph=True)
    orc.run(3)
    worst, bad = compare(dev, orc, m, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0)
    assert not bad, bad
