"""The mixed-precision form (precision "mixed"): FP32 control-volume and
edge fluxes (ADV_T/ADV_S/ADV_Q faces, ADVAVE/ADVECTION/ADV_UV edge terms)
with FP64 sums of those fluxes, FP64 state, FP64 elevation update
(EXTEL_EDGE), FP64 budgets and FP64 vertical solves.

Bounds (stated here, checked against the FP64 path, which reproduces the
oracle bit for bit): after 60 internal steps of the 81x21 estuary channel
(50 simulated minutes) the RMS differences stay below 1e-5 m for EL,
1e-5 m/s for UA/VA, 1e-4 m/s for U/V and 1e-4 (degC, psu) for T/S.  Volume,
salt and heat stay conserved to FP64 round-off in a closed basin, because
every FP32 face flux enters both of its nodes with opposite signs."""
import numpy as np
import pytest

from helpers import interior

pytestmark = pytest.mark.gpu

BOUNDS = {"el": 1e-5, "ua": 1e-5, "va": 1e-5, "u": 1e-4, "v": 1e-4, "t1": 1e-4, "s1": 1e-4}


def rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def test_mixed_error_bounds(built):
    import myfvcomcode_amd as F
    m = F.Model.build(0)
    print("case: 81x21 estuary channel, 60 steps, mixed vs fp64", flush=True)
    d64 = F.DeviceModel(m, precision="fp64")
    dmx = F.DeviceModel(m, precision="mixed")
    d64.run(60, graph=True)
    dmx.run(60, graph=True)
    rep = {}
    for name, bound in BOUNDS.items():
        a = interior(name, dmx.get(name), m)
        b = interior(name, d64.get(name), m)
        assert np.isfinite(a).all(), name
        rep[name] = rms(a, b)
    print(rep, flush=True)
    assert any(v > 0 for v in rep.values()), "the mixed form should differ from FP64"
    for name, bound in BOUNDS.items():
        assert rep[name] <= bound, (name, rep[name], bound)


def test_mixed_conserves_closed_basin(built):
    import myfvcomcode_amd as F
    from test_oracle import closed_basin
    m, _, _ = closed_basin(scalar_positivity=0, noflux_bot=0)
    vx = m.field("vx")
    m.field("el")[1:] = 0.1 * (vx[1:] / vx.max() - 0.5)
    m.field3("s1")[:-1, 1:] = 30.0 * vx[1:] / vx.max()
    m.cold_start()
    print("case: closed basin, 10 steps, mixed budgets", flush=True)
    dev = F.DeviceModel(m, precision="mixed")
    d0 = dev.diagnostics()
    dev.run(10, graph=False)
    d1 = dev.diagnostics()
    assert np.abs(dev.get("ua")).max() > 1e-3
    assert abs(d1.volume - d0.volume) <= 1e-13 * d0.volume
    assert abs(d1.salt - d0.salt) <= 1e-12 * abs(d0.salt)
    assert abs(d1.heat - d0.heat) <= 1e-12 * abs(d0.heat)


def test_mixed_high_degree_node(built):
    """The mixed form on the mesh with a node of degree 12 (the widest
    register lists in FP32): closed-basin budgets to round-off and the same
    bounds against FP64 as above after 30 steps."""
    import myfvcomcode_amd as F
    from helpers import wheel_mesh
    m = wheel_mesh(scalar_positivity=0, noflux_bot=0)
    print("case: wheel mesh (node of degree 12), 30 steps, mixed vs fp64 and budgets", flush=True)
    d64 = F.DeviceModel(m, precision="fp64")
    dmx = F.DeviceModel(m, precision="mixed")
    b0 = dmx.diagnostics()
    d64.run(30, graph=True)
    dmx.run(30, graph=False)
    b1 = dmx.diagnostics()
    assert abs(b1.volume - b0.volume) <= 1e-13 * b0.volume
    assert abs(b1.salt - b0.salt) <= 1e-12 * abs(b0.salt)
    assert abs(b1.heat - b0.heat) <= 1e-12 * abs(b0.heat)
    rep = {}
    for name, bound in BOUNDS.items():
        a = interior(name, dmx.get(name), m)
        b = interior(name, d64.get(name), m)
        assert np.isfinite(a).all(), name
        rep[name] = rms(a, b)
    print(rep, flush=True)
    for name, bound in BOUNDS.items():
        assert rep[name] <= bound, (name, rep[name], bound)
