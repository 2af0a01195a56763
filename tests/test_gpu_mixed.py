"""The mixed-precision form (precision "mixed"): FP32 control-volume and
edge fluxes (ADV_T/ADV_S/ADV_Q faces, ADVAVE/ADVECTION/ADV_UV edge terms)
with FP64 sums of those fluxes, FP64 state, FP64 elevation update
(EXTEL_EDGE), FP64 budgets and FP64 vertical solves.

Bounds (stated here, checked against the FP64 path, which reproduces the
oracle bit for bit), about ten times the RMS differences measured after one
simulated day of the 81x21 estuary channel (1728 internal steps: EL 1.1e-7
m, UA/VA 1.4e-6 / 1.1e-6 m/s, U/V 2.4e-6 / 1.7e-6 m/s, T 1.3e-12 degC, S
6.0e-6 psu): EL 1e-6 m, UA/VA 1.5e-5 m/s, U/V 3e-5 m/s, T 1e-10 degC,
S 6e-5 psu.  Volume, salt and heat stay conserved to FP64 round-off in a
closed basin, because every FP32 face flux enters both of its nodes with
opposite signs."""
import numpy as np
import pytest

from helpers import interior

pytestmark = pytest.mark.gpu

BOUNDS = {"el": 1e-6, "ua": 1.5e-5, "va": 1.5e-5, "u": 3e-5, "v": 3e-5, "t1": 1e-10, "s1": 6e-5}
WHEEL_BOUNDS = {"el": 3e-10, "ua": 3e-10, "va": 3e-10, "u": 1e-9, "v": 1e-9, "t1": 7e-8, "s1": 9e-8}


def rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def test_mixed_error_bounds(built):
    """One simulated day (1728 internal steps) of the estuary channel."""
    import myfvcomcode_amd as F
    m = F.Model.build(0)
    print("case: 81x21 estuary channel, 1728 steps (1 day), mixed vs fp64", flush=True)
    d64 = F.DeviceModel(m, precision="fp64")
    dmx = F.DeviceModel(m, precision="mixed")
    d64.run(1728, graph=True)
    dmx.run(1728, graph=True)
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
    register lists in FP32): closed-basin budgets to round-off, and after 30
    steps about ten times the differences measured there (EL 2.7e-11 m,
    UA/VA 2.9e-11 m/s, U/V 8.2e-11 m/s, T 6.5e-9 degC, S 8.7e-9 psu; its
    stratified T carries FP32 face rounding that the channel's uniform T
    does not)."""
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
    for name, bound in WHEEL_BOUNDS.items():
        a = interior(name, dmx.get(name), m)
        b = interior(name, d64.get(name), m)
        assert np.isfinite(a).all(), name
        rep[name] = rms(a, b)
    print(rep, flush=True)
    for name, bound in WHEEL_BOUNDS.items():
        assert rep[name] <= bound, (name, rep[name], bound)
