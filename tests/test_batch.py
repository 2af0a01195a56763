"""Batched instances as one mesh (fv_model_union): every instance of the
disjoint union advances exactly as it does alone.  The CPU test runs the
oracle on the union and on each part; the GPU test runs the HIP path on the
union and compares it with the oracle bit for bit."""
import numpy as np
import pytest

import myfvcomcode_amd as F

FIELDS = ("el", "d", "ua", "va", "u", "v", "w", "t1", "s1", "q2", "q2l", "km", "kh", "wts")


def parts():
    # three instances of the ensemble-member generator at different sizes and
    # seeds (few seamounts, so that the small domains stay shallow and stable)
    return [F.Model.build(3, nx=21, ny=11, kb=6, seed=11, n_seamounts=1),
            F.Model.build(3, nx=31, ny=13, kb=6, seed=12, n_seamounts=1),
            F.Model.build(3, nx=17, ny=9, kb=6, seed=13, n_seamounts=0)]


def instance_slice(a, union, off, kb):
    om, on, m, n = off
    if a.size == union.m + 1:
        return a[om + 1: om + m + 1]
    if a.size == union.n + 1:
        return a[on + 1: on + n + 1]
    if a.size == (union.m + 1) * kb:
        return a.reshape(kb, -1)[:, om + 1: om + m + 1]
    return a.reshape(kb, -1)[:, on + 1: on + n + 1]


def own(a, part, kb):
    if a.size in (part.m + 1, part.n + 1):
        return a[1:]
    return a.reshape(kb, -1)[:, 1:]


def test_union_sizes(built):
    ps = parts()
    u, offs = F.Model.union(ps)
    assert u.m == sum(p.m for p in ps) and u.n == sum(p.n for p in ps)
    assert u.hdr.ne == sum(p.hdr.ne for p in ps) and u.hdr.ncv == sum(p.hdr.ncv for p in ps)
    assert u.hdr.iobcn == sum(p.hdr.iobcn for p in ps)
    # renumbered topology: every triangle of instance p uses instance p's nodes
    nv = u.field("nv").reshape(3, -1)
    for (om, on, m, n) in offs:
        t = nv[:, on + 1: on + n + 1]
        assert t.min() == om + 1 and t.max() == om + m


def test_union_rejects_mismatched_parts(built):
    a = F.Model.build(3, nx=17, ny=9, kb=6, n_seamounts=0)
    b = F.Model.build(3, nx=17, ny=9, kb=7, n_seamounts=0)
    with pytest.raises(RuntimeError):
        F.Model.union([a, b])


def test_union_instances_evolve_alone_oracle(built):
    from oracle import Oracle
    ps = parts()
    u, offs = F.Model.union(ps)
    ou = Oracle(u)
    ou.run(3)
    for p, off in zip(ps, offs):
        op = Oracle(p)
        op.run(3)
        for name in FIELDS:
            a = instance_slice(ou.get(name), u, off, u.kb)
            b = own(op.get(name), p, u.kb)
            assert np.isfinite(b).all(), name
            assert np.array_equal(a, b), (name, off)


@pytest.mark.gpu
def test_union_gpu_parity(built):
    from oracle import Oracle
    ps = parts()
    u, offs = F.Model.union(ps)
    print("case: union of 3 ensemble-generator instances, 4 steps (graph)", flush=True)
    dev = F.DeviceModel(u)
    dev.run(4, graph=True)
    orc = Oracle(u)
    orc.run(4)
    for name in FIELDS:
        a, b = dev.get(name), orc.get(name)
        assert np.array_equal(a, b), (name, float(np.abs(a - b).max()))


@pytest.mark.gpu
def test_mixed_union_member_shape(built):
    """The mixed-precision form on a union of ensemble-member instances at
    the member's 20 layers (KB = 21) plus two smaller instances of the same
    generator: every instance of the union advances bit for bit as its own
    mixed-precision run, and stays within the mixed form's bounds of the
    FP64 union."""
    ps = [F.Model.build(3, nx=41, ny=21, kb=21, seed=21, n_seamounts=1),
          F.Model.build(3, nx=41, ny=21, kb=21, seed=22, n_seamounts=1),
          F.Model.build(3, nx=25, ny=13, kb=21, seed=23, n_seamounts=0),
          F.Model.build(3, nx=17, ny=9, kb=21, seed=24, n_seamounts=0)]
    u, offs = F.Model.union(ps)
    n = 6
    print("case: mixed union of 4 member-generator instances KB=21, %d steps (graph)" % n, flush=True)
    dmx = F.DeviceModel(u, precision="mixed")
    dmx.run(n, graph=True)
    d64 = F.DeviceModel(u, precision="fp64")
    d64.run(n, graph=True)
    assert not dmx.diagnostics().nonfinite
    for p, off in zip(ps, offs):
        alone = F.DeviceModel(p, precision="mixed")
        alone.run(n, graph=True)
        for name in FIELDS:
            a = instance_slice(dmx.get(name), u, off, u.kb)
            b = own(alone.get(name), p, u.kb)
            assert np.array_equal(a, b), (name, off)
        del alone
    for name, bound in (("el", 1e-6), ("ua", 1e-5), ("u", 1e-4), ("t1", 1e-4), ("s1", 1e-4)):
        a, b = dmx.get(name), d64.get(name)
        err = float(np.sqrt(np.mean((a - b) ** 2)))
        print("  mixed vs fp64 %-3s rms %.3e" % (name, err), flush=True)
        assert err <= bound, (name, err)


def test_union_carries_obc_types_and_edge_rivers(built):
    """Per-instance boundary set-ups survive the union: TYPE_OBC of every
    kind, NEXT_OBC2 (node numbers, offset), edge rivers (cell and node
    numbers, offset) beside a node river; each instance of the union still
    evolves exactly as it does alone (oracle)."""
    from oracle import Oracle
    from test_river_edge import _west_cell
    a = F.Model.build(0, nx=21, ny=7, kb=6)
    a.set_obc_type([(k % 10) + 1 for k in range(a.hdr.iobcn)])
    b = F.Model.build(0, nx=17, ny=9, kb=6, river=0)
    b.set_river_edge(_west_cell(b), 30.0, 12.0, 1.0)
    b.set_obc_type([5 + (k % 2) for k in range(b.hdr.iobcn)])
    for p in (a, b):
        p.cold_start()
    ps = [a, b]
    u, offs = F.Model.union(ps)
    assert list(u.field("type_obc")) == list(a.field("type_obc")) + list(b.field("type_obc"))
    om_b = offs[1][0]
    nob_a = a.hdr.iobcn
    assert np.array_equal(u.field("next_obc2")[nob_a:], b.field("next_obc2") + om_b)
    nq = u.hdr.numqbc
    assert nq == 2
    assert int(u.field("icellq")[1]) == int(b.field("icellq")[0]) + offs[1][1]
    nic = u.field("n_icellq").reshape(2, nq)
    assert list(nic[:, 1]) == [int(x) + om_b for x in b.field("n_icellq")]
    ou = Oracle(u)
    ou.run(3)
    for p, off in zip(ps, offs):
        op = Oracle(p)
        op.run(3)
        for name in FIELDS:
            x = instance_slice(ou.get(name), u, off, u.kb)
            y = own(op.get(name), p, u.kb)
            assert np.isfinite(y).all(), name
            assert np.array_equal(x, y), (name, off)


@pytest.mark.gpu
def test_union_with_boundary_kinds_gpu_parity(built):
    """The same union on the HIP path against the oracle, bit for bit."""
    from oracle import Oracle
    from helpers import ELEM2, ELEM3, NODE2, NODE3, compare
    from test_river_edge import _west_cell
    a = F.Model.build(0, nx=21, ny=7, kb=6)
    a.set_obc_type([(k % 10) + 1 for k in range(a.hdr.iobcn)])
    b = F.Model.build(0, nx=17, ny=9, kb=6, river=0)
    b.set_river_edge(_west_cell(b), 30.0, 12.0, 1.0)
    b.set_obc_type([5 + (k % 2) for k in range(b.hdr.iobcn)])
    for p in (a, b):
        p.cold_start()
    u, offs = F.Model.union([a, b])
    print("case: union of two channels with OBC types 1..10 / 5-6 and an edge river, 4 steps (graph)", flush=True)
    dev, orc = F.DeviceModel(u), Oracle(u)
    dev.run(4, graph=True)
    orc.run(4)
    rep = {}
    worst, bad = compare(dev, orc, u, NODE2 + ELEM2 + ELEM3 + NODE3, 0.0, rep)
    assert not bad, bad
