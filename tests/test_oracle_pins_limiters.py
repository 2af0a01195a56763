"""Pinning of the oracle's external momentum update, upwind face choice,
flux limiters, vertical-velocity mass correction and river-mouth adjustment
(oracle.Oracle.routine runs one routine on a state the test sets), each
against a property that does not come from the code under test:

  EXTUV_EDGE    one RK stage of d(UD)/dt = f V D + tau_s + tau_b on the
                f-plane (extuv_edge.F:112-140); four stages of the inertial
                oscillation give the degree-4 Taylor polynomial of
                exp(-i f dt) (ALPHA_RK = 1/4, 1/3, 1/2, 1); the sponge
                divides by 1 + c u^2 (extuv_edge.F:162-163)
  ADV_T, ADV_Q  upwind face values (adv_t.F:499-500, adv_q.F): advection of
                a bump by a uniform current with diffusion off does not raise
                the discrete variance sum(ART1 D DZ T^2), which a downwind
                choice does, and leaves no new extremum at a sharp front
  FCT_T/S       a local max principle: TF1 clipped into the range of the
                node's ring and of the column linearly interpolated to the
                layer's interfaces (fct_t.F:74-101)
  FCT_Q2/Q2L    the same for the interface fields (fct_q2.F)
  VERTVL_EDGE   with no horizontal flux the bottom omega of a column that
                does not close is removed by the linear correction and the
                depth becomes the old depth (vertvl_edge.F:500-533); closed
                basin volume is kept by DTFA
  ADJUST_TS     a river node's T1/S1 raised to the river's values, the
                elements around it the mean of their nodes (adjust_ts.F)
"""
import numpy as np
import pytest

import myfvcomcode_amd as F
from oracle import Oracle

ALPHA = {1: 0.25, 2: 1.0 / 3.0, 3: 0.5, 4: 1.0}


def _flat_basin(kb=6, depth=12.0, nx=21, ny=7, **pkw):
    """The estuary channel's mesh as a closed basin of uniform depth with no
    open boundary and no river; keyword arguments override parameters."""
    c = F.Model.build(0, nx=nx, ny=ny, kb=kb, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    p = F.Model.case_defaults(0).p
    for k, v in pkw.items():
        setattr(p, k, v)
    m = F.Model.from_mesh(x, y, nv, np.full(c.m, depth), kb, [], p)
    m.cold_start()
    return m


def _zero_2d_terms(o, n1):
    for nm in ("advua", "advva", "adx2d", "ady2d", "drx2d", "dry2d", "pstx", "psty",
               "wusurf2", "wvsurf2", "wubot", "wvbot"):
        o.put(nm, np.zeros(n1))


# ------------------------------------------------------------- EXTUV_EDGE
@pytest.mark.parametrize("K", [1, 2, 3, 4])
def test_extuv_stage_momentum_balance(built, K):
    """One RK stage K of the depth-integrated momentum equation without
    advection and pressure: UAF (H+ELF1) = UARK (H+ELRK1) + alpha_K DTE
    (f VA D1 + WUSURF2 + WUBOT), and the y component with -f UA D1.  The
    Coriolis force turns the current to the right (f > 0), the bottom stress
    WUBOT = -CBC |u| u brakes it, and the depth changes from the stage's base
    depth to the new one."""
    m = _flat_basin()
    n1 = m.n + 1
    f0 = 1.2e-4
    m.field("cor")[1:] = f0
    o = Oracle(m)
    rng = np.random.default_rng(7)
    H = 12.0
    ua, va = 0.5 * rng.standard_normal(n1), 0.5 * rng.standard_normal(n1)
    uark, vark = 0.5 * rng.standard_normal(n1), 0.5 * rng.standard_normal(n1)
    elrk1, elf1 = 0.3 * rng.standard_normal(n1), 0.3 * rng.standard_normal(n1)
    d1 = H + 0.2 * rng.standard_normal(n1)
    cbc = 2.5e-3
    wubot, wvbot = -cbc * np.hypot(ua, va) * ua, -cbc * np.hypot(ua, va) * va
    wusurf2, wvsurf2 = 1e-4 * rng.standard_normal(n1), 1e-4 * rng.standard_normal(n1)
    _zero_2d_terms(o, n1)
    for nm, a in (("ua", ua), ("va", va), ("uark", uark), ("vark", vark), ("elrk1", elrk1), ("elf1", elf1),
                  ("d1", d1), ("wubot", wubot), ("wvbot", wvbot), ("wusurf2", wusurf2), ("wvsurf2", wvsurf2)):
        o.put(nm, a)
    o.routine("extuv", K)
    uaf, vaf = o.get("uaf")[1:], o.get("vaf")[1:]
    e = slice(1, n1)
    dte, a = m.hdr.p.dte, ALPHA[K]
    eu = (uark[e] * (H + elrk1[e]) + a * dte * (f0 * va[e] * d1[e] + wusurf2[e] + wubot[e])) / (H + elf1[e])
    ev = (vark[e] * (H + elrk1[e]) + a * dte * (-f0 * ua[e] * d1[e] + wvsurf2[e] + wvbot[e])) / (H + elf1[e])
    assert np.max(np.abs(uaf - eu)) < 1e-14
    assert np.max(np.abs(vaf - ev)) < 1e-14


def test_extuv_four_stages_inertial_oscillation(built):
    """The four stages of an external step (EXTUV with UA, VA the previous
    stage's result, UARK, VARK the step's start) on a uniform current over
    an f-plane: z = UA + i VA obeys dz/dt = -i f z, and the stages give
    z0 T4(-i f DTE) with T4 the degree-4 Taylor polynomial of exp; repeated
    over an inertial period the current turns once clockwise."""
    m = _flat_basin()
    n1 = m.n + 1
    f0 = 1.0e-3                    # short period: 6283 s = 1257 steps of 5 s
    m.field("cor")[1:] = f0
    o = Oracle(m)
    _zero_2d_terms(o, n1)
    for nm in ("elrk1", "elf1"):
        o.put(nm, np.zeros(n1))
    o.put("d1", np.full(n1, 12.0))
    dte = m.hdr.p.dte
    z0 = 0.4 - 0.25j
    z = z0
    lam = -1j * f0 * dte
    t4 = 1 + lam + lam ** 2 / 2 + lam ** 3 / 6 + lam ** 4 / 24
    nsteps = int(round(2 * np.pi / f0 / dte))
    for step in range(nsteps):
        base = z
        o.put("uark", np.full(n1, base.real))
        o.put("vark", np.full(n1, base.imag))
        cur = base
        for K in (1, 2, 3, 4):
            o.put("ua", np.full(n1, cur.real))
            o.put("va", np.full(n1, cur.imag))
            o.routine("extuv", K)
            cur = complex(o.get("uaf")[5], o.get("vaf")[5])
        if step == 0:
            assert abs(cur - z0 * t4) < 1e-15
            assert np.all(o.get("uaf")[1:] == cur.real) and np.all(o.get("vaf")[1:] == cur.imag)
        z = cur
    # one inertial period later: back where it started, amplitude kept to
    # the scheme's O((f dt)^6) damping per step
    period_phase = nsteps * f0 * dte - 2 * np.pi
    assert abs(z - z0 * np.exp(-1j * period_phase)) < 1e-9
    assert abs(abs(z) - abs(z0)) < 1e-10


def test_extuv_sponge(built):
    """CC_SPONGE = c at an element divides its new velocity by 1 + c u^2
    per component; elements without sponge are untouched."""
    m = _flat_basin()
    n1 = m.n + 1
    rng = np.random.default_rng(3)
    spg = np.zeros(n1)
    pick = rng.choice(np.arange(1, n1), 40, replace=False)
    spg[pick] = rng.uniform(0.001, 0.05, 40)
    uark, vark = rng.standard_normal(n1), rng.standard_normal(n1)
    res = {}
    for name, c in (("free", np.zeros(n1)), ("sponge", spg)):
        m.field("cc_sponge")[:] = c
        o = Oracle(m)
        _zero_2d_terms(o, n1)
        for nm, a in (("ua", uark), ("va", vark), ("uark", uark), ("vark", vark), ("elrk1", np.zeros(n1)),
                      ("elf1", np.zeros(n1)), ("d1", np.full(n1, 12.0))):
            o.put(nm, a)
        o.routine("extuv", 2)
        res[name] = o.get("uaf"), o.get("vaf")
    (u0, v0), (u1, v1) = res["free"], res["sponge"]
    e = np.arange(1, n1)
    assert np.max(np.abs(u1[e] - u0[e] / (1 + spg[e] * u0[e] ** 2))) < 1e-15
    assert np.max(np.abs(v1[e] - v0[e] / (1 + spg[e] * v0[e] ** 2))) < 1e-15
    assert np.any(np.abs(u1[pick] - u0[pick]) > 1e-4)


# ------------------------------------------------- upwind face: ADV_T/ADV_Q
def _variance_case(kb=6):
    m = _flat_basin(kb=kb, nx=31, ny=25, horizontal_closure=0)
    m.field("nn_hvc")[:] = 0.0            # horizontal diffusion off
    return m


def _bump(m, width=2500.0):
    x, y = m.field("vx"), m.field("vy")
    xm, ym = 0.5 * (x[1:].max() + x[1:].min()), 0.5 * (y[1:].max() + y[1:].min())
    r = np.hypot(x - xm, y - ym) / width
    b = np.where(r < 1, np.cos(0.5 * np.pi * r) ** 2, 0.0)
    b[0] = 0.0
    return b


def _set_flow(o, m, U0, V0, D):
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    u = np.zeros((kb, n1))
    v = np.zeros((kb, n1))
    u[:-1, 1:] = U0
    v[:-1, 1:] = V0
    for nm, arr in (("u", u.ravel()), ("v", v.ravel()), ("wts", np.zeros(kb * m1)), ("d", np.full(m1, D)),
                    ("dt", np.full(m1, D)), ("dtfa", np.full(m1, D)), ("dt1", np.full(n1, D)),
                    ("viscofh", np.zeros(kb * m1))):
        o.put(nm, arr)


@pytest.mark.parametrize("U0,V0", [(0.8, 0.0), (0.0, -0.5), (0.6, 0.45)])
def test_adv_t_upwind_does_not_raise_variance(built, U0, V0):
    """A cosine bump of T carried by a uniform current over a flat bottom,
    diffusion off: the volume flux is divergence free node by node (the
    control volumes close), so sum(ART1 D DZ T) is kept and an upwind-biased
    face value makes sum(ART1 D DZ T^2) fall, step after step; taking the
    downwind node's face value makes it grow."""
    m = _variance_case()
    kb, m1 = m.kb, m.m + 1
    o = Oracle(m)
    D = 12.0
    _set_flow(o, m, U0, V0, D)
    art1, dz = m.field("art1"), m.field("dz").reshape(kb, m1)
    t = np.zeros((kb, m1))
    t[:-1] = 3.0 * _bump(m)[None, :]
    o.put("tmean1", np.zeros(kb * m1))
    mass0 = np.sum(art1[1:] * D * dz[:-1, 1:] * t[:-1, 1:])
    var = [np.sum(art1[1:] * D * dz[:-1, 1:] * t[:-1, 1:] ** 2)]
    for step in range(5):
        o.put("t1", t.ravel())
        o.routine("adv_ts", 1)
        t = o.get("tf1").reshape(kb, m1).copy()
        var.append(np.sum(art1[1:] * D * dz[:-1, 1:] * t[:-1, 1:] ** 2))
    mass = np.sum(art1[1:] * D * dz[:-1, 1:] * t[:-1, 1:])
    assert abs(mass - mass0) < 1e-12 * mass0
    dv = np.diff(var)
    assert np.all(dv < 0), dv
    assert var[0] - var[-1] > 1e-6 * var[0]


def test_adv_q_upwind_does_not_raise_variance(built):
    """The same for Q2 on the layer interfaces (ADV_Q, adv_q.F): no new
    variance under a uniform current with diffusion off."""
    m = _variance_case(kb=7)
    kb, m1 = m.kb, m.m + 1
    o = Oracle(m)
    D = 12.0
    _set_flow(o, m, 0.5, 0.35, D)
    art1, dzz = m.field("art1"), m.field("dzz").reshape(kb, m1)
    q = np.zeros((kb, m1))
    q[1:-1] = 1e-3 * _bump(m)[None, :]
    w = art1[1:] * D * dzz[:kb - 2, 1:]
    var = [np.sum(w * q[1:-1, 1:] ** 2)]
    for step in range(5):
        o.put("q2", q.ravel())
        o.routine("adv_q")
        q = o.get("q2f").reshape(kb, m1).copy()
        var.append(np.sum(w * q[1:-1, 1:] ** 2))
    dv = np.diff(var)
    assert np.all(dv < 0), dv


def test_adv_t_sharp_front_with_fct_stays_in_range(built):
    """A step in T (0 west, 1 east of a line) under a current of Courant
    number near 1 against it: ADV_T alone leaves values outside [0, 1];
    with FCT_T every node lies in [0, 1] again (a discrete max principle),
    and nodes away from the front are untouched."""
    m = _variance_case()
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    o = Oracle(m)
    D = 12.0
    U0 = -0.9 * 500.0 / dti                 # westward, 0.9 of a 500 m cell per step
    _set_flow(o, m, U0, 0.0, D)
    x = m.field("vx")
    xf = 0.5 * (x[1:].max() + x[1:].min()) + 130.0
    t = np.zeros((kb, m1))
    t[:-1, 1:] = (x[1:] > xf).astype(float)
    o.put("tmean1", np.zeros(kb * m1))
    o.put("t1", t.ravel())
    o.routine("adv_ts", 1)
    raw = o.get("tf1").reshape(kb, m1)[:-1, 1:]
    assert raw.max() > 1 + 1e-3 or raw.min() < -1e-3
    o.routine("fct_ts", 1)
    tf = o.get("tf1").reshape(kb, m1)[:-1, 1:]
    assert tf.min() >= 0.0 and tf.max() <= 1.0
    far = np.abs(x[1:] - xf) > 2000.0
    assert np.all(tf[:, far] == t[:-1, 1:][:, far])


# --------------------------------------------------------------- FCT_T/S, Q
def _ring(m, i):
    nbsn, off, ntsn = m.field("nbsn"), m.field("nbsn_off"), m.field("ntsn")
    return nbsn[off[i]:off[i] + ntsn[i]]


def _skip_nodes(m):
    return set(int(j) for j in m.field("i_obc_n")) | set(int(j) for j in m.field("inodeq"))


@pytest.mark.parametrize("which", [1, 2])
def test_fct_ts_local_max_principle_horizontal(built, which):
    """T1 uniform in the vertical and random across nodes, TF1 far outside
    it: FCT leaves TF1 clipped into [min, max] of T1 over the node and its
    ring (the vertical interpolants are the node's own value), to
    round-off; open-boundary and river nodes are skipped."""
    m = F.Model.build(0, nx=21, ny=7, kb=6)          # with its OBC and river node
    kb, m1 = m.kb, m.m + 1
    o = Oracle(m)
    rng = np.random.default_rng(11 + which)
    col = rng.uniform(10.0, 20.0, m1)
    f1 = np.tile(col, (kb, 1))
    ff = f1 + rng.normal(0.0, 6.0, (kb, m1))
    src, dst = ("t1", "tf1") if which == 1 else ("s1", "sf1")
    o.put(src, f1.ravel())
    o.put(dst, ff.ravel())
    o.routine("fct_ts", which)
    out = o.get(dst).reshape(kb, m1)
    skip = _skip_nodes(m)
    assert len(skip) > 5
    nclip = 0
    for i in range(1, m1):
        if i in skip:
            assert np.all(out[:-1, i] == ff[:-1, i])
            continue
        vals = np.append(col[_ring(m, i)], col[i])
        lo, hi = vals.min(), vals.max()
        exp = np.clip(ff[:-1, i], lo, hi)
        assert np.max(np.abs(out[:-1, i] - exp)) <= 4e-15 * hi, i
        nclip += int(np.sum((ff[:-1, i] < lo) | (ff[:-1, i] > hi)))
    assert nclip > 100


def _nonuniform_dz(m, weights):
    """Layer thicknesses proportional to `weights` at every node."""
    kb, m1 = m.kb, m.m + 1
    w = np.asarray(weights, float)
    dz = np.zeros((kb, m1))
    dz[:-1, :] = (w / w.sum())[:, None]
    m.field("dz")[:] = dz.ravel()
    return dz


def test_fct_t_vertical_bounds_are_interface_interpolants(built):
    """T1 the same at every node and increasing with depth, layers of
    unequal thickness: the ring gives only the layer's own value, so the
    bounds are the column linearly interpolated (in z, between the layer
    centres) to the layer's upper and lower interfaces; a TF1 beyond them
    lands on them."""
    m = F.Model.build(0, nx=21, ny=7, kb=7)
    kb, m1 = m.kb, m.m + 1
    dz = _nonuniform_dz(m, [1.0, 1.5, 2.0, 1.0, 3.0, 1.5])
    o = Oracle(m)
    tk = np.array([20.0, 18.5, 15.0, 14.0, 11.0, 10.5])
    f1 = np.zeros((kb, m1))
    f1[:-1] = tk[:, None]
    # centre depths and interfaces of the layer column (unit depth)
    zi = np.concatenate([[0.0], -np.cumsum(dz[:-1, 1])])
    zc = 0.5 * (zi[:-1] + zi[1:])
    up = np.interp(-zi[1:-1], -zc, tk)        # value at interface between k and k+1
    lo_b = np.empty(kb - 1)
    hi_b = np.empty(kb - 1)
    for k in range(kb - 1):
        cand = [tk[k]] if False else []
        if k > 0:
            cand.append(up[k - 1])
        if k < kb - 2:
            cand.append(up[k])
        cand.append(tk[k])                    # the ring's value at this level
        lo_b[k], hi_b[k] = min(cand), max(cand)
    rng = np.random.default_rng(5)
    ff = f1.copy()
    ff[:-1, 1:] += rng.choice([-5.0, 5.0], (kb - 1, m1 - 1))
    o.put("t1", f1.ravel())
    o.put("tf1", ff.ravel())
    o.routine("fct_ts", 1)
    out = o.get("tf1").reshape(kb, m1)
    skip = _skip_nodes(m)
    nodes = np.array([i for i in range(1, m1) if i not in skip])
    exp = np.clip(ff[:-1, nodes], lo_b[:, None], hi_b[:, None])
    assert np.max(np.abs(out[:-1, nodes] - exp)) < 1e-13
    # the interpolated bounds are strictly inside the neighbouring layers'
    # values, so raw neighbour values as bounds would fail this
    assert np.all(hi_b[1:-1] < np.maximum(tk[:-2], tk[2:]))


@pytest.mark.parametrize("which", [1, 2])
def test_fct_q_local_max_principle(built, which):
    """FCT_Q2 / FCT_Q2L on the interfaces 2..KBM1: a field uniform in the
    vertical and random across nodes is clipped into [min, max] over the
    node and its ring; a field the same at every node and increasing with
    depth is clipped to the means of neighbouring interfaces (equal
    sigma spacing)."""
    m = F.Model.build(0, nx=21, ny=7, kb=8)
    kb, m1 = m.kb, m.m + 1
    o = Oracle(m)
    src, dst = ("q2", "q2f") if which == 1 else ("q2l", "q2lf")
    rng = np.random.default_rng(21 + which)
    col = rng.uniform(1e-4, 1e-2, m1)
    q = np.tile(col, (kb, 1))
    qf = q * rng.uniform(-1.0, 3.0, (kb, m1))
    o.put(src, q.ravel())
    o.put(dst, qf.ravel())
    o.routine("fct_q", which)
    out = o.get(dst).reshape(kb, m1)
    skip = _skip_nodes(m)
    for i in range(1, m1):
        if i in skip:
            assert np.all(out[1:-1, i] == qf[1:-1, i])
            continue
        vals = np.append(col[_ring(m, i)], col[i])
        exp = np.clip(qf[1:kb - 1, i], vals.min(), vals.max())
        assert np.max(np.abs(out[1:kb - 1, i] - exp)) <= 4e-15 * vals.max(), i
    # vertical: q_k increasing with k at every node
    qk = np.array([0.0, 1.0, 2.0, 4.0, 7.0, 8.0, 10.0, 0.0]) * 1e-3
    q = np.tile(qk[:, None], (1, m1))
    qf = q.copy()
    qf[1:-1, 1:] += rng.choice([-5e-3, 5e-3], (kb - 2, m1 - 1))
    o.put(src, q.ravel())
    o.put(dst, qf.ravel())
    o.routine("fct_q", which)
    out = o.get(dst).reshape(kb, m1)
    nodes = np.array([i for i in range(1, m1) if i not in skip])
    for k in range(1, kb - 1):                 # interfaces 2..KBM1 (0-based rows 1..KBM1-1)
        cand = [qk[k]]
        if k > 1:
            cand.append(0.5 * (qk[k] + qk[k - 1]))
        if k < kb - 2:
            cand.append(0.5 * (qk[k] + qk[k + 1]))
        exp = np.clip(qf[k, nodes], min(cand), max(cand))
        assert np.max(np.abs(out[k, nodes] - exp)) < 1e-9, k


# ------------------------------------------------------------- VERTVL_EDGE
def test_vertvl_correction_without_horizontal_flux(built):
    """U = V = 0 and a depth D different from DT: the raw integral gives
    omega(k) = (k-1) DZ (D - DT)/DTI, which does not vanish at the bottom;
    the linear correction takes it out at every level (omega = 0) and the
    depth the column's volume supports is the old one, DTFA = DT."""
    m = _flat_basin(kb=9)
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    o = Oracle(m)
    rng = np.random.default_rng(9)
    H = 12.0
    el = 0.3 * rng.standard_normal(m1)
    et = el + 0.05 * rng.standard_normal(m1)
    for nm, a in (("u", np.zeros(kb * n1)), ("v", np.zeros(kb * n1)), ("elf", el), ("el", el),
                  ("d", H + el), ("dt", H + et), ("dtfa", np.full(m1, -1.0)), ("dt1", np.full(n1, H)),
                  ("wts", np.full(kb * m1, 7.0))):
        o.put(nm, a)
    o.routine("vertvl")
    wts = o.get("wts").reshape(kb, m1)
    dtfa = o.get("dtfa")
    assert np.max(np.abs(wts[:, 1:])) < 1e-15
    assert np.max(np.abs(dtfa[1:] - (H + et[1:]))) < 1e-12
    # W at elements is the node mean of WTS
    assert np.max(np.abs(o.get("w").reshape(kb, n1)[:, 1:])) < 1e-15


def test_vertvl_uniform_current_closed_basin(built):
    """A uniform current over a flat closed basin with D = DT: interior
    control volumes close (omega = 0 and DTFA untouched); at the walls the
    current converges or diverges, the layer fluxes are equal, so the
    corrected omega is 0 at every level there too, and DTFA takes up the
    column's volume change: sum(ART1 (DTFA - DT)) = 0 over the basin and
    DTFA - DT = -DTI sum_k XFLUX_k / ART1 has the sign of the convergence."""
    m = _flat_basin(kb=7)
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    o = Oracle(m)
    H = 12.0
    U0, V0 = 0.4, 0.1
    u = np.zeros((kb, n1))
    v = np.zeros((kb, n1))
    u[:-1, 1:] = U0
    v[:-1, 1:] = V0
    for nm, a in (("u", u.ravel()), ("v", v.ravel()), ("elf", np.zeros(m1)), ("el", np.zeros(m1)),
                  ("d", np.full(m1, H)), ("dt", np.full(m1, H)), ("dtfa", np.full(m1, -1.0)),
                  ("dt1", np.full(n1, H))):
        o.put(nm, a)
    o.routine("vertvl")
    wts = o.get("wts").reshape(kb, m1)
    dtfa = o.get("dtfa")
    isonb = m.field("isonb")
    art1, x, y = m.field("art1"), m.field("vx"), m.field("vy")
    assert np.max(np.abs(wts[:, 1:])) < 1e-15
    inner = np.where(isonb[1:] == 0)[0] + 1
    wall = np.where(isonb[1:] == 1)[0] + 1
    assert np.all(dtfa[inner] == -1.0)
    assert np.all(dtfa[wall] > 0)
    change = art1[wall] * (dtfa[wall] - H)
    assert abs(change.sum()) < 1e-9 * np.abs(change).sum()
    # water piles up against the walls the current runs into: the east
    # wall (x max) and the north wall (y max) gain, west and south lose
    east = wall[x[wall] > x[1:].max() - 1.0]
    west = wall[x[wall] < x[1:].min() + 1.0]
    corner = lambda i: (abs(y[i] - y[1:].max()) < 1.0) or (abs(y[i] - y[1:].min()) < 1.0)
    assert all(dtfa[i] > H for i in east if not corner(i))
    assert all(dtfa[i] < H for i in west if not corner(i))
    # the east wall's gain is the current's transport through the half
    # edges into its control volume: U0 H dy DTI per column (to the single
    # precision of the sigma levels, which DTFA's formula divides by)
    ew = [i for i in east if not corner(i)]
    dy = np.median(np.diff(np.sort(y[ew])))
    assert np.allclose(art1[ew] * (dtfa[ew] - H), U0 * H * dy * dti, rtol=1e-7)


# --------------------------------------------------------------- ADJUST_TS
def test_adjust_ts_raises_river_node(built):
    """RIVER_TS_SETTING 'calculated': at the river node every layer's T1
    below TDIS is raised to it, those above are kept (the same for S1), no
    other node changes, and the element T around it is the mean of its
    nodes."""
    m = F.Model.build(0, nx=21, ny=7, kb=6, river_t=15.0, river_s=4.0)
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    j = int(m.field("inodeq")[0])
    o = Oracle(m)
    rng = np.random.default_rng(4)
    t1 = rng.uniform(10.0, 20.0, (kb, m1))
    s1 = rng.uniform(0.0, 8.0, (kb, m1))
    o.put("t1", t1.ravel())
    o.put("s1", s1.ravel())
    o.routine("adjust_ts")
    t1n, s1n = o.get("t1").reshape(kb, m1), o.get("s1").reshape(kb, m1)
    tdis, sdis = m.field("tdis")[0], m.field("sdis")[0]
    assert tdis == 15.0 and sdis == 4.0
    assert np.all(t1n[:-1, j] == np.maximum(t1[:-1, j], tdis))
    assert np.all(s1n[:-1, j] == np.maximum(s1[:-1, j], sdis))
    assert np.any(t1[:-1, j] < tdis) and np.any(t1[:-1, j] > tdis)
    others = np.arange(1, m1) != j
    assert np.all(t1n[:-1, 1:][:, others] == t1[:-1, 1:][:, others])
    nv = m.field("nv").reshape(3, n1)
    t = o.get("t").reshape(kb, n1)
    near = [e for e in range(1, n1) if j in nv[:, e]]
    for e in near:
        assert np.allclose(t[:-1, e], t1n[:-1, nv[:, e]].mean(axis=1), rtol=1e-15, atol=0)


def test_calculated_river_keeps_river_node_at_or_above_river_values(built):
    """A whole internal step with RIVER_TS_SETTING 'calculated' (the river's
    T/S enter through the flux of ADV_T/ADV_S): afterwards the river node's
    T1 and S1 are at least the river's on every layer."""
    m = F.Model.build(0, nx=21, ny=7, kb=6, river_t=25.0, river_s=31.0)
    m.hdr.p.river_ts_specified = 0
    kb, m1 = m.kb, m.m + 1
    j = int(m.field("inodeq")[0])
    o = Oracle(m)
    o.run(3)
    t1, s1 = o.get("t1").reshape(kb, m1), o.get("s1").reshape(kb, m1)
    assert np.all(t1[:-1, j] >= 25.0) and np.all(s1[:-1, j] >= 31.0)


# ------------------------------------------------- wind at each ExtTime
def _drag(u, v):
    """Large and Pond drag law of PSIMPLE_DRAG (mod_bulk.F:58-90):
    tau = rho_air Cd |w| w with Cd = 1.2e-3 below 11 m/s, (0.49 + 0.065|w|)
    1e-3 up to 25 m/s, constant above; rho_air = 1.2."""
    w = np.hypot(u, v)
    cd = np.where(w >= 25.0, (0.49 + 0.065 * 25.0) * 1e-3,
                  np.where(w >= 11.0, (0.49 + 0.065 * w) * 1e-3, float(np.float32(1.2e-3))))
    return 1.2 * cd * u * w, 1.2 * cd * v * w


def test_external_wind_interpolated_to_ext_time(built):
    """With different winds at the start and end of the internal step, the
    external step IEXT takes the wind at its ExtTime, the fraction
    IEXT/ISPLIT of the way (UPDATE_WIND at ExtTime, bcond_gcn.F:878-895),
    and the internal surface stress the wind at the end (IntTime)."""
    m = F.Model.build(1, nx=31, ny=17, kb=5, n_seamounts=0)
    assert m.hdr.p.wind_on
    n1 = m.n + 1
    uu, vv, uu0, vv0 = (m.field(x) for x in ("uuwind", "vvwind", "uuwind0", "vvwind0"))
    rng = np.random.default_rng(2)
    uu0[1:] = 5.0 + 3 * rng.standard_normal(n1 - 1)
    vv0[1:] = -4.0 + 3 * rng.standard_normal(n1 - 1)
    uu[1:] = 12.0 + 8 * rng.standard_normal(n1 - 1)     # across the 11 and 25 m/s branches
    vv[1:] = 9.0 + 8 * rng.standard_normal(n1 - 1)
    o = Oracle(m)
    isplit = m.hdr.p.isplit
    for iext in (1, 4, isplit):
        o.external_step(iext)
        ramp = o.get("ramp")[0]
        f = iext / isplit
        sx, sy = _drag(uu0[1:] + f * (uu[1:] - uu0[1:]), vv0[1:] + f * (vv[1:] - vv0[1:]))
        np.testing.assert_allclose(o.get("wusurf2")[1:], sx * ramp * 0.001, rtol=1e-13, atol=1e-18)
        np.testing.assert_allclose(o.get("wvsurf2")[1:], sy * ramp * 0.001, rtol=1e-13, atol=1e-18)
    o.internal_step()
    sx, sy = _drag(uu[1:], vv[1:])
    ramp = o.get("ramp")[0]
    np.testing.assert_allclose(o.get("wusurf")[1:], -sx * ramp * 0.001, rtol=1e-13, atol=1e-18)
