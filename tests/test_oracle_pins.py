"""Pinning of the oracle's boundary conditions, mixing and diagnostic
routines one at a time (oracle.Oracle.routine runs a single routine on a
state the test sets), each against a closed form or a published result that
does not come from the code under test:

  VISCOF_H      Smagorinsky deformation of a linear velocity field is exact
                at interior nodes (viscofh.F:60-119)
  WREAL         sigma-to-z vertical velocity of linear D and elevation
                fields, w = omega + u.grad(z) + dz/dt (wreal.F:106-125)
  BCOND_GCN     wall projection parallel to the solid boundary edge, the
                river-mouth rule in the river's frame (bcond_gcn.F:325-616)
  VDIF_Q        MY2.5 (Mellor and Yamada 1982): implicit dissipation q^3/B1 l
                with the wall function W, buoyancy production, Galperin et
                al. (1988) stability functions and length-scale limit
                l <= 0.53 q/N, Craig-Banner surface and B1^(2/3) u*^2 bottom
                values (vdif_q.F:34-551)
  ADVAVE_EDGE_GCN  pressure terms of a linear elevation; closing momentum
                sums of divergence-free linear velocities
  ADVECTION_EDGE_GCN  the same, layer by layer
  ADV_Q, ADV_T  advection of linear fields
  ADV_UV_EDGE_GCN  UF/VF term by term: advection, pressure, Coriolis, omega
  BAROPG        hydrostatic baroclinic force on a flat bottom (baropg.F)
  VDIF_UV       column momentum budget with surface and implicit bottom
                stress; decay of the discrete Neumann cosine modes
  ADJUST2D3D    depth mean = UA, transport = UARD / ISPLIT
  BOTTOM_ROUGHNESS  law of the wall with the CBCMIN floor
  BCOND_TS      Blumberg and Kantha (1985) implicit radiation of the
                perturbation with 3 h relaxation about the 2-D upstream
                update of the column mean (bcond_ts.F:96-160,
                mod_obcs.F:1223-1295)

The routines are written into the step exactly as the HIP path runs them,
so these checks pin both.
"""
import numpy as np
import pytest

import myfvcomcode_amd as F
from oracle import Oracle

A1, B1, A2, B2, C1, E1, E2 = 0.92, 16.6, 0.74, 10.1, 0.08, 1.80, 1.33
KAPPA, SMALL = 0.40, 1e-8
# MY82 neutral stability functions (published: S_M = 0.393, S_H = 0.494)
SM0 = A1 * (1 - 3 * C1 - 6 * A1 / B1)
SH0 = A2 * (1 - 6 * A1 / B1)


def _linear_uv(m, a, b, c, d):
    kb, n1 = m.kb, m.n + 1
    xc, yc = m.field("xc"), m.field("yc")
    u = np.zeros((kb, n1))
    v = np.zeros((kb, n1))
    for k in range(kb - 1):
        s = 1 + 0.3 * k
        u[k, 1:] = s * (a * xc[1:] + b * yc[1:] + 0.3)
        v[k, 1:] = s * (c * xc[1:] + d * yc[1:] - 0.2)
    return u, v


def test_viscofh_linear_field_exact_at_interior_nodes(built):
    """For u = a x + b y, v = c x + d y at the element centroids, the
    centroid-ring circulation over ART1 is the exact gradient at interior
    nodes, so VISCOFH = sqrt(a^2 + d^2 + (b + c)^2 / 2) * ART1."""
    m = F.Model.build(0, nx=21, ny=7, kb=6)
    o = Oracle(m)
    a, b, c, d = 2e-4, -1.5e-4, 0.7e-4, 3e-4
    u, v = _linear_uv(m, a, b, c, d)
    o.put("u", u.ravel())
    o.put("v", v.ravel())
    o.routine("viscof_h")
    kb, m1 = m.kb, m.m + 1
    vh = o.get("viscofh").reshape(kb, m1)
    art1 = m.field("art1")
    inner = np.where(m.field("isonb")[1:] == 0)[0] + 1
    assert len(inner) > 50
    for k in range(kb - 1):
        exp = (1 + 0.3 * k) * np.sqrt(a * a + d * d + 0.5 * (b + c) ** 2) * art1[inner]
        assert np.max(np.abs(vh[k, inner] / exp - 1)) < 1e-13, k


def test_wreal_sigma_to_z_velocity(built):
    """With D, ELF and EL linear in x, y the element gradients and the
    three-node mean are exact: WW = (W_k + W_k+1)/2 + u (sigma dD/dx +
    dzeta/dx) + v (sigma dD/dy + dzeta/dy) + (1 + sigma) (zeta_c - ET1)/DTI."""
    m = F.Model.build(0, nx=21, ny=7, kb=8)
    kb, n1 = m.kb, m.n + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    o = Oracle(m)
    vx, vy, xc, yc = m.field("vx"), m.field("vy"), m.field("xc"), m.field("yc")
    zz1 = m.field("zz1").reshape(kb, n1)
    gD, gF, gE = (1e-4, -2e-4), (3e-6, 1e-6), (-2e-6, 4e-6)
    o.put("d", 10 + gD[0] * vx + gD[1] * vy)
    o.put("elf", 0.2 + gF[0] * vx + gF[1] * vy)
    o.put("el", 0.1 + gE[0] * vx + gE[1] * vy)
    rng = np.random.default_rng(2)
    et1 = 0.05 * rng.standard_normal(n1)
    u, v = rng.standard_normal((kb, n1)), rng.standard_normal((kb, n1))
    w = 1e-3 * rng.standard_normal((kb, n1))
    o.put("et1", et1)
    for nm, a in (("u", u), ("v", v), ("w", w)):
        o.put(nm, a.ravel())
    o.routine("wreal")
    ww = o.get("ww").reshape(kb, n1)
    e = slice(1, n1)
    zeta_c = 0.1 + gE[0] * xc[e] + gE[1] * yc[e]
    for k in range(kb - 1):
        s = zz1[k, e]
        exp = (0.5 * (w[k, e] + w[k + 1, e]) + u[k, e] * (s * gD[0] + gF[0]) + v[k, e] * (s * gD[1] + gF[1])
               + (1 + s) * (zeta_c - et1[e]) / dti)
        assert np.max(np.abs(ww[k, e] - exp)) < 1e-15, k


def _mouth_elements(m):
    nbve, off, ntve = m.field("nbve"), m.field("nbve_off"), m.field("ntve")
    mouth = {}
    for j, i1 in enumerate(m.field("inodeq")):
        i1 = int(i1)
        for e in (nbve[off[i1]], nbve[off[i1] + ntve[i1] - 1]):
            mouth.setdefault(int(e), j)
    return mouth


def _rotated_estuary(deg, kb=4):
    """The estuary channel turned by `deg` degrees, so that no wall is
    parallel to an axis, with its open boundary and its river node."""
    c = F.Model.build(0, nx=21, ny=7, kb=kb)
    x, y = c.field("vx")[1:], c.field("vy")[1:]
    th = np.deg2rad(deg)
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    m = F.Model.from_mesh(x * np.cos(th) - y * np.sin(th), x * np.sin(th) + y * np.cos(th), nv,
                          c.field("h")[1:].copy(), kb, c.field("i_obc_n").copy(), F.Model.case_defaults(0).p)
    m.set_river(int(c.field("inodeq")[0]), 20.0)
    m.cold_start()
    return m


@pytest.mark.parametrize("which", ["bcond_gcn2", "bcond_gcn5"])
def test_wall_projection_follows_the_boundary_edge(built, which):
    """ISBCE = 1 elements: the velocity becomes parallel to the element's
    solid edge with its tangential component kept (ALPHA carries the
    reference's 3.1415926/2, hence the 2e-7 bound); river-mouth elements
    keep the along-river component and lose only inflow against it; every
    other element is untouched."""
    m = _rotated_estuary(30.0)
    assert m.hdr.numqbc >= 1
    n1, kb = m.n + 1, m.kb
    o = Oracle(m)
    rng = np.random.default_rng(3)
    if which == "bcond_gcn2":
        u0, v0 = rng.standard_normal(n1), rng.standard_normal(n1)
        u0[0] = v0[0] = 0.0
        o.put("uaf", u0)
        o.put("vaf", v0)
        o.routine(which)
        u1, v1 = o.get("uaf")[None, :], o.get("vaf")[None, :]
        u0, v0 = u0[None, :], v0[None, :]
    else:
        u0, v0 = rng.standard_normal((kb, n1)), rng.standard_normal((kb, n1))
        u0[-1] = v0[-1] = 0.0
        u0[:, 0] = v0[:, 0] = 0.0
        o.put("u", u0.ravel())
        o.put("v", v0.ravel())
        o.routine(which)
        u1, v1 = o.get("u").reshape(kb, n1)[:-1], o.get("v").reshape(kb, n1)[:-1]
        u0, v0 = u0[:-1], v0[:-1]
    isbce, nbe = m.field("isbce"), m.field("nbe").reshape(3, n1)
    nv, vx, vy, angleq = m.field("nv").reshape(3, n1), m.field("vx"), m.field("vy"), m.field("angleq")
    mouth = _mouth_elements(m)
    n_wall = n_mouth = clipped = 0
    for e in range(1, n1):
        if isbce[e] == 1 and e in mouth:
            aq = angleq[mouth[e]]
            nrm, tan = (np.cos(aq), np.sin(aq)), (-np.sin(aq), np.cos(aq))
            un0 = u0[:, e] * nrm[0] + v0[:, e] * nrm[1]
            assert np.allclose(u1[:, e] * nrm[0] + v1[:, e] * nrm[1], np.maximum(un0, 0.0), atol=1e-15)
            assert np.allclose(u1[:, e] * tan[0] + v1[:, e] * tan[1], u0[:, e] * tan[0] + v0[:, e] * tan[1],
                               atol=1e-15)
            n_mouth += 1
            clipped += int(np.sum(un0 < 0))
        elif isbce[e] == 1:
            j = [q for q in range(3) if nbe[q, e] == 0][0]
            a, b = nv[(j + 1) % 3, e], nv[(j + 2) % 3, e]
            t = np.array([vx[b] - vx[a], vy[b] - vy[a]])
            t /= np.hypot(*t)
            mag = np.hypot(u0[:, e], v0[:, e])
            assert np.all(np.abs(u1[:, e] * t[1] - v1[:, e] * t[0]) <= 2e-7 * mag)
            tan0 = u0[:, e] * t[0] + v0[:, e] * t[1]
            assert np.all(np.abs(u1[:, e] * t[0] + v1[:, e] * t[1] - tan0) <= 2e-7 * mag)
            n_wall += 1
        elif isbce[e] == 3:
            assert np.all(u1[:, e] == 0) and np.all(v1[:, e] == 0)
        else:
            assert np.array_equal(u1[:, e], u0[:, e]) and np.array_equal(v1[:, e], v0[:, e])
    assert n_wall > 20 and n_mouth >= 1 and clipped >= 1


def _vq_case(kb=8):
    return F.Model.build(0, nx=21, ny=7, kb=kb, p_umol=0.0)


def _vq_setup(m, o, rho_k, q2k, lk, wind=(0.0, 0.0), bot=(0.0, 0.0), kh=0.0):
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    q2 = np.zeros((kb, m1))
    q2l = np.zeros((kb, m1))
    q2[:, 1:] = q2k[:, None]
    q2l[:, 1:] = (q2k * lk)[:, None]
    for nm, a in (("q2", q2), ("q2l", q2l), ("q2f", q2), ("q2lf", q2l)):
        o.put(nm, a.ravel())
    for nm, val in (("km", 0.0), ("kh", kh), ("kq", 0.0)):
        a = np.zeros((kb, m1))
        a[:, 1:] = val
        o.put(nm, a.ravel())
    rho = np.zeros((kb, m1))
    rho[:, 1:] = rho_k[:, None]
    o.put("rho1", rho.ravel())
    for nm, val in (("wusurf", wind[0]), ("wvsurf", wind[1]), ("wubot", bot[0]), ("wvbot", bot[1])):
        a = np.full(n1, val)
        a[0] = 0.0
        o.put(nm, a)
    return q2, q2l


def _stab(gh):
    """Galperin et al. (1988) quasi-equilibrium stability functions."""
    sh = SH0 / (1 - 3 * A2 * (B2 + 6 * A1) * gh)
    sm = (SM0 + sh * (18 * A1 * A1 + 9 * A1 * A2) * gh) / (1 - 9 * A1 * A2 * gh)
    return sm, sh


@pytest.mark.parametrize("case", ["neutral", "unstable"])
def test_vdif_q_my25_column(built, case):
    """No shear, no vertical diffusion (KQ = UMOL = 0): each level solves
    q2' (1 + 2 dt q / (B1 l)) = q2 + 2 dt P_b and
    q2l' (1 + dt q W / (B1 l)) = q2l + dt E1 l P_b, with P_b = KH N^2 of the
    old KH and W = 1 + E2 (l / kappa (1/d_surface + 1/d_bottom))^2; the new
    KM, KH are the means of the old values and l q S_M, l q S_H with
    Galperin's S_M(G_H), S_H(G_H), G_H capped at 0.0233."""
    m = _vq_case()
    kb, m1 = m.kb, m.m + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    Z = m.field("z").reshape(kb, m1)
    DZZ = m.field("dzz").reshape(kb, m1)
    grav = m.field("grav_n")
    o = Oracle(m)
    d = o.get("d")
    q2k = np.linspace(1e-4, 4e-4, kb)
    lk = np.linspace(0.5, 2.0, kb)
    if case == "neutral":
        rho_k, kh0 = np.full(kb, 0.02), 0.0
    else:
        rho_k = 0.025 - np.concatenate([[0.0], np.cumsum(np.geomspace(1e-10, 1e-6, kb - 1))])
        kh0 = 1e-3
    q2, q2l = _vq_setup(m, o, rho_k, q2k, lk, kh=kh0)
    o.routine("vdif_q")
    Q2F = o.get("q2f").reshape(kb, m1)
    Q2LF = o.get("q2lf").reshape(kb, m1)
    KM = o.get("km").reshape(kb, m1)
    KH = o.get("kh").reshape(kb, m1)
    q2a = np.abs(q2 + SMALL)
    la = q2l / q2a
    i = slice(1, m1)
    n_cap = 0
    for k in range(1, kb - 1):
        boygr = grav[i] * (rho_k[k - 1] - rho_k[k]) / (DZZ[k - 1, i] * d[i])
        gh = np.minimum(la[k, i] ** 2 / q2a[k, i] * boygr, 0.0233)
        n_cap += int(np.sum(gh == 0.0233))
        prod = kh0 * boygr
        diss = np.sqrt(q2a[k, i]) / (B1 * la[k, i] + SMALL)
        W = 1 + E2 * ((1 / np.abs(Z[k, i] - Z[0, i]) + 1 / np.abs(Z[k, i] - Z[kb - 1, i]))
                      * la[k, i] / (d[i] * KAPPA)) ** 2
        assert np.max(np.abs(Q2F[k, i] / ((q2[k, i] + 2 * dti * prod) / (1 + 2 * dti * diss)) - 1)) < 1e-14
        assert np.max(np.abs(Q2LF[k, i] / ((q2l[k, i] + dti * prod * la[k, i] * E1) / (1 + dti * diss * W)) - 1)) < 1e-14
        sm, sh = _stab(gh)
        kn = la[k, i] * np.sqrt(q2a[k, i])
        assert np.max(np.abs(KM[k, i] / (0.5 * kn * sm) - 1)) < 1e-14
        assert np.max(np.abs(KH[k, i] / (0.5 * (kn * sh * m.hdr.p.vprnu + kh0)) - 1)) < 1e-14
    if case == "neutral":
        assert round(SM0, 3) == 0.393 and round(SH0, 3) == 0.494
    else:
        assert 0 < n_cap < (kb - 2) * (m1 - 1)   # both branches of the cap taken


def test_vdif_q_stable_length_scale_limit(built):
    """Stable stratification: l is limited to sqrt(0.281) q / N = 0.53 q / N
    (Galperin et al. 1988) and Q2L follows as Q2 * L."""
    m = _vq_case()
    kb, m1 = m.kb, m.m + 1
    DZZ = m.field("dzz").reshape(kb, m1)
    grav = m.field("grav_n")
    o = Oracle(m)
    d = o.get("d")
    rho_k = 0.020 + 2e-3 * np.arange(kb)
    _vq_setup(m, o, rho_k, np.linspace(1e-4, 4e-4, kb), np.linspace(0.5, 2.0, kb))
    o.routine("vdif_q")
    L = o.get("l").reshape(kb, m1)
    Q2L = o.get("q2l").reshape(kb, m1)
    Q2 = o.get("q2").reshape(kb, m1)
    i = slice(1, m1)
    for k in range(1, kb - 1):
        n2 = -grav[i] * (rho_k[k - 1] - rho_k[k]) / (DZZ[k - 1, i] * d[i])
        r = L[k, i] * np.sqrt(n2 / Q2[k, i])
        assert np.all(r <= np.sqrt(0.281) * (1 + 1e-12))
        assert np.all(r > 0.5300)          # the limit binds at this stratification
        assert np.array_equal(Q2L[k, i], Q2[k, i] * L[k, i])


def test_vdif_q_surface_and_bottom_values(built):
    """Wind stress |tau|: q2 = (15.8 alpha_CB)^(2/3) |tau| at the surface
    (Craig and Banner 1994, alpha_CB = 100), L(1) = kappa * 2e5 |tau| / g
    and L >= kappa l0 in the upper half; bottom stress: q2 = B1^(2/3) |tau_b|.
    The reference's exponent 0.6666667 bounds the agreement at 3e-7."""
    m = _vq_case()
    kb, m1 = m.kb, m.m + 1
    Z = m.field("z").reshape(kb, m1)
    grav = m.field("grav_n")
    o = Oracle(m)
    q2k, lk = np.linspace(1e-4, 4e-4, kb), np.linspace(0.01, 0.05, kb)
    q2, q2l = _vq_setup(m, o, np.full(kb, 0.02), q2k, lk, wind=(3e-4, 4e-4), bot=(-3e-4, 4e-4))
    o.routine("vdif_q")
    Q2F = o.get("q2f").reshape(kb, m1)
    L = o.get("l").reshape(kb, m1)
    tau = 5e-4
    l0 = 2e5 * tau / grav[1:]
    assert np.max(np.abs(Q2F[0, 1:] / ((15.8 * 100) ** (2 / 3) * tau) - 1)) < 3e-7
    assert np.max(np.abs(Q2F[kb - 1, 1:] / (16.6 ** (2 / 3) * tau) - 1)) < 3e-7
    assert np.max(np.abs(L[0, 1:] / (KAPPA * l0) - 1)) < 1e-15
    assert np.all(L[kb - 1, 1:] == 0.0)
    la = q2l / np.abs(q2 + SMALL)
    raised = 0
    for k in range(1, kb - 1):
        up = Z[k, 1:] > -0.5
        assert np.array_equal(L[k, 1:], np.where(up, np.maximum(la[k, 1:], KAPPA * l0), la[k, 1:]))
        raised += int(np.sum(up & (KAPPA * l0 > la[k, 1:])))
    assert raised > 0


def test_bcond_ts_blumberg_kantha(built):
    """TYPE_TSOBC = 3 at outflow OBC nodes: the column mean is the 2-D
    upstream update (T2D DT - (XFLUX2D + T2D UARD_OBCN) DTI / ART1) / D and
    the perturbation about it radiates implicitly with c = sqrt(g H) and
    relaxes on 3 h: p' = (cc p_next + p (1 - DTI / 10800)) / (1 + cc),
    cc = c DTI / DLTN (Blumberg and Kantha 1985); inflow nodes keep T1.  The
    ring values bracket the result, so the clip to the local range is
    inactive here."""
    m = F.Model.build(0, nx=21, ny=7, kb=8, river=0)
    assert m.hdr.p.type_tsobc == 3
    kb, m1, nob = m.kb, m.m + 1, m.hdr.iobcn
    dti = m.hdr.p.dte * m.hdr.p.isplit
    o = Oracle(m)
    obc, nxt, dln = m.field("i_obc_n"), m.field("next_obc"), m.field("dltn_obc")
    DZ = m.field("dz").reshape(kb, m1)
    art1, h, g = m.field("art1"), m.field("h"), m.field("grav_n")
    nbsn, off, ntsn = m.field("nbsn"), m.field("nbsn_off"), m.field("ntsn")
    rng = np.random.default_rng(5)
    P = 15 + rng.standard_normal(kb - 1)
    P2 = 14 + rng.standard_normal(kb - 1)
    t1 = np.zeros((kb, m1))
    tf1 = np.zeros((kb, m1))
    t1[:-1, 1:] = P[:, None]
    obcs, used = set(int(j) for j in obc), set()
    for q, j in enumerate(obc):
        c = 20.0 if q % 2 == 0 else -20.0
        t1[:-1, j] = P + c
        ring = [int(x) for x in nbsn[off[j]:off[j] + ntsn[j]]
                if int(x) not in obcs and int(x) != j and int(x) not in used]
        t1[:-1, ring[0]] = P + 2 * c
        used.add(ring[0])
    tf1[:-1, 1:] = P2[:, None]
    d = o.get("d")
    dt = d * (1 + 0.01 * rng.standard_normal(m1))
    uo = np.abs(rng.standard_normal(nob)) * 50.0
    uo[1] = -1.0
    xf = np.zeros((kb, nob))
    xf[:-1] = rng.standard_normal((kb - 1, nob)) * 30.0
    for nm, a in (("dt", dt), ("t1", t1.ravel()), ("tf1", tf1.ravel()), ("uard_obcn", uo),
                  ("xflux_obc", xf.ravel())):
        o.put(nm, a)
    o.routine("bcond_ts", 1)
    out = o.get("tf1").reshape(kb, m1)
    for q, j in enumerate(obc):
        if uo[q] <= 0:
            assert np.array_equal(out[:-1, j], t1[:-1, j])
            continue
        j1 = nxt[q]
        t2d = np.sum(t1[:-1, j] * DZ[:-1, j])
        t2n = np.sum(tf1[:-1, j1] * DZ[:-1, j1])
        mean = (t2d * dt[j] - (np.sum(xf[:-1, q]) + t2d * uo[q]) * dti / art1[j]) / d[j]
        cc = np.sqrt(g[j] * h[j]) * dti / dln[q]
        exp = mean + (cc * (tf1[:-1, j1] - t2n) + (t1[:-1, j] - t2d) * (1 - dti / 10800.0)) / (1 + cc)
        ring = nbsn[off[j]:off[j] + ntsn[j]]
        assert np.all(t1[:-1, ring].min(axis=1) < exp) and np.all(exp < t1[:-1, ring].max(axis=1))
        assert np.max(np.abs(out[:-1, j] - exp)) < 1e-12
        # the perturbation carries no depth mean
        assert abs(np.sum(out[:-1, j] * DZ[:-1, j]) - mean) < 1e-12


def test_adjust2d3d_depth_mean_matches_external_mode(built):
    """ADJUST2D3D (adjust2d3d.F): type 1 shifts every layer of U by one
    amount so that sum_k U DZ1 = UA; type 2 makes the transport
    sum_k U DZ1 DT1 equal the external mode's time mean UARD / ISPLIT.  The
    vertical shear U_k - U_k+1 is kept in both."""
    m = F.Model.build(0, nx=21, ny=7, kb=8)
    kb, n1 = m.kb, m.n + 1
    dz1 = m.field("dz1").reshape(kb, n1)
    rng = np.random.default_rng(7)
    for typ in (1, 2):
        o = Oracle(m)
        u0, v0 = rng.standard_normal((kb, n1)), rng.standard_normal((kb, n1))
        u0[-1] = v0[-1] = 0.0
        u0[:, 0] = v0[:, 0] = 0.0
        ua, va = rng.standard_normal(n1), rng.standard_normal(n1)
        dt1 = 5 + rng.random(n1)
        for nm, a in (("u", u0.ravel()), ("v", v0.ravel()), ("ua", ua), ("va", va), ("uard", ua * 10),
                      ("vard", va * 10), ("dt1", dt1)):
            o.put(nm, a)
        o.routine("adjust2d3d", typ)
        u, v = o.get("u").reshape(kb, n1), o.get("v").reshape(kb, n1)
        e = slice(1, n1)
        for x, x0, mean2d in ((u, u0, ua), (v, v0, va)):
            depth_sum = np.sum(x[:-1, e] * dz1[:-1, e], axis=0)
            target = mean2d[e] if typ == 1 else mean2d[e] * 10 / m.hdr.p.isplit / dt1[e]
            assert np.max(np.abs(depth_sum - target)) < 1e-13
            assert np.max(np.abs(np.diff(x[:-1, e] - x0[:-1, e], axis=0))) < 1e-13
        if typ == 2:
            assert np.max(np.abs(o.get("uard")[e] - ua[e] * 10 / m.hdr.p.isplit)) < 1e-15


def test_bottom_roughness_law_of_the_wall(built):
    """BR_ORIG (brough.F:41-215): CBC = max((kappa / ln(z_b / z0))^2, CBCMIN)
    with z_b the height of the lowest layer centre above the bottom (the
    depth taken as at least 3 m), and the quadratic drag
    (WUBOT, WVBOT) = -CBC |u_b| u_b."""
    m = F.Model.build(0, nx=21, ny=7, kb=8)
    kb, n1 = m.kb, m.n + 1
    o = Oracle(m)
    rng = np.random.default_rng(8)
    dt1 = np.concatenate([[0.0], rng.uniform(0.5, 400.0, n1 - 1)])
    u, v = rng.standard_normal((kb, n1)), rng.standard_normal((kb, n1))
    for nm, a in (("dt1", dt1), ("u", u.ravel()), ("v", v.ravel())):
        o.put(nm, a)
    o.routine("bottom_roughness")
    zz1, z1 = m.field("zz1").reshape(kb, n1), m.field("z1").reshape(kb, n1)
    z0 = m.field("cc_z0b")
    e = slice(1, n1)
    zb = (zz1[kb - 2, e] - z1[kb - 1, e]) * np.maximum(dt1[e], 3.0)
    cbc = np.maximum((0.4 / np.log(zb / z0[e])) ** 2, m.hdr.p.cbcmin)
    got = o.get("cbc")[e]
    assert np.max(np.abs(got / cbc - 1)) < 1e-12
    assert np.any(got > m.hdr.p.cbcmin) and np.any(got == m.hdr.p.cbcmin)
    ub, vb = u[kb - 2, e], v[kb - 2, e]
    sp = np.hypot(ub, vb)
    assert np.max(np.abs(o.get("wubot")[e] + got * sp * ub)) < 1e-15
    assert np.max(np.abs(o.get("wvbot")[e] + got * sp * vb)) < 1e-15


def test_baropg_hydrostatic_flat_bottom(built):
    """Flat bottom, density anomaly linear in x and y plus any function of
    sigma: the layer's baroclinic force is g RAMP (d rho/dx) (-sigma_k D)
    (D DZ1_k) ART, the hydrostatic pressure gradient g rho_x (eta - z)
    integrated over the layer (baropg.F:33-166); the sigma part cancels
    around each element."""
    c = F.Model.build(0, nx=21, ny=7, kb=8, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    m = F.Model.from_mesh(x, y, nv, np.full(c.m, 12.0), 8, [], F.Model.case_defaults(0).p)
    m.cold_start()
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    o = Oracle(m)
    vx, vy = m.field("vx"), m.field("vy")
    zz = m.field("zz").reshape(kb, m1)
    zz1 = m.field("zz1").reshape(kb, n1)
    gx, gy, beta, D, ramp = 2e-7, -3e-7, 1e-3, 12.5, 0.7
    rho1 = np.zeros((kb, m1))
    rho1[:-1, 1:] = 0.02 + gx * vx[1:] + gy * vy[1:] + beta * zz[:-1, 1:]
    rho = np.zeros((kb, n1))
    rho[:-1, 1:] = 0.02 + beta * zz1[:-1, 1:]
    for nm, a in (("rho1", rho1.ravel()), ("rmean1", np.zeros(kb * m1)), ("rho", rho.ravel()),
                  ("rmean", np.zeros(kb * n1)), ("dt", np.full(m1, D)), ("dt1", np.full(n1, D)),
                  ("ramp", np.array([ramp]))):
        o.put(nm, a)
    o.routine("baropg")
    drx, dry = o.get("drhox").reshape(kb, n1), o.get("drhoy").reshape(kb, n1)
    art, dz1, ge = m.field("art"), m.field("dz1").reshape(kb, n1), m.field("grav_e")
    e = slice(1, n1)
    for k in range(kb - 1):
        f = ge[e] * ramp * (-zz1[k, e]) * D * D * dz1[k, e] * art[e]
        assert np.max(np.abs(drx[k, e] / (gx * f) - 1)) < 1e-11, k
        assert np.max(np.abs(dry[k, e] / (gy * f) - 1)) < 1e-11, k
    # RHO1 and RHO come back unchanged (the mean is removed and restored)
    assert np.max(np.abs(o.get("rho1") - rho1.ravel())) < 1e-17


def test_vdif_uv_column_budget_and_modes(built):
    """VDIF_UV (vdif_uv.F:34-189), sponge off: the implicit solve changes
    the column momentum by the surface stress (none at elements touching the
    open boundary) and the implicit quadratic bottom drag,
    sum_k (UF' - UF) DZ1 D / DTI = -WUSURF - CBC |u_b| UF'_KBM1; with no
    stress and a uniform K the cosine modes of the Neumann problem decay by
    1 / (1 + 4 DTI (K + UMOL) sin^2(pi m / 2N) / (dsigma D)^2)."""
    m = F.Model.build(0, nx=21, ny=7, kb=8)
    m.field("cc_sponge")[:] = 0.0
    kb, n1 = m.kb, m.n + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    dz1 = m.field("dz1").reshape(kb, n1)
    o = Oracle(m)
    rng = np.random.default_rng(4)
    uf0, vf0 = rng.standard_normal((kb, n1)), rng.standard_normal((kb, n1))
    uf0[-1] = vf0[-1] = 0.0
    uf0[:, 0] = vf0[:, 0] = 0.0
    u, v = rng.standard_normal((kb, n1)), rng.standard_normal((kb, n1))
    ws, vs = 1e-4 * rng.standard_normal(n1), 1e-4 * rng.standard_normal(n1)
    cbc = 2.5e-3 + 1e-3 * rng.random(n1)
    for nm, a in (("uf", uf0.ravel()), ("vf", vf0.ravel()), ("u", u.ravel()), ("v", v.ravel()),
                  ("km1", (np.abs(rng.standard_normal((kb, n1))) * 1e-2).ravel()),
                  ("wusurf", ws), ("wvsurf", vs), ("cbc", cbc)):
        o.put(nm, a)
    o.routine("vdif_uv")
    uf, vf, d1 = o.get("uf").reshape(kb, n1), o.get("vf").reshape(kb, n1), o.get("d1")
    e = slice(1, n1)
    isonb, nv = m.field("isonb"), m.field("nv").reshape(3, n1)
    open_el = (isonb[nv[0]] == 2) | (isonb[nv[1]] == 2) | (isonb[nv[2]] == 2)
    assert np.any(open_el[1:])
    tmp = np.where(open_el, 0.0, 1.0)
    tps = cbc * np.hypot(u[kb - 2], v[kb - 2])
    for x, x0, s in ((uf, uf0, ws), (vf, vf0, vs)):
        change = np.sum((x[:-1, e] - x0[:-1, e]) * dz1[:-1, e], axis=0) * d1[e] / dti
        assert np.max(np.abs(change - (-s[e] * tmp[e] - tps[e] * x[kb - 2, e]))) < 1e-15
    assert np.max(np.abs(o.get("wubot")[e] + tps[e] * u[kb - 2, e])) < 1e-17
    # modes (the sigma levels are single-precision fractions: 3e-8 spread)
    N, K = kb - 1, 3e-2
    for mode in (1, 3):
        prof = np.cos(np.pi * mode * (np.arange(1, N + 1) - 0.5) / N)
        x0 = np.zeros((kb, n1))
        x0[:-1, 1:] = prof[:, None]
        for nm, a in (("uf", x0.ravel()), ("vf", 2 * x0.ravel()), ("km1", np.full(kb * n1, K)),
                      ("wusurf", np.zeros(n1)), ("wvsurf", np.zeros(n1)), ("cbc", np.zeros(n1))):
            o.put(nm, a)
        o.routine("vdif_uv")
        uf, vf, d1 = o.get("uf").reshape(kb, n1), o.get("vf").reshape(kb, n1), o.get("d1")
        fac = 1 + 4 * dti * (K + m.hdr.p.umol) * np.sin(np.pi * mode / (2 * N)) ** 2 / (d1[1:] / N) ** 2
        assert np.max(np.abs(uf[:-1, 1:] - prof[:, None] / fac[None, :])) < 1e-7
        assert np.max(np.abs(vf[:-1, 1:] - 2 * prof[:, None] / fac[None, :])) < 2e-7
        assert np.min(fac) > 1.01        # the damping is well resolved by the bound


def test_advave_pressure_and_uniform_momentum(built):
    """ADVAVE_EDGE_GCN (advave_edge_gcn.F:546-884) on a flat bottom with
    constant horizontal viscosity: for a linear elevation the edge-midpoint
    pressure terms sum to (PSTX, PSTY) = g D1 grad(eta) ART on every element;
    for UA = U0 and VA = V0 + c x (divergence-free, reconstructed exactly by
    the least-squares gradients) the x-momentum flux U0 (u.n) and the
    constant stress c nu close around every element whose neighbours are
    interior too, so ADVUA = 0 there, while ADVVA = D U0 c ART; for
    UA = U0 + a y, VA = V0 the y-momentum closes (ADVVA = 0)."""
    c = F.Model.build(0, nx=21, ny=7, kb=6, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    p = F.Model.case_defaults(0).p
    p.horizontal_closure = 0
    m = F.Model.from_mesh(x, y, nv, np.full(c.m, 12.0), 6, [], p)
    m.cold_start()
    m.field("cc_hvc")[:] = 7.0
    m1, n1 = m.m + 1, m.n + 1
    o = Oracle(m)
    xc, vx, vy = m.field("xc"), m.field("vx"), m.field("vy")
    U0, V0, cx, D, gx, gy = 0.8, -0.3, 2e-4, 12.0, 3e-6, -2e-6
    ua, va = np.full(n1, U0), V0 + cx * xc
    ua[0] = va[0] = 0.0
    for nm, a in (("ua", ua), ("va", va), ("d", np.full(m1, D)), ("d1", np.full(n1, D)),
                  ("el", 0.1 + gx * vx + gy * vy)):
        o.put(nm, a)
    o.routine("advave")
    art, ge = m.field("art"), m.field("grav_e")
    e = np.arange(1, n1)
    assert np.max(np.abs(o.get("pstx")[e] / (ge[e] * D * gx * art[e]) - 1)) < 1e-12
    assert np.max(np.abs(o.get("psty")[e] / (ge[e] * D * gy * art[e]) - 1)) < 1e-12
    isbce, nbe = m.field("isbce"), m.field("nbe").reshape(3, n1)
    deep = np.array([i for i in e if isbce[i] == 0 and all(nbe[j, i] != 0 and isbce[nbe[j, i]] == 0
                                                           for j in range(3))])
    assert len(deep) > 100
    # terms of one edge are D |u.n| U0 ~ 1e4 here
    assert np.max(np.abs(o.get("advua")[deep])) < 1e-10
    # y-momentum: v (u.n) is linear along each edge, so the midpoint rule is
    # exact and ADVVA = D U0 dv/dx ART
    assert np.max(np.abs(o.get("advva")[deep] / (D * U0 * cx * art[deep]) - 1)) < 1e-12
    # UA = U0 + a y, VA = V0: V0 (u.n) and the constant stress a nu close
    ua = U0 + 3e-4 * m.field("yc")
    va = np.full(n1, V0)
    ua[0] = va[0] = 0.0
    o.put("ua", ua)
    o.put("va", va)
    o.routine("advave")
    assert np.max(np.abs(o.get("advva")[deep])) < 1e-10


def test_advection_edge_layers_close(built):
    """ADVECTION_EDGE_GCN (advection_edge_gcn.F:29-734) per layer, flat
    bottom, constant viscosity: U = U0_k, V = V0 + c_k x gives ADVX = 0 and
    ADVY = D DZ_k U0_k c_k ART at elements with interior neighbours (the
    2-D closed forms with the layer's thickness)."""
    c = F.Model.build(0, nx=21, ny=7, kb=6, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    p = F.Model.case_defaults(0).p
    p.horizontal_closure = 0
    m = F.Model.from_mesh(x, y, nv, np.full(c.m, 12.0), 6, [], p)
    m.cold_start()
    m.field("cc_hvc")[:] = 7.0
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    o = Oracle(m)
    xc, D = m.field("xc"), 11.0
    u, v = np.zeros((kb, n1)), np.zeros((kb, n1))
    U0 = 0.8 - 0.1 * np.arange(kb - 1)
    cx = 2e-4 * (1 + np.arange(kb - 1))
    u[:-1, 1:] = U0[:, None]
    v[:-1, 1:] = -0.3 + cx[:, None] * xc[None, 1:]
    o.put("u", u.ravel())
    o.put("v", v.ravel())
    o.put("dt", np.full(m1, D))
    o.routine("advection")
    advx, advy = o.get("advx").reshape(kb, n1), o.get("advy").reshape(kb, n1)
    dz1, art = m.field("dz1").reshape(kb, n1), m.field("art")
    isbce, nbe = m.field("isbce"), m.field("nbe").reshape(3, n1)
    deep = np.array([i for i in range(1, n1) if isbce[i] == 0 and all(
        nbe[j, i] != 0 and isbce[nbe[j, i]] == 0 for j in range(3))])
    assert len(deep) > 100
    for k in range(kb - 1):
        assert np.max(np.abs(advx[k, deep])) < 1e-10, k
        exp = D * dz1[k, deep] * U0[k] * cx[k] * art[deep]
        assert np.max(np.abs(advy[k, deep] / exp - 1)) < 1e-12, k


def test_adv_uv_momentum_terms(built):
    """ADV_UV_EDGE_GCN (adv_uv_edge_gcn.F:31-1198) on a flat bottom with
    constant viscosity, depth-uniform U = U0 and V = V0 + c x, a linear
    EGF, Coriolis f and any omega W: each term of the momentum equation has
    its closed form,
      UF = DT1/D1 (U0 + DTI f V - DTI g EGF_x) - DTI U0 (W_k - W_k+1) / (D1 DZ_k)
      VF = DT1/D1 (V - DTI f U0 - DTI g EGF_y) - DTI DT U0 c / D1
           - DTI V (W_k - W_k+1) / (D1 DZ_k)
    at elements with interior neighbours (W_1 = W_KB = 0)."""
    c = F.Model.build(0, nx=21, ny=7, kb=6, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    p = F.Model.case_defaults(0).p
    p.horizontal_closure = 0
    m = F.Model.from_mesh(x, y, nv, np.full(c.m, 12.0), 6, [], p)
    m.cold_start()
    m.field("cc_hvc")[:] = 7.0
    f = 1e-4
    m.field("cor")[:] = f
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    o = Oracle(m)
    xc, vx, vy = m.field("xc"), m.field("vx"), m.field("vy")
    D, D1, U0, V0, cx, gx, gy = 11.0, 11.3, 0.6, -0.3, 2e-4, 3e-6, -2e-6
    rng = np.random.default_rng(9)
    u, v, w = np.zeros((kb, n1)), np.zeros((kb, n1)), np.zeros((kb, n1))
    u[:-1, 1:] = U0
    v[:-1, 1:] = V0 + cx * xc[None, 1:]
    w[1:-1, 1:] = 1e-4 * rng.standard_normal((kb - 2, n1 - 1))
    for nm, a in (("u", u.ravel()), ("v", v.ravel()), ("w", w.ravel()), ("dt", np.full(m1, D)),
                  ("dt1", np.full(n1, D)), ("d1", np.full(n1, D1)), ("egf", gx * vx + gy * vy),
                  ("drhox", np.zeros(kb * n1)), ("drhoy", np.zeros(kb * n1))):
        o.put(nm, a)
    o.routine("adv_uv")
    uf, vf = o.get("uf").reshape(kb, n1), o.get("vf").reshape(kb, n1)
    dz1, ge = m.field("dz1").reshape(kb, n1), m.field("grav_e")
    isbce, nbe = m.field("isbce"), m.field("nbe").reshape(3, n1)
    e = np.array([i for i in range(1, n1) if isbce[i] == 0 and all(
        nbe[j, i] != 0 and isbce[nbe[j, i]] == 0 for j in range(3))])
    assert len(e) > 100
    for k in range(kb - 1):
        dw = (w[k, e] - w[k + 1, e]) / (D1 * dz1[k, e])
        vk = v[k, e]
        eu = D / D1 * (U0 + dti * f * vk - dti * ge[e] * gx) - dti * U0 * dw
        ev = D / D1 * (vk - dti * f * U0 - dti * ge[e] * gy) - dti * D * U0 * cx / D1 - dti * vk * dw
        assert np.max(np.abs(uf[k, e] - eu)) < 1e-12, k
        assert np.max(np.abs(vf[k, e] - ev)) < 1e-12, k


@pytest.mark.parametrize("routine", ["advave", "advection", "adv_uv"])
@pytest.mark.parametrize("closure", [0, 1])
def test_edge_viscosity_closed_forms(built, routine, closure):
    """The horizontal viscosity of the three edge routines (ADVAVE_EDGE_GCN
    advave_edge_gcn.F:166-320, ADVECTION_EDGE_GCN advection_edge_gcn.F,
    ADV_UV_EDGE_GCN adv_uv_edge_gcn.F), isolated as the difference between
    CC_HVC = 0.3 and CC_HVC = 0 (the advective terms do not depend on it),
    for the shear flow u = a_k y^2, v = 0 (a_k per layer) on a uniform flat
    mesh, HPRNU = 2, at elements with interior neighbours:
      constant (HORIZONTAL_CLOSURE off): nu = CC/HPRNU and the viscous
        acceleration is nu lap(u) = 2 a nu, exactly (the least-squares
        gradients of a quadratic are exact in the mean of the two sides);
      Smagorinsky (on): nu = CC ART |S| / HPRNU with |S| = |u_y|/sqrt2
        = sqrt2 a y, so d/dy(nu u_y) = 4 sqrt2 CC ART a^2 y / HPRNU, within
        the scheme's discretisation error (here 6 %, 0.3 % on average).
    The y-force of the scheme alternates between neighbouring elements but
    sums to zero over them.  Terms: ADVUA = -D ART acc, ADVX = -D DZ_k ART
    acc, UF gains DTI acc D/D1."""
    c = F.Model.build(0, nx=21, ny=7, kb=6, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    D, D1, CC, HPRNU = 12.0, 12.4, 0.3, 2.0
    p = F.Model.case_defaults(0).p
    p.horizontal_closure = closure
    p.hprnu = HPRNU
    m = F.Model.from_mesh(x, y + 20000.0, nv, np.full(c.m, D), 6, [], p)
    m.cold_start()
    m.field("cor")[:] = 0.0
    kb, n1, m1 = m.kb, m.n + 1, m.m + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    yc, art, dz1 = m.field("yc"), m.field("art"), m.field("dz1").reshape(kb, n1)
    assert np.ptp(art[1:]) == 0.0
    a = (1e-9 if closure else 1e-7) * (1 + np.arange(kb - 1))
    out = {}
    for cc in (0.0, CC):
        m.field("cc_hvc")[:] = cc
        o = Oracle(m)
        if routine == "advave":
            ua = a[0] * yc ** 2
            ua[0] = 0.0
            for nm, arr in (("ua", ua), ("va", np.zeros(n1)), ("d", np.full(m1, D)), ("d1", np.full(n1, D)),
                            ("el", np.zeros(m1))):
                o.put(nm, arr)
            o.routine("advave")
            out[cc] = (o.get("advua")[None, :], o.get("advva")[None, :])
            continue
        u = np.zeros((kb, n1))
        u[:-1, 1:] = a[:, None] * yc[None, 1:] ** 2
        arrs = [("u", u.ravel()), ("v", np.zeros(kb * n1)), ("dt", np.full(m1, D))]
        if routine == "adv_uv":
            arrs += [("w", np.zeros(kb * n1)), ("dt1", np.full(n1, D)), ("d1", np.full(n1, D1)),
                     ("egf", np.zeros(m1)), ("drhox", np.zeros(kb * n1)), ("drhoy", np.zeros(kb * n1))]
        for nm, arr in arrs:
            o.put(nm, arr)
        o.routine(routine)
        names = ("advx", "advy") if routine == "advection" else ("uf", "vf")
        out[cc] = tuple(o.get(nm).reshape(kb, n1)[:-1] for nm in names)
    du, dv = out[CC][0] - out[0.0][0], out[CC][1] - out[0.0][1]
    isbce, nbe = m.field("isbce"), m.field("nbe").reshape(3, n1)
    e = np.array([i for i in range(1, n1) if isbce[i] == 0 and all(
        nbe[j, i] != 0 and isbce[nbe[j, i]] == 0 for j in range(3))])
    assert len(e) > 100
    for k in range(du.shape[0]):
        if closure:
            acc = 4 * np.sqrt(2) * CC * art[e] * a[k] ** 2 * yc[e] / HPRNU
        else:
            acc = 2 * a[k] * CC / HPRNU
        if routine == "advave":
            want = -acc * D * art[e]
        elif routine == "advection":
            want = -acc * D * dz1[k, e] * art[e]
        else:
            want = dti * acc * D / D1
        r = du[k, e] / want
        if closure:
            assert abs(r.mean() - 1) < 0.01 and np.all(np.abs(r - 1) < 0.08), (k, r.min(), r.max(), r.mean())
        else:
            assert np.max(np.abs(r - 1)) < 1e-9, (k, r.min(), r.max())
        assert abs(dv[k, e].sum()) < 1e-9 * np.abs(du[k, e]).sum(), k


@pytest.mark.parametrize("routine", ["advave", "advection", "adv_uv"])
def test_edge_viscosity_independent_of_numbering(built, routine):
    """The Smagorinsky viscosity of an edge averages ART |S| CC of its two
    elements, so the viscous terms cannot depend on which element is the
    edge's IA: numbering the elements of a jittered mesh (areas varying by
    over 100 %) in reverse order swaps IA and IB on every interior edge and
    leaves them unchanged up to round-off.  (A term taking one side's area
    twice changes them by percents.)"""
    c = F.Model.build(1, nx=31, ny=17, kb=6, n_seamounts=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    p = F.Model.case_defaults(0).p
    D = 12.0

    def viscous(order):
        m = F.Model.from_mesh(x, y, nv[order], np.full(c.m, D), 6, [], p)
        m.cold_start()
        m.field("cor")[:] = 0.0
        kb, n1, m1 = m.kb, m.n + 1, m.m + 1
        xc, yc = m.field("xc"), m.field("yc")
        ua = 0.3 * np.sin(2 * np.pi * yc / y.max()) + 0.1 * np.cos(3 * xc / x.max())
        va = 0.2 * np.cos(2 * np.pi * xc / x.max()) * np.sin(yc / y.max())
        ua[0] = va[0] = 0.0
        out = {}
        for cc in (0.0, 0.3):
            m.field("cc_hvc")[:] = cc
            o = Oracle(m)
            if routine == "advave":
                for nm, arr in (("ua", ua), ("va", va), ("d", np.full(m1, D)), ("d1", np.full(n1, D)),
                                ("el", np.zeros(m1))):
                    o.put(nm, arr)
                o.routine("advave")
                out[cc] = np.stack([o.get("advua")[1:], o.get("advva")[1:]])
                continue
            u, v = np.zeros((kb, n1)), np.zeros((kb, n1))
            u[:-1] = ua[None, :] * (1 + 0.1 * np.arange(kb - 1))[:, None]
            v[:-1] = va[None, :]
            arrs = [("u", u.ravel()), ("v", v.ravel()), ("dt", np.full(m1, D))]
            if routine == "adv_uv":
                arrs += [("w", np.zeros(kb * n1)), ("dt1", np.full(n1, D)), ("d1", np.full(n1, D)),
                         ("egf", np.zeros(m1)), ("drhox", np.zeros(kb * n1)), ("drhoy", np.zeros(kb * n1))]
            for nm, arr in arrs:
                o.put(nm, arr)
            o.routine(routine)
            names = ("advx", "advy") if routine == "advection" else ("uf", "vf")
            out[cc] = np.stack([o.get(nm).reshape(kb, n1)[:-1, 1:] for nm in names])
        art = m.field("art")[1:]
        return out[0.3] - out[0.0], np.ptp(art) / art.mean()

    n = nv.shape[0]
    fwd, spread = viscous(np.arange(n))
    rev, _ = viscous(np.arange(n)[::-1])
    rev = rev[..., ::-1]
    assert spread > 1.0
    assert np.abs(fwd).max() > 0.0
    assert np.abs(fwd - rev).max() < 1e-12 * np.abs(fwd).max()


def test_bcond_ts_river_values_and_local_bounds(built):
    """BCOND_TS (bcond_ts.F:60-95, 96-160): with the river's T and S given,
    every layer of the river node takes them (both tracers in one call, as
    the reference does); at an outflow OBC node the Blumberg-Kantha value is
    kept where it lies inside the range of the node's ring and column
    neighbours and clipped into that range where a large inflowing flux
    drives it out (the FCT principle: no new extrema)."""
    m = F.Model.build(0, nx=21, ny=7, kb=8)
    assert m.hdr.p.river_ts_specified == 1 and m.hdr.numqbc == 1
    kb, m1, nob = m.kb, m.m + 1, m.hdr.iobcn
    dti = m.hdr.p.dte * m.hdr.p.isplit
    o = Oracle(m)
    j = int(m.field("inodeq")[0])
    t1 = np.zeros((kb, m1))
    t1[:-1, 1:] = 12.0 + np.random.default_rng(6).uniform(-0.5, 0.5, (kb - 1, m1 - 1))
    o.put("t1", t1.ravel())
    o.put("tf1", t1.ravel())
    o.put("sf1", t1.ravel())
    uo = np.full(nob, 10.0)
    o.put("uard_obcn", uo)
    xf = np.zeros((kb, nob))
    xf[:-1, 0] = -5e4            # OBC node 0: a large inflow of tracer content
    o.put("xflux_obc", xf.ravel())
    o.routine("bcond_ts", 1)
    tf = o.get("tf1").reshape(kb, m1)
    sf = o.get("sf1").reshape(kb, m1)
    assert np.all(tf[:-1, j] == m.field("tdis")[0]) and np.all(sf[:-1, j] == m.field("sdis")[0])
    obc, nxt, dln = m.field("i_obc_n"), m.field("next_obc"), m.field("dltn_obc")
    DZ, art1, h, g = m.field("dz").reshape(kb, m1), m.field("art1"), m.field("h"), m.field("grav_n")
    d, dt = o.get("d"), o.get("dt")
    nbsn, off, ntsn = m.field("nbsn"), m.field("nbsn_off"), m.field("ntsn")
    kept = clipped = 0
    for q, jo in enumerate(obc):
        j1 = nxt[q]
        t2d = np.sum(t1[:-1, jo] * DZ[:-1, jo])
        t2n = np.sum(t1[:-1, j1] * DZ[:-1, j1])
        mean = (t2d * dt[jo] - (np.sum(xf[:-1, q]) + t2d * uo[q]) * dti / art1[jo]) / d[jo]
        cc = np.sqrt(g[jo] * h[jo]) * dti / dln[q]
        raw = mean + (cc * (t1[:-1, j1] - t2n) + (t1[:-1, jo] - t2d) * (1 - dti / 10800.0)) / (1 + cc)
        ring = list(nbsn[off[jo]:off[jo] + ntsn[jo]])
        for k in range(kb - 1):
            col = t1[max(k - 1, 0):min(k + 2, kb - 1), jo]
            lo = min(t1[k, ring].min(), col.min())
            hi = max(t1[k, ring].max(), col.max())
            assert lo <= tf[k, jo] <= hi
            if t1[k, ring].min() < raw[k] < t1[k, ring].max():
                assert abs(tf[k, jo] - raw[k]) < 1e-12
                kept += 1
            elif raw[k] > hi:
                assert tf[k, jo] < raw[k]
                clipped += 1
    assert kept > 0 and clipped > 0


def test_adv_q_linear_fields(built):
    """ADV_Q (adv_q.F:35-418) on a flat bottom with DT = D and constant
    horizontal diffusivity: a field linear in x carried by a uniform U0
    gives exactly Q2F = Q2 - DTI U0 dQ/dx at nodes whose ring is interior
    (the face reconstruction is exact, the CV fluxes close, the midpoint
    rule is exact); a field linear in the level index under a uniform omega
    W0 gives Q2F = Q2 + DTI W0 (dQ/dk) / (D dsigma) (the central vertical
    flux is exact for it)."""
    c = F.Model.build(0, nx=21, ny=7, kb=8, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    p = F.Model.case_defaults(0).p
    p.horizontal_closure = 0
    m = F.Model.from_mesh(x, y, nv, np.full(c.m, 12.0), 8, [], p)
    m.cold_start()
    m.field("nn_hvc")[:] = 3.0
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    o = Oracle(m)
    vx, D, U0, ax = m.field("vx"), 12.0, 0.3, 2e-8
    q = np.zeros((kb, m1))
    q[:, 1:] = 3e-4 + ax * vx[1:]
    u = np.zeros((kb, n1))
    u[:-1, 1:] = U0
    for nm, arr in (("q2", q.ravel()), ("u", u.ravel()), ("v", np.zeros(kb * n1)), ("wts", np.zeros(kb * m1)),
                    ("d", np.full(m1, D)), ("dt", np.full(m1, D)), ("dt1", np.full(n1, D))):
        o.put(nm, arr)
    o.routine("adv_q")
    qf = o.get("q2f").reshape(kb, m1)
    isonb = m.field("isonb")
    nbsn, off, ntsn = m.field("nbsn"), m.field("nbsn_off"), m.field("ntsn")
    deep = np.array([i for i in range(1, m1) if isonb[i] == 0 and
                     all(isonb[j] == 0 for j in nbsn[off[i]:off[i] + ntsn[i]])])
    assert len(deep) > 30
    assert np.max(np.abs(qf[1:-1, deep] - (q[1:-1, deep] - dti * U0 * ax))) < 1e-16
    # vertical
    W0, beta = 1e-4, 1e-5
    q[:, 1:] = (3e-4 + beta * np.arange(kb))[:, None]
    for nm, arr in (("q2", q.ravel()), ("u", np.zeros(kb * n1)), ("wts", np.full(kb * m1, W0))):
        o.put(nm, arr)
    o.routine("adv_q")
    qf = o.get("q2f").reshape(kb, m1)
    dsig = m.field("dz").reshape(kb, m1)[0, 1]
    assert np.max(np.abs(qf[1:-1, 1:] - (q[1:-1, 1:] + dti * W0 * beta / (D * dsig)))) < 1e-13


def test_adv_t_linear_field(built):
    """ADV_T (adv_t.F:34-1077) under the same flat-bottom set-up as ADV_Q:
    a temperature linear in x carried by a uniform U0 gives exactly
    TF1 = T1 - DTI U0 dT/dx at nodes with interior rings, every layer."""
    c = F.Model.build(0, nx=21, ny=7, kb=8, river=0)
    x, y = c.field("vx")[1:].copy(), c.field("vy")[1:].copy()
    nv = c.field("nv").reshape(3, c.n + 1)[:, 1:].T.copy()
    p = F.Model.case_defaults(0).p
    p.horizontal_closure = 0
    m = F.Model.from_mesh(x, y, nv, np.full(c.m, 12.0), 8, [], p)
    m.cold_start()
    m.field("nn_hvc")[:] = 3.0
    kb, m1, n1 = m.kb, m.m + 1, m.n + 1
    dti = m.hdr.p.dte * m.hdr.p.isplit
    o = Oracle(m)
    vx, D, U0, ax = m.field("vx"), 12.0, 0.3, 2e-5
    t1 = np.zeros((kb, m1))
    t1[:-1, 1:] = 10.0 + ax * vx[1:]
    u = np.zeros((kb, n1))
    u[:-1, 1:] = U0
    for nm, arr in (("t1", t1.ravel()), ("tmean1", np.zeros(kb * m1)), ("u", u.ravel()),
                    ("v", np.zeros(kb * n1)), ("wts", np.zeros(kb * m1)), ("d", np.full(m1, D)),
                    ("dt", np.full(m1, D)), ("dtfa", np.full(m1, D)), ("dt1", np.full(n1, D))):
        o.put(nm, arr)
    o.routine("adv_ts", 1)
    tf = o.get("tf1").reshape(kb, m1)
    isonb = m.field("isonb")
    nbsn, off, ntsn = m.field("nbsn"), m.field("nbsn_off"), m.field("ntsn")
    deep = np.array([i for i in range(1, m1) if isonb[i] == 0 and
                     all(isonb[j] == 0 for j in nbsn[off[i]:off[i] + ntsn[i]])])
    assert len(deep) > 30
    assert np.max(np.abs(tf[:-1, deep] - (t1[:-1, deep] - dti * U0 * ax))) < 1e-12
