// fvgpu_kernels_2d.h -- the barotropic (external) mode on the device.
//
// One Runge-Kutta stage of EXTERNAL_STEP (FVCOM_source/external_step.F:159-438)
// is three kernels:
//   k_ext_elevation  node-parallel: EXTEL_EDGE (extel_edge.F:32-229) as a
//                    gather of each node's control-volume sub-edge terms
//                    (formed per sub-edge by the previous stage's
//                    k_ext_momentum, or by k_ext_phi), the
//                    OBC tide and RK blend (BCOND_GCN(1,K), bcond_gcn.F:116;
//                    BCOND_ASL, mod_obcs.F:509-592), the OBC flux of FLUX_OBN
//                    (mod_obcs.F:1181-1215), D = H + EL, DEPTH_CHECK, and the
//                    stage-3 EGF / stage-4 UARD_OBCN accumulations.
//   k_advave_edge    edge-parallel: the ADVAVE_EDGE_GCN edge terms
//                    (advave_edge_gcn.F:546-884) of the stage input, once per edge;
//   k_ext_momentum   element-parallel: N2E2D of the new elevation, the
//                    element's three edge terms gathered in edge order,
//                    EXTUV_EDGE (extuv_edge.F:33-196),
//                    the wall projection BCOND_GCN(2,K) (bcond_gcn.F:325-420),
//                    D1 = H1 + EL1, the stage-3 UARD/VARD accumulation and
//                    the EXTEL terms of its sub-edges for the next stage.
// The per-node and per-element gather lists are sorted by sub-edge / edge
// index, so every sum is formed in the order of the reference's scatter loop.
#pragma once
#include "fvgpu_common.h"
#include "../../include/fv_detmath.h"

struct Ext2DState {      // one buffer set of the 2-D state
  double *el, *d;        // nodes
  double *ua, *va, *el1, *d1;  // elements
};

// Per edge (and layer) momentum-flux terms of the edge loops of
// ADVAVE_EDGE_GCN / ADVECTION_EDGE_GCN / ADV_UV_EDGE_GCN, written once per
// edge by an edge kernel and gathered by the two adjacent elements.
template <typename T>
struct __align__(16) EdgeFlux4T {
  T xadv, yadv, fx, fy;
};

struct ExtArgs {
  Ext2DState cur;        // stage input
  Ext2DState base;       // RK base (ELRK, UARK, VARK, ELRK1)
  Ext2DState nxt;        // stage output (may alias base: pointwise use only)
  int K;                 // RK stage 1..4
  double alpha;          // ALPHA_RK(K)
  double *xflux_obcn, *fobn, *uard_obcn, *egf, *fluxobn, *uard, *vard;
  const double *adx2d, *ady2d, *drx2d, *dry2d, *wubot, *wvbot;
  const double *qdis;    // QDIS (internal-mode river discharge)
  void* ef;              // per-edge ADVAVE terms of the stage input, EdgeFlux4T<FT>
  double* elij;          // per-edge mean elevation of the stage input
  double* phi;           // per sub-edge EXTEL terms (read for the stage input,
                         // written by k_ext_momentum for the next stage)
  const double *ws2x, *ws2y;   // wind stress of the external step (before RAMP*0.001)
  double *elm1, *elm2;         // the elevation one and two RK stages back (Orlanski OBC)
  DevClock* clk;
};

// PSIMPLE_DRAG (mod_bulk.F:58-90): wind stress of a wind speed (Large and
// Pond drag coefficient, 1.2 kg/m^3 air), on the host and the device.
__host__ __device__ __forceinline__ void psimple_drag(double tx, double ty, double& sx, double& sy) {
  const double wds = sqrt(tx * tx + ty * ty);
  double cd = (double)1.2E-3f;
  if (wds >= 11.0) cd = (0.49 + 0.065 * wds) * 1.E-3;
  if (wds >= 25.0) cd = (0.49 + 0.065 * 25.0) * 1.E-3;
  sx = 1.2 * cd * tx * wds;
  sy = 1.2 * cd * ty * wds;
}

// BCOND_GCN(9) wind of EXTERNAL_STEP iext (bcond_gcn.F:878-895): the speed
// interpolated linearly in time between the internal step's start (u0, v0)
// and end (u1, v1) to ExtTime, fraction iext/ISPLIT, then its stress.
__global__ void k_ext_wind(DevGrid g, const double* u0, const double* v0, const double* u1, const double* v1,
                           int iext, double* sx, double* sy) {
  const int i = (int)xcd_thread() + 1;
  if (i > g.N) return;
  const double frac = (double)iext / (double)g.isplit;
  const double uu = u0[i] + frac * (u1[i] - u0[i]);
  const double vv = v0[i] + frac * (v1[i] - v0[i]);
  psimple_drag(uu, vv, sx[i], sy[i]);
}

// Wall projection of a velocity pair (BCOND_GCN cases 2/3/5): the element's
// case, found once (wall_prep), then applied to any number of pairs of its
// column (wall_apply).  Case 3 zeroes the pair; on a solid wall (ISBCE 1)
// an element that is the first or last of a river node's NBVE ring keeps
// only the inflowing part along the river, any other the component along
// the wall.
struct WallOp {
  int kind;   // 0: unchanged, 1: river (p, q = COS, SIN of its angle), 2: wall (p, q = SIN, COS of ALPHA), 3: zero
  double p, q;
};

__device__ __forceinline__ WallOp wall_prep(const DevGrid& g, int i) {
  const int b = g.isbce[i];
  if (b == 3) return {3, 0.0, 0.0};
  if (b != 1) return {0, 0.0, 0.0};
  // the first river (in river order) whose mouth is this cell: the first
  // and last cells around a node river's node, an edge river's own cell
  if (g.re_off[i] < g.re_off[i + 1]) {
    const int j = g.re_riv[g.re_off[i]];
    return {1, g.cos_angleq[j], g.sin_angleq[j]};
  }
  return {2, g.sin_alpha[i], g.cos_alpha[i]};
}

__device__ __forceinline__ void wall_apply(const WallOp& w, double& u, double& v) {
  if (w.kind == 3) {
    u = 0.0;
    v = 0.0;
  } else if (w.kind == 1) {
    const double ca = w.p, sa = w.q;
    double untmp = u * ca + v * sa;
    const double vntmp = -u * sa + v * ca;
    untmp = fmax(untmp, 0.0);
    u = untmp * ca - vntmp * sa;
    v = untmp * sa + vntmp * ca;
  } else if (w.kind == 2) {
    const double s = w.p, c = w.q;
    const double ui = u * (s * s) - v * s * c;
    const double vi = -u * s * c + v * (c * c);
    u = ui;
    v = vi;
  }
}

__device__ __forceinline__ void wall_project(const DevGrid& g, int i, double& u, double& v) {
  wall_apply(wall_prep(g, i), u, v);
}

// Spectral tide at OBC node index `obc` (BCOND_ASL, mod_obcs.F:552-582).
__device__ __forceinline__ double tide_force(const DevGrid& g, int obc, double rk_time, double ramp) {
  const double PI2 = 6.283185307179586;
  double force = 0.0;
  for (int j = 0; j < g.NTIDE; ++j) {
    const double phai_ij = g.phai[(long)j * g.IOBCN + obc] * PI2 / 360.0;
    const double time_flt = rk_time * PI2 / g.period[j];
    force = g.apt[(long)j * g.IOBCN + obc] * fv_cos(time_flt - phai_ij) + force;
  }
  force = force + g.emean[obc];
  return force * ramp;
}

// FLUX_OBN's flux defect of OBC node i (mod_obcs.F:1196) and, in the last
// stage, its UARD_OBCN share (external_step.F:423-430).
__device__ __forceinline__ void ext_obc_flux(const DevGrid& g, const ExtArgs& a, int i, int obc, double elf,
                                             double elrk, double xobc) {
  a.fobn[obc] = -(elf - elrk) * g.art1[i] / (a.alpha * g.dte) - xobc;
  if (a.K == 4) {
    const double tmp = -(elf - elrk) * g.art1[i] / g.dte - xobc;
    a.uard_obcn[obc] = a.uard_obcn[obc] + tmp / (double)g.isplit;
  }
}

// The new elevation of node i: EL, D = H + EL, the stage-3 EGF sum and
// DEPTH_CHECK.
__device__ __forceinline__ void ext_node_store(const DevGrid& g, const ExtArgs& a, int i, double elf) {
  const double dnew = g.h[i] + elf;
  a.nxt.el[i] = elf;
  a.nxt.d[i] = dnew;
  if (a.K == 3) a.egf[i] = a.egf[i] + elf / (double)g.isplit;
  if (i >= 1 && dnew < g.min_depth) a.clk->depth_fail = 1;
}

template <int SMAX>
__global__ void k_ext_elevation(DevGrid g, ExtArgs a) {
  const int i = (int)xcd_thread();
  if (i > g.M) return;
  // EXTEL_EDGE sum over the node's sub-edges in sub-edge order.  The term of
  // sub-edge c, D1*(-UA*DLTYE + VA*DLTXE) of its element for side IB, is
  // exactly the negative for side IA (DLTXE/DLTYE negated; round to
  // nearest is symmetric), so the stored term is added or subtracted.
  // SMAX >= the node's sub-edge count: the slots past it re-read slot 0,
  // so all list and term loads are unconditional and issue together.
  const int n = g.cv_cnt[i];
  int ce[SMAX];
  static_for<SMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    ce[j] = g.cvl_ent[(long)(j < n ? j : 0) * g.m1 + i];
  });
  double t[SMAX];
  static_for<SMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    t[j] = a.phi[ce[j] >> 1];
  });
  double xflux = 0.0;
  static_for<SMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < n) xflux = (ce[j] & 1) ? xflux + t[j] : xflux - t[j];
  });
  const int obc = g.node_obc[i];
  double xobc = 0.0;
  if (obc >= 0) { xobc = xflux; xflux = 0.0; a.xflux_obcn[obc] = xobc; }
  const int rv = g.node_riv[i];
  if (rv >= 0) {
    for (int j = rv; j < g.NUMQBC; ++j) {
      double w;
      if (river_share(g, j, i, w)) xflux = xflux - a.qdis[j] * w;
    }
  }
  const double dtk = a.alpha * g.dte;
  const double elrk = a.base.el[i];
  double elf = elrk - dtk * xflux / g.art1[i];
  if (obc >= 0) {
    const int ty = g.type_obc[obc];
    if (ty >= 5) return;   // radiation conditions: k_ext_obc_radiate, once every interior ELF is formed
    if (ty <= 2) {         // BCOND_ASL, then the RK blend (external_step.F:262-266)
      if (g.obc_elev) {
        const double rk_time = a.clk->ext_time + g.dte * (a.alpha - 1.0);
        elf = tide_force(g, obc, rk_time, a.clk->ramp_ext);
      }
      elf = elrk + a.alpha * (elf - elrk);
    } else {               // BCOND_ASL_CLP (mod_obcs.F:599-617)
      elf = 0.0;
    }
    ext_obc_flux(g, a, i, obc, elf, elrk, xobc);
  }
  ext_node_store(g, a, i, elf);
}

// BCOND_GWI (TYPE_OBC 5/6, mod_obcs.F:624-658) and BCOND_BKI (7/8,
// mod_obcs.F:665-699): the OBC node's ELF from its interior neighbour
// NEXT_OBC's new ELF (formed by k_ext_elevation, which leaves these nodes
// alone); BCOND_ORE (9/10, mod_obcs.F:706-737): from the stage input EL
// and the elevations ELM1/ELM2 one and two stages back.  One thread per
// OBC node.
__global__ void k_ext_obc_radiate(DevGrid g, ExtArgs a) {
  const int obc = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (obc >= g.IOBCN) return;
  const int ty = g.type_obc[obc];
  if (ty < 5) return;
  const int i1 = g.i_obc_n[obc], i2 = g.next_obc[obc];
  const double elrk = a.base.el[i1];
  double elf;
  if (ty >= 9) {
    const int i3 = g.next_obc2[obc];
    const double cl = (a.elm2[i2] - a.cur.el[i2]) / (a.cur.el[i2] + a.elm2[i2] - 2.0 * a.elm1[i3]);
    double mu;                            // a NaN CL (0/0 at rest) gives MU = 0, as in the reference
    if (cl >= 1.0) mu = 1.0;
    else if (cl > 0.0 && cl < 1.0) mu = cl;
    else mu = 0.0;
    elf = (a.elm1[i1] * (1.0 - mu) + 2.0 * mu * a.cur.el[i2]) / (1.0 + mu);
  } else {
    const double cc = sqrt(g.grav_n[i1] * g.h[i1]) * g.dte / g.dltn_obc[obc] * a.alpha;
    const double cp = cc + 1.0;
    elf = ty <= 6 ? (cc * a.nxt.el[i2] + elrk) / cp : (cc * a.nxt.el[i2] + elrk * (1.0 - g.dte / 10800.0)) / cp;
  }
  ext_obc_flux(g, a, i1, obc, elf, elrk, a.xflux_obcn[obc]);
  ext_node_store(g, a, i1, elf);
}

// ASSIGN_ELM1_TO_ELM2 (mod_obcs.F:178-199) at the end of a stage
// (external_step.F:333): ELM2 = ELM1, ELM1 = the stage input EL.  Run
// after k_ext_obc_radiate, the only reader.
__global__ void k_elm_shift(DevGrid g, ExtArgs a) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > g.M) return;
  a.elm2[i] = a.elm1[i];
  a.elm1[i] = a.cur.el[i];
}

// ADVAVE_EDGE_GCN edge loop body (advave_edge_gcn.F:546-884) for edge e of
// the state (ua, va, d, el), one thread per edge: least-squares velocity
// gradients on both sides, face velocities, upwind advective flux, the
// Smagorinsky stresses; the side-dependent factors (EPOR, IUCP, ISBC, the
// pressure term) are applied by the gathering element.  Arithmetic in the
// flux type FT (FP64: the reference's expressions bit for bit).
template <typename FT>
__global__ __launch_bounds__(256) void k_advave_edge(DevGrid g, const double* ua_, const double* va_,
                                                     const double* d, const double* el,
                                                     void* ef_, double* elij_out) {
  const int e = (int)xcd_thread();
  if (e < 1 || e > g.NE) return;
  EdgeFlux4T<FT>* ef = (EdgeFlux4T<FT>*)ef_;
  const FT half = (FT)0.5, one = (FT)1.0;
  const int ia = IEC_(e, 1), ib = IEC_(e, 2), j1 = IENODE_(e, 1), j2 = IENODE_(e, 2);
  const FT dij = half * ((FT)d[j1] + (FT)d[j2]);
  const double elij = 0.5 * (el[j1] + el[j2]);
  auto U = [&](int i) { return (FT)ua_[i]; };
  auto V = [&](int i) { return (FT)va_[i]; };
  auto A1 = [&](int i, int j) { return (FT)A1U_(i, j); };
  auto A2 = [&](int i, int j) { return (FT)A2U_(i, j); };
  int k1 = NBE_(ia, 1), k2 = NBE_(ia, 2), k3 = NBE_(ia, 3);
  const FT cofa1 = A1(ia, 1) * U(ia) + A1(ia, 2) * U(k1) + A1(ia, 3) * U(k2) + A1(ia, 4) * U(k3);
  const FT cofa2 = A2(ia, 1) * U(ia) + A2(ia, 2) * U(k1) + A2(ia, 3) * U(k2) + A2(ia, 4) * U(k3);
  const FT cofa5 = A1(ia, 1) * V(ia) + A1(ia, 2) * V(k1) + A1(ia, 3) * V(k2) + A1(ia, 4) * V(k3);
  const FT cofa6 = A2(ia, 1) * V(ia) + A2(ia, 2) * V(k1) + A2(ia, 3) * V(k2) + A2(ia, 4) * V(k3);
  FT xij = (FT)(g.xijc[e] - g.xc[ia]), yij = (FT)(g.yijc[e] - g.yc[ia]);
  const FT uij1 = U(ia) + cofa1 * xij + cofa2 * yij;
  const FT vij1 = V(ia) + cofa5 * xij + cofa6 * yij;
  k1 = NBE_(ib, 1); k2 = NBE_(ib, 2); k3 = NBE_(ib, 3);
  const FT cofa3 = A1(ib, 1) * U(ib) + A1(ib, 2) * U(k1) + A1(ib, 3) * U(k2) + A1(ib, 4) * U(k3);
  const FT cofa4 = A2(ib, 1) * U(ib) + A2(ib, 2) * U(k1) + A2(ib, 3) * U(k2) + A2(ib, 4) * U(k3);
  const FT cofa7 = A1(ib, 1) * V(ib) + A1(ib, 2) * V(k1) + A1(ib, 3) * V(k2) + A1(ib, 4) * V(k3);
  const FT cofa8 = A2(ib, 1) * V(ib) + A2(ib, 2) * V(k1) + A2(ib, 3) * V(k2) + A2(ib, 4) * V(k3);
  xij = (FT)(g.xijc[e] - g.xc[ib]); yij = (FT)(g.yijc[e] - g.yc[ib]);
  const FT uij2 = U(ib) + cofa3 * xij + cofa4 * yij;
  const FT vij2 = V(ib) + cofa7 * xij + cofa8 * yij;
  const FT uij = half * (uij1 + uij2), vij = half * (vij1 + vij2);
  const FT dlx = (FT)g.dltxc[e], dly = (FT)g.dltyc[e];
  const FT un_tmp = -uij * dly + vij * dlx;
  const FT viscof1 = (FT)g.art[ia] * sqrt(sq(cofa1) + sq(cofa6) + half * sq(cofa2 + cofa5));
  const FT viscof2 = (FT)g.art[ib] * sqrt(sq(cofa3) + sq(cofa8) + half * sq(cofa4 + cofa7));
  const FT cca = (FT)g.cc_hvc[ia], ccb = (FT)g.cc_hvc[ib];
  const FT viscof = ((FT)g.fact * half * (viscof1 * cca + viscof2 * ccb) +
                     (FT)g.fm1 * half * (cca + ccb)) / (FT)g.hprnu;
  const FT txxij = (cofa1 + cofa3) * viscof;
  const FT tyyij = (cofa6 + cofa8) * viscof;
  const FT txyij = half * (cofa2 + cofa4 + cofa5 + cofa7) * viscof;
  EdgeFlux4T<FT> r;
  r.fx = dij * (txxij * dly - txyij * dlx);
  r.fy = dij * (txyij * dly - tyyij * dlx);
  const FT sgn = fsign1(un_tmp);
  r.xadv = dij * un_tmp * ((one - sgn) * uij2 + (one + sgn) * uij1) * half;
  r.yadv = dij * un_tmp * ((one - sgn) * vij2 + (one + sgn) * vij1) * half;
  ef[e] = r;
  elij_out[e] = elij;
}

// ADVAVE_EDGE_GCN for element i: its three edges' terms gathered in edge
// order (the reference's scatter order), OBC flux and rivers
// (advave_edge_gcn.F:746-992).  The element sums are FP64.
template <typename FT>
__device__ __forceinline__ void advave_gather(const DevGrid& g, int i, const void* ef_, const double* elij,
                                              const double* ua, const double* va, const double* d1,
                                              const double* fluxobn, const double* qdis,
                                              double& xf, double& yf, double& px, double& py) {
  const EdgeFlux4T<FT>* ef = (const EdgeFlux4T<FT>*)ef_;
  xf = 0.0; yf = 0.0; px = 0.0; py = 0.0;
  const double ep = g.epor[i], iu = (double)g.iucp[i], gdi = g.grav_e[i] * d1[i];
  // the three edges' loads first, then the sums with the side chosen by a
  // select: no branch separates the loads, so they issue together
  int ent[3];
  EdgeFlux4T<FT> r[3];
  double isbc_tmp[3], el[3], dlx[3], dly[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) ent[q] = g.ed_ent[3 * i + q];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int e = ent[q] >> 1;
    r[q] = ef[e];
    isbc_tmp[q] = (double)g.isbc[e];
    el[q] = elij[e];
    dlx[q] = g.dltxc[e];
    dly[q] = g.dltyc[e];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double xadv = (double)r[q].xadv, yadv = (double)r[q].yadv, fx = (double)r[q].fx, fy = (double)r[q].fy;
    const double gd = gdi * el[q];
    const double tx = (xadv + fx * ep) * (1.0 - isbc_tmp[q]) * iu;
    const double ty = (yadv + fy * ep) * (1.0 - isbc_tmp[q]) * iu;
    const double gy = gd * dly[q], gx = gd * dlx[q];
    const bool ib = (ent[q] & 1) != 0;
    xf = ib ? xf - tx : xf + tx;
    yf = ib ? yf - ty : yf + ty;
    px = ib ? px + gy : px - gy;
    py = ib ? py - gx : py + gx;
  }
  if (g.isbce[i] == 2) {
    xf = (xf + fluxobn[i] * ua[i]) * (double)g.iucp[i];
    yf = (yf + fluxobn[i] * va[i]) * (double)g.iucp[i];
  }
  for (int p = g.re_off[i]; p < g.re_off[i + 1]; ++p) {
    // a node river's momentum is split between two cells, an edge river's
    // enters its cell whole (advave_edge_gcn.F:971-992)
    const int k = g.re_riv[p];
    const double fq = g.icellq[k] > 0 ? 1.0 : 0.5;
    const double vlctyq = qdis[k] / g.qarea[k];
    xf = xf - fq * qdis[k] * vlctyq * g.cos_angleq[k];
    yf = yf - fq * qdis[k] * vlctyq * g.sin_angleq[k];
  }
}

// EXTEL terms of element i's three sub-edges (side IB), element-major.
__device__ __forceinline__ void ext_phi_store(const DevGrid& g, int i, double ua, double va, double d1,
                                              double* phi) {
  const long p0 = 3L * (i - 1) + 1;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const long q = (long)j * g.n1 + i;
    phi[p0 + j] = d1 * (-ua * g.ecv_dy[q] + va * g.ecv_dx[q]);
  }
}

// The EXTEL terms of a 2-D state (before its first stage).
__global__ __launch_bounds__(256) void k_ext_phi(DevGrid g, const double* ua, const double* va, const double* d1,
                                                 double* phi) {
  const int i = (int)xcd_thread() + 1;
  if (i > g.N) return;
  ext_phi_store(g, i, ua[i], va[i], d1[i], phi);
}

template <typename FT>
__global__ void k_ext_momentum(DevGrid g, ExtArgs a) {
  const int i = (int)xcd_thread();
  if (i < 1 || i > g.N) return;
  const double one_third = 1.0 / 3.0;
  const double elf1 = one_third * (a.nxt.el[NV_(i, 1)] + a.nxt.el[NV_(i, 2)] + a.nxt.el[NV_(i, 3)]);
  // FLUX_OBN: contributions to this open-boundary element, in OBC order
  if (g.isbce[i] == 2) {
    double f = 0.0;
    for (int p = g.fo_off[i]; p < g.fo_off[i + 1]; ++p) {
      const int ent = g.fo_obc[p], ob = ent >> 1, slot = ent & 1;
      f = f + g.fluxf_obc[(long)slot * g.IOBCN + ob] * a.fobn[ob];
    }
    a.fluxobn[i] = f;
  }
  double xf, yf, px, py;
  advave_gather<FT>(g, i, a.ef, a.elij, a.cur.ua, a.cur.va, a.cur.d1, a.fluxobn, a.qdis, xf, yf, px, py);
  // EXTUV_EDGE
  const double ua = a.cur.ua[i], va = a.cur.va[i], d1 = a.cur.d1[i], art = g.art[i], cor = g.cor[i];
  const double ramp = a.clk->ramp_ext;
  const double wus2 = g.wind_on ? a.ws2x[i] * ramp * 0.001 : 0.0;
  const double wvs2 = g.wind_on ? a.ws2y[i] * ramp * 0.001 : 0.0;
  const double resx = a.adx2d[i] + xf + a.drx2d[i] + px - cor * va * d1 * art - (wus2 + a.wubot[i]) * art;
  const double resy = a.ady2d[i] + yf + a.dry2d[i] + py + cor * ua * d1 * art - (wvs2 + a.wvbot[i]) * art;
  const double h1 = g.h1[i], elrk1 = a.base.el1[i];
  double uaf = (a.base.ua[i] * (h1 + elrk1) - a.alpha * g.dte * resx / art) / (h1 + elf1);
  double vaf = (a.base.va[i] * (h1 + elrk1) - a.alpha * g.dte * resy / art) / (h1 + elf1);
  uaf = uaf / (1.0 + g.cc_sponge[i] * uaf * uaf);
  vaf = vaf / (1.0 + g.cc_sponge[i] * vaf * vaf);
  wall_project(g, i, uaf, vaf);
  const double d1n = h1 + elf1;
  a.nxt.ua[i] = uaf;
  a.nxt.va[i] = vaf;
  a.nxt.el1[i] = elf1;
  a.nxt.d1[i] = d1n;
  ext_phi_store(g, i, uaf, vaf, d1n, a.phi);
  if (a.K == 3) {
    a.uard[i] = a.uard[i] + uaf * d1n;
    a.vard[i] = a.vard[i] + vaf * d1n;
  }
}

// Start of EXTERNAL_STEP: clock and RAMP (external_step.F:83-88).
__global__ void k_ext_clock(DevClock* c, int iext) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  c->iext = iext;
  c->ext_time = c->time0 + (double)(c->iint - 1) * c->dti + (double)iext * c->dte;
  if (c->iramp != 0) {
    const double tmp = (double)(c->iint - 1) + (double)iext / (double)c->isplit;
    c->ramp_ext = fv_tanh(tmp / (double)c->iramp);
  } else {
    c->ramp_ext = 1.0;
  }
}
