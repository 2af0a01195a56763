// fvgpu_kernels_edge.h -- the 3-D momentum edge loops in two phases.
//
// ADVECTION_EDGE_GCN and ADV_UV_EDGE_GCN loop over edges and scatter each
// edge's flux to its two elements.  Here k_uvedge evaluates every edge once
// per layer (one thread per edge, the edge's geometry and least-squares
// coefficients in registers across the layers) into EdgeFlux4 records, and
// the element kernels gather their three edges in increasing edge index --
// the reference's scatter order -- so the sums are bit-identical.
#pragma once
#include "fvgpu_kernels_cv.h"

// MODE 0: ADVECTION_EDGE_GCN (advection_edge_gcn.F:397-734);
// MODE 1: ADV_UV_EDGE_GCN (adv_uv_edge_gcn.F:512-848), which also keeps the
// edge viscosity for VISCOFM.  The two differ only in how the upwind flux is
// grouped, which is kept as each reference file writes it.
template <int MODE, typename FT>
__global__ __launch_bounds__(256) void k_uvedge(DevGrid g, IntArgs a, const double* U, const double* V) {
  const int e = (int)xcd_thread();
  if (e < 1 || e > g.NE) return;
  EdgeFlux4T<FT>* ef3 = (EdgeFlux4T<FT>*)a.ef3;
  FT* efvisc = (FT*)a.efvisc;
  const FT half = (FT)0.5, one = (FT)1.0;
  const int ia = IEC_(e, 1), ib = IEC_(e, 2), j1 = IENODE_(e, 1), j2 = IENODE_(e, 2);
  const int k1 = NBE_(ia, 1), k2 = NBE_(ia, 2), k3 = NBE_(ia, 3), k4 = NBE_(ib, 1), k5 = NBE_(ib, 2), k6 = NBE_(ib, 3);
  const FT xija = (FT)(g.xijc[e] - g.xc[ia]), yija = (FT)(g.yijc[e] - g.yc[ia]);
  const FT xijb = (FT)(g.xijc[e] - g.xc[ib]), yijb = (FT)(g.yijc[e] - g.yc[ib]);
  const FT a11 = (FT)A1U_(ia, 1), a12 = (FT)A1U_(ia, 2), a13 = (FT)A1U_(ia, 3), a14 = (FT)A1U_(ia, 4);
  const FT a21 = (FT)A2U_(ia, 1), a22 = (FT)A2U_(ia, 2), a23 = (FT)A2U_(ia, 3), a24 = (FT)A2U_(ia, 4);
  const FT b11 = (FT)A1U_(ib, 1), b12 = (FT)A1U_(ib, 2), b13 = (FT)A1U_(ib, 3), b14 = (FT)A1U_(ib, 4);
  const FT b21 = (FT)A2U_(ib, 1), b22 = (FT)A2U_(ib, 2), b23 = (FT)A2U_(ib, 3), b24 = (FT)A2U_(ib, 4);
  const FT dlx = (FT)g.dltxc[e], dly = (FT)g.dltyc[e];
  const FT arta = (FT)g.art[ia], artb = (FT)g.art[ib], cca = (FT)g.cc_hvc[ia], ccb = (FT)g.cc_hvc[ib];
  const FT visc0 = (FT)g.fm1 * half * (cca + ccb);
  const FT fact = (FT)g.fact, hprnu = (FT)g.hprnu;
  const double dtj1 = a.dt[j1], dtj2 = a.dt[j2];
  const FT isbc_tmp = (FT)g.isbc[e];
  const long es = (long)g.NE + 1;
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.n1;
    const FT dij = (FT)(0.5 * (dtj1 * DZ(j1, k) + dtj2 * DZ(j2, k)));
    const FT uia = (FT)U[ko + ia], vIA = (FT)V[ko + ia], uib = (FT)U[ko + ib], vIB = (FT)V[ko + ib];
    const FT u1 = (FT)U[ko + k1], u2 = (FT)U[ko + k2], u3 = (FT)U[ko + k3];
    const FT u4 = (FT)U[ko + k4], u5 = (FT)U[ko + k5], u6 = (FT)U[ko + k6];
    const FT v1 = (FT)V[ko + k1], v2 = (FT)V[ko + k2], v3 = (FT)V[ko + k3];
    const FT v4 = (FT)V[ko + k4], v5 = (FT)V[ko + k5], v6 = (FT)V[ko + k6];
    const FT cofa1 = a11 * uia + a12 * u1 + a13 * u2 + a14 * u3;
    const FT cofa2 = a21 * uia + a22 * u1 + a23 * u2 + a24 * u3;
    const FT cofa5 = a11 * vIA + a12 * v1 + a13 * v2 + a14 * v3;
    const FT cofa6 = a21 * vIA + a22 * v1 + a23 * v2 + a24 * v3;
    const FT uij1 = uia + cofa1 * xija + cofa2 * yija;
    const FT vij1 = vIA + cofa5 * xija + cofa6 * yija;
    const FT cofa3 = b11 * uib + b12 * u4 + b13 * u5 + b14 * u6;
    const FT cofa4 = b21 * uib + b22 * u4 + b23 * u5 + b24 * u6;
    const FT cofa7 = b11 * vIB + b12 * v4 + b13 * v5 + b14 * v6;
    const FT cofa8 = b21 * vIB + b22 * v4 + b23 * v5 + b24 * v6;
    const FT uij2 = uib + cofa3 * xijb + cofa4 * yijb;
    const FT vij2 = vIB + cofa7 * xijb + cofa8 * yijb;
    const FT uij = half * (uij1 + uij2), vij = half * (vij1 + vij2);
    const FT viscof1 = arta * sqrt(sq(cofa1) + sq(cofa6) + half * sq(cofa2 + cofa5));
    const FT viscof2 = artb * sqrt(sq(cofa3) + sq(cofa8) + half * sq(cofa4 + cofa7));
    const FT viscof = (fact * half * (viscof1 * cca + viscof2 * ccb) + visc0) / hprnu;
    const FT txxij = (cofa1 + cofa3) * viscof, tyyij = (cofa6 + cofa8) * viscof;
    const FT txyij = half * (cofa2 + cofa4 + cofa5 + cofa7) * viscof;
    const FT fxx = dij * (txxij * dly - txyij * dlx);
    const FT fyy = dij * (txyij * dly - tyyij * dlx);
    EdgeFlux4T<FT> r;
    if (MODE == 0) {
      const FT un_tmp = vij * dlx - uij * dly;
      const FT sgn = fsign1(un_tmp);
      r.xadv = dij * un_tmp * ((one - sgn) * uij2 + (one + sgn) * uij1) * half;
      r.yadv = dij * un_tmp * ((one - sgn) * vij2 + (one + sgn) * vij1) * half;
    } else {
      const FT exflux = dij * (-uij * dly + vij * dlx);
      const FT sgn = fsign1(exflux);
      r.xadv = exflux * ((one - sgn) * uij2 + (one + sgn) * uij1) * half;
      r.yadv = exflux * ((one - sgn) * vij2 + (one + sgn) * vij1) * half;
      efvisc[(long)(k - 1) * es + e] = viscof;
    }
    r.fx = fxx + (FT)3.0 * fxx * isbc_tmp;
    r.fy = fyy + (FT)3.0 * fyy * isbc_tmp;
    ef3[(long)(k - 1) * es + e] = r;
  }
}

// ADVECTION_EDGE_GCN summed over the column into ADX2D/ADY2D
// (internal_step.F:374-384): per layer the three edge terms in edge order,
// OBC elements zeroed, river momentum (advection_edge_gcn.F:600-734).
template <typename FT>
__global__ __launch_bounds__(256) void k_adv3d_gather(DevGrid g, IntArgs a) {
  const int i = (int)xcd_thread();
  if (i < 1 || i > g.N) return;
  int ee[3];
  double tp[3];
  const double ep = g.epor[i];
  for (int q = 0; q < 3; ++q) {
    ee[q] = g.ed_ent[3 * i + q];
    tp[q] = (double)(1 - g.isbc[ee[q] >> 1]) * ep;
  }
  const bool obc = g.isbce[i] == 2;
  const int r0 = g.re_off[i], r1 = g.re_off[i + 1];
  const long es = (long)g.NE + 1;
  double sx = 0.0, sy = 0.0;
  for (int k = 1; k <= g.KBM1; ++k) {
    const EdgeFlux4T<FT>* fk = (const EdgeFlux4T<FT>*)a.ef3 + (long)(k - 1) * es;
    double xf = 0.0, yf = 0.0;
    // loads first, side by select: no branch separates the three loads
    EdgeFlux4T<FT> r[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) r[q] = fk[ee[q] >> 1];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double xadv = (double)r[q].xadv, yadv = (double)r[q].yadv, fx = (double)r[q].fx, fy = (double)r[q].fy;
      const double ax = xadv * tp[q], ay = yadv * tp[q], bx = fx * ep, by = fy * ep;
      const bool ib = (ee[q] & 1) != 0;
      xf = ib ? (xf - ax) - bx : (xf + ax) + bx;
      yf = ib ? (yf - ay) - by : (yf + ay) + by;
    }
    if (obc) { xf = 0.0; yf = 0.0; }
    for (int p = r0; p < r1; ++p) {
      const int r = g.re_riv[p];
      const int j = g.inodeq[r];
      const double vlctyq = a.qdis[r] / g.qarea[r];
      const double vq = g.vqdist[(long)(k - 1) * g.NUMQBC + r];
      const double temp = 0.5 * a.qdis[r] * vq * vq * vlctyq / DZ(j, k);
      xf = xf - temp * g.cos_angleq[r];
      yf = yf - temp * g.sin_angleq[r];
    }
    sx = sx + xf;
    sy = sy + yf;
  }
  a.adx2d[i] = sx;
  a.ady2d[i] = sy;
}

// ADV_UV_EDGE_GCN per element column (adv_uv_edge_gcn.F:776-1194): the three
// edge terms in edge order with the internal barotropic pressure gradient
// from EGF, vertical advection, baroclinic pressure gradient and Coriolis,
// OBC zeroing, rivers, and the update of UF/VF and VISCOFM.
template <typename FT>
__global__ __launch_bounds__(256) void k_advuv_gather(DevGrid g, IntArgs a) {
  const int i = (int)xcd_thread();
  if (i > g.N) return;
  if (i == 0) {
    for (int k = 1; k <= g.KB; ++k) a.viscofm[EI(g, 0, k)] = 0.0;
    return;
  }
  int ee[3];
  double tp[3], elij[3], dlx[3], dly[3];
  const double ep = g.epor[i];
  for (int q = 0; q < 3; ++q) {
    ee[q] = g.ed_ent[3 * i + q];
    const int e = ee[q] >> 1;
    tp[q] = (double)(1 - g.isbc[e]) * ep;
    elij[q] = 0.5 * (a.egf[IENODE_(e, 1)] + a.egf[IENODE_(e, 2)]);
    dlx[q] = g.dltxc[e];
    dly[q] = g.dltyc[e];
  }
  const double art = g.art[i], dt1 = a.dt1[i], d1 = a.d1[i], cor = g.cor[i], ge = g.grav_e[i];
  const bool obc = g.isbce[i] == 2;
  const int r0 = g.re_off[i], r1 = g.re_off[i + 1];
  const double* U = a.u; const double* V = a.v; const double* W = a.w;
  const long es = (long)g.NE + 1;
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.n1;
    const EdgeFlux4T<FT>* fk = (const EdgeFlux4T<FT>*)a.ef3 + (long)(k - 1) * es;
    const FT* vk_ = (const FT*)a.efvisc + (long)(k - 1) * es;
    const double dzk = DZ1(i, k);
    double xf = 0.0, yf = 0.0, px = 0.0, py = 0.0, vm = 0.0;
    // loads first, side by select: no branch separates the six loads
    EdgeFlux4T<FT> r[3];
    FT vis[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      r[q] = fk[ee[q] >> 1];
      vis[q] = vk_[ee[q] >> 1];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double xadv = (double)r[q].xadv, yadv = (double)r[q].yadv, fx = (double)r[q].fx, fy = (double)r[q].fy;
      vm = vm + (double)vis[q];
      const double gd = ge * dt1 * dzk * elij[q];
      const double ax = xadv * tp[q], ay = yadv * tp[q], bx = fx * ep, by = fy * ep;
      const double gy = gd * dly[q], gx = gd * dlx[q];
      const bool ib = (ee[q] & 1) != 0;
      xf = ib ? (xf - ax) - bx : (xf + ax) + bx;
      yf = ib ? (yf - ay) - by : (yf + ay) + by;
      px = ib ? px + gy : px - gy;
      py = ib ? py - gx : py + gx;
    }
    xf = xf + px;
    yf = yf + py;
    // vertical advection (adv_uv_edge_gcn.F:927-946)
    const double uk = U[ko + i], vk = V[ko + i];
    // the layer's neighbours are loaded unconditionally (layer 1 reads
    // itself as its upper neighbour; level KB exists below KBM1) so that
    // they issue with the edge loads
    const long kp = ko + g.n1, km = k != 1 ? ko - g.n1 : ko;
    const double dzp = DZ1(i, k + 1), dzm = DZ1(i, k != 1 ? k - 1 : k);
    const double wk = W[ko + i], wp = W[kp + i];
    const double up = U[kp + i], vp = V[kp + i], um = U[km + i], vmm = V[km + i];
    double xfluxv, yfluxv;
    if (k == 1) {
      xfluxv = -wp * (uk * dzp + up * dzk) / (dzk + dzp);
      yfluxv = -wp * (vk * dzp + vp * dzk) / (dzk + dzp);
    } else if (k == g.KBM1) {
      xfluxv = wk * (uk * dzm + um * dzk) / (dzk + dzm);
      yfluxv = wk * (vk * dzm + vmm * dzk) / (dzk + dzm);
    } else {
      xfluxv = wk * (uk * dzm + um * dzk) / (dzk + dzm) - wp * (uk * dzp + up * dzk) / (dzk + dzp);
      yfluxv = wk * (vk * dzm + vmm * dzk) / (dzk + dzm) - wp * (vk * dzp + vp * dzk) / (dzk + dzp);
    }
    xf = xf + xfluxv * art + a.drhox[ko + i] - cor * vk * dt1 * dzk * art;
    yf = yf + yfluxv * art + a.drhoy[ko + i] + cor * uk * dt1 * dzk * art;
    if (obc) { xf = 0.0; yf = 0.0; }
    for (int p = r0; p < r1; ++p) {
      const int r = g.re_riv[p];
      const int j = g.inodeq[r];
      const double vlctyq = a.qdis[r] / g.qarea[r];
      const double vq = g.vqdist[(long)(k - 1) * g.NUMQBC + r];
      const double temp = 0.5 * a.qdis[r] * vq * vq * vlctyq / DZ(j, k);
      xf = xf - temp * g.cos_angleq[r];
      yf = yf - temp * g.sin_angleq[r];
    }
    a.uf[ko + i] = uk * dt1 / d1 - g.dti * xf / art / (d1 * dzk);
    a.vf[ko + i] = vk * dt1 / d1 - g.dti * yf / art / (d1 * dzk);
    a.viscofm[ko + i] = vm / art;
  }
  a.viscofm[EI(g, i, g.KB)] = 0.0 / art;
}

// ---------------------------------------------------------------------------
// Element-centric form of the two edge loops above: one thread per element
// column evaluates the terms of its three edges itself at every layer --
// each edge is evaluated by both of its elements, from the same operands in
// the same roles (IA / IB) and with the same expressions as k_uvedge, so
// both elements see the bits the edge kernel would have stored -- and sums
// them in edge order as the gathers do.  No per edge-layer buffer crosses
// HBM.  The element's own least-squares gradient is formed once per layer
// and serves as the IA or IB side of each of its edges; the other side's
// coefficients, geometry and neighbour list stay in registers across the
// layers.
template <typename FT>
struct EdgeSide {
  int o, on1, on2, on3;           // the other element and its NBE(:,1..3)
  int j1, j2;                     // IENODE(e,1..2)
  bool own_ib;                    // this element is IB of the edge
  FT oa1[4], oa2[4];              // A1U/A2U of the other element
  FT xija, yija, xijb, yijb, dlx, dly, arto, cco, isbc;
  double dtj1, dtj2;
};

template <typename FT>
__device__ __forceinline__ void edge_side_load(const DevGrid& g, const IntArgs& a, int i, int ent, EdgeSide<FT>& s) {
  const int e = ent >> 1;
  s.own_ib = (ent & 1) != 0;
  const int ia = IEC_(e, 1), ib = IEC_(e, 2);
  s.o = s.own_ib ? ia : ib;
  s.on1 = NBE_(s.o, 1); s.on2 = NBE_(s.o, 2); s.on3 = NBE_(s.o, 3);
  s.j1 = IENODE_(e, 1); s.j2 = IENODE_(e, 2);
#pragma unroll
  for (int j = 0; j < 4; ++j) { s.oa1[j] = (FT)A1U_(s.o, j + 1); s.oa2[j] = (FT)A2U_(s.o, j + 1); }
  s.xija = (FT)(g.xijc[e] - g.xc[ia]); s.yija = (FT)(g.yijc[e] - g.yc[ia]);
  s.xijb = (FT)(g.xijc[e] - g.xc[ib]); s.yijb = (FT)(g.yijc[e] - g.yc[ib]);
  s.dlx = (FT)g.dltxc[e]; s.dly = (FT)g.dltyc[e];
  s.arto = (FT)g.art[s.o]; s.cco = (FT)g.cc_hvc[s.o];
  s.isbc = (FT)g.isbc[e];
  s.dtj1 = a.dt[s.j1]; s.dtj2 = a.dt[s.j2];
}

// MODE 0: ADVECTION_EDGE_GCN summed over the column into ADX2D/ADY2D (as
// k_uvedge<0> + k_adv3d_gather); MODE 1: ADV_UV_EDGE_GCN with UF/VF and
// VISCOFM (as k_uvedge<1> + k_advuv_gather).
template <int MODE, typename FT>
__global__ __launch_bounds__(256) void k_uv_elem(DevGrid g, IntArgs a) {
  const int i = (int)xcd_thread();
  if (i > g.N) return;
  if (i == 0) {
    if (MODE == 1)
      for (int k = 1; k <= g.KB; ++k) a.viscofm[EI(g, 0, k)] = 0.0;
    return;
  }
  const FT half = (FT)0.5, one = (FT)1.0;
  const double* U = a.u;
  const double* V = a.v;
  // own element
  const int n1 = NBE_(i, 1), n2 = NBE_(i, 2), n3 = NBE_(i, 3);
  FT sa1[4], sa2[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) { sa1[j] = (FT)A1U_(i, j + 1); sa2[j] = (FT)A2U_(i, j + 1); }
  const FT arts = (FT)g.art[i], ccs = (FT)g.cc_hvc[i];
  EdgeSide<FT> es[3];
  int ent[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    ent[q] = g.ed_ent[3 * i + q];
    edge_side_load<FT>(g, a, i, ent[q], es[q]);
  }
  const FT fact = (FT)g.fact, hprnu = (FT)g.hprnu, fm1 = (FT)g.fm1;
  const double ep = g.epor[i];
  double tp[3], elij[3], dlxd[3], dlyd[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int e = ent[q] >> 1;
    tp[q] = (double)(1 - g.isbc[e]) * ep;
    dlxd[q] = g.dltxc[e];
    dlyd[q] = g.dltyc[e];
    elij[q] = MODE == 1 ? 0.5 * (a.egf[es[q].j1] + a.egf[es[q].j2]) : 0.0;
  }
  const bool obc = g.isbce[i] == 2;
  const int r0 = g.re_off[i], r1 = g.re_off[i + 1];
  const double art = g.art[i], dt1 = a.dt1[i], d1 = a.d1[i], cor = g.cor[i], ge = g.grav_e[i];
  double sx = 0.0, sy = 0.0;
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.n1;
    // every velocity of the layer first (own, own neighbours, each other
    // element and its neighbours), then the arithmetic
    const FT us = (FT)U[ko + i], vs = (FT)V[ko + i];
    const FT u1 = (FT)U[ko + n1], u2 = (FT)U[ko + n2], u3 = (FT)U[ko + n3];
    const FT v1 = (FT)V[ko + n1], v2 = (FT)V[ko + n2], v3 = (FT)V[ko + n3];
    FT uo[3], vo[3], uo1[3], uo2[3], uo3[3], vo1[3], vo2[3], vo3[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      uo[q] = (FT)U[ko + es[q].o]; vo[q] = (FT)V[ko + es[q].o];
      uo1[q] = (FT)U[ko + es[q].on1]; uo2[q] = (FT)U[ko + es[q].on2]; uo3[q] = (FT)U[ko + es[q].on3];
      vo1[q] = (FT)V[ko + es[q].on1]; vo2[q] = (FT)V[ko + es[q].on2]; vo3[q] = (FT)V[ko + es[q].on3];
    }
    const FT sc1 = sa1[0] * us + sa1[1] * u1 + sa1[2] * u2 + sa1[3] * u3;
    const FT sc2 = sa2[0] * us + sa2[1] * u1 + sa2[2] * u2 + sa2[3] * u3;
    const FT sc5 = sa1[0] * vs + sa1[1] * v1 + sa1[2] * v2 + sa1[3] * v3;
    const FT sc6 = sa2[0] * vs + sa2[1] * v1 + sa2[2] * v2 + sa2[3] * v3;
    const FT svisc = arts * sqrt(sq(sc1) + sq(sc6) + half * sq(sc2 + sc5));
    double xf = 0.0, yf = 0.0, px = 0.0, py = 0.0, vm = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const EdgeSide<FT>& s = es[q];
      const FT oc1 = s.oa1[0] * uo[q] + s.oa1[1] * uo1[q] + s.oa1[2] * uo2[q] + s.oa1[3] * uo3[q];
      const FT oc2 = s.oa2[0] * uo[q] + s.oa2[1] * uo1[q] + s.oa2[2] * uo2[q] + s.oa2[3] * uo3[q];
      const FT oc5 = s.oa1[0] * vo[q] + s.oa1[1] * vo1[q] + s.oa1[2] * vo2[q] + s.oa1[3] * vo3[q];
      const FT oc6 = s.oa2[0] * vo[q] + s.oa2[1] * vo1[q] + s.oa2[2] * vo2[q] + s.oa2[3] * vo3[q];
      const FT ovisc = s.arto * sqrt(sq(oc1) + sq(oc6) + half * sq(oc2 + oc5));
      // the edge kernel's operands by role
      const bool b = s.own_ib;
      const FT uia = b ? uo[q] : us, vIA = b ? vo[q] : vs, uib = b ? us : uo[q], vIB = b ? vs : vo[q];
      const FT cofa1 = b ? oc1 : sc1, cofa2 = b ? oc2 : sc2, cofa5 = b ? oc5 : sc5, cofa6 = b ? oc6 : sc6;
      const FT cofa3 = b ? sc1 : oc1, cofa4 = b ? sc2 : oc2, cofa7 = b ? sc5 : oc5, cofa8 = b ? sc6 : oc6;
      const FT viscof1 = b ? ovisc : svisc, viscof2 = b ? svisc : ovisc;
      const FT cca = b ? s.cco : ccs, ccb = b ? ccs : s.cco;
      const FT dij = (FT)(0.5 * (s.dtj1 * DZ(s.j1, k) + s.dtj2 * DZ(s.j2, k)));
      const FT uij1 = uia + cofa1 * s.xija + cofa2 * s.yija;
      const FT vij1 = vIA + cofa5 * s.xija + cofa6 * s.yija;
      const FT uij2 = uib + cofa3 * s.xijb + cofa4 * s.yijb;
      const FT vij2 = vIB + cofa7 * s.xijb + cofa8 * s.yijb;
      const FT uij = half * (uij1 + uij2), vij = half * (vij1 + vij2);
      const FT visc0 = fm1 * half * (cca + ccb);
      const FT viscof = (fact * half * (viscof1 * cca + viscof2 * ccb) + visc0) / hprnu;
      const FT txxij = (cofa1 + cofa3) * viscof, tyyij = (cofa6 + cofa8) * viscof;
      const FT txyij = half * (cofa2 + cofa4 + cofa5 + cofa7) * viscof;
      const FT fxx = dij * (txxij * s.dly - txyij * s.dlx);
      const FT fyy = dij * (txyij * s.dly - tyyij * s.dlx);
      FT xadv_, yadv_;
      if (MODE == 0) {
        const FT un_tmp = vij * s.dlx - uij * s.dly;
        const FT sgn = fsign1(un_tmp);
        xadv_ = dij * un_tmp * ((one - sgn) * uij2 + (one + sgn) * uij1) * half;
        yadv_ = dij * un_tmp * ((one - sgn) * vij2 + (one + sgn) * vij1) * half;
      } else {
        const FT exflux = dij * (-uij * s.dly + vij * s.dlx);
        const FT sgn = fsign1(exflux);
        xadv_ = exflux * ((one - sgn) * uij2 + (one + sgn) * uij1) * half;
        yadv_ = exflux * ((one - sgn) * vij2 + (one + sgn) * vij1) * half;
      }
      const FT rfx = fxx + (FT)3.0 * fxx * s.isbc;
      const FT rfy = fyy + (FT)3.0 * fyy * s.isbc;
      // the element's sum (k_adv3d_gather / k_advuv_gather)
      const double xadv = (double)xadv_, yadv = (double)yadv_, fx = (double)rfx, fy = (double)rfy;
      const double ax = xadv * tp[q], ay = yadv * tp[q], bx = fx * ep, by = fy * ep;
      xf = b ? (xf - ax) - bx : (xf + ax) + bx;
      yf = b ? (yf - ay) - by : (yf + ay) + by;
      if (MODE == 1) {
        vm = vm + (double)viscof;
        const double gd = ge * dt1 * DZ1(i, k) * elij[q];
        const double gy = gd * dlyd[q], gx = gd * dlxd[q];
        px = b ? px + gy : px - gy;
        py = b ? py - gx : py + gx;
      }
    }
    if (MODE == 0) {
      if (obc) { xf = 0.0; yf = 0.0; }
      for (int p = r0; p < r1; ++p) {
        const int r = g.re_riv[p];
        const int j = g.inodeq[r];
        const double vlctyq = a.qdis[r] / g.qarea[r];
        const double vq = g.vqdist[(long)(k - 1) * g.NUMQBC + r];
        const double temp = 0.5 * a.qdis[r] * vq * vq * vlctyq / DZ(j, k);
        xf = xf - temp * g.cos_angleq[r];
        yf = yf - temp * g.sin_angleq[r];
      }
      sx = sx + xf;
      sy = sy + yf;
    } else {
      const double dzk = DZ1(i, k);
      xf = xf + px;
      yf = yf + py;
      // vertical advection (adv_uv_edge_gcn.F:927-946)
      const double uk = U[ko + i], vk = V[ko + i];
      const long kp = ko + g.n1, km = k != 1 ? ko - g.n1 : ko;
      const double dzp = DZ1(i, k + 1), dzm = DZ1(i, k != 1 ? k - 1 : k);
      const double wk = a.w[ko + i], wp = a.w[kp + i];
      const double up = U[kp + i], vp = V[kp + i], um = U[km + i], vmm = V[km + i];
      double xfluxv, yfluxv;
      if (k == 1) {
        xfluxv = -wp * (uk * dzp + up * dzk) / (dzk + dzp);
        yfluxv = -wp * (vk * dzp + vp * dzk) / (dzk + dzp);
      } else if (k == g.KBM1) {
        xfluxv = wk * (uk * dzm + um * dzk) / (dzk + dzm);
        yfluxv = wk * (vk * dzm + vmm * dzk) / (dzk + dzm);
      } else {
        xfluxv = wk * (uk * dzm + um * dzk) / (dzk + dzm) - wp * (uk * dzp + up * dzk) / (dzk + dzp);
        yfluxv = wk * (vk * dzm + vmm * dzk) / (dzk + dzm) - wp * (vk * dzp + vp * dzk) / (dzk + dzp);
      }
      xf = xf + xfluxv * art + a.drhox[ko + i] - cor * vk * dt1 * dzk * art;
      yf = yf + yfluxv * art + a.drhoy[ko + i] + cor * uk * dt1 * dzk * art;
      if (obc) { xf = 0.0; yf = 0.0; }
      for (int p = r0; p < r1; ++p) {
        const int r = g.re_riv[p];
        const int j = g.inodeq[r];
        const double vlctyq = a.qdis[r] / g.qarea[r];
        const double vq = g.vqdist[(long)(k - 1) * g.NUMQBC + r];
        const double temp = 0.5 * a.qdis[r] * vq * vq * vlctyq / DZ(j, k);
        xf = xf - temp * g.cos_angleq[r];
        yf = yf - temp * g.sin_angleq[r];
      }
      a.uf[ko + i] = uk * dt1 / d1 - g.dti * xf / art / (d1 * dzk);
      a.vf[ko + i] = vk * dt1 / d1 - g.dti * yf / art / (d1 * dzk);
      a.viscofm[ko + i] = vm / art;
    }
  }
  if (MODE == 0) {
    a.adx2d[i] = sx;
    a.ady2d[i] = sy;
  } else {
    a.viscofm[EI(g, i, g.KB)] = 0.0 / art;
  }
}
