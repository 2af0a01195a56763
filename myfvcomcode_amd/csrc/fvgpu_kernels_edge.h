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
