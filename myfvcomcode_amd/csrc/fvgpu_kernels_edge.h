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
template <int MODE>
__global__ __launch_bounds__(256) void k_uvedge(DevGrid g, IntArgs a, const double* U, const double* V) {
  const int e = (int)xcd_thread();
  if (e < 1 || e > g.NE) return;
  const int ia = IEC_(e, 1), ib = IEC_(e, 2), j1 = IENODE_(e, 1), j2 = IENODE_(e, 2);
  const int k1 = NBE_(ia, 1), k2 = NBE_(ia, 2), k3 = NBE_(ia, 3), k4 = NBE_(ib, 1), k5 = NBE_(ib, 2), k6 = NBE_(ib, 3);
  const double xija = g.xijc[e] - g.xc[ia], yija = g.yijc[e] - g.yc[ia];
  const double xijb = g.xijc[e] - g.xc[ib], yijb = g.yijc[e] - g.yc[ib];
  const double a11 = A1U_(ia, 1), a12 = A1U_(ia, 2), a13 = A1U_(ia, 3), a14 = A1U_(ia, 4);
  const double a21 = A2U_(ia, 1), a22 = A2U_(ia, 2), a23 = A2U_(ia, 3), a24 = A2U_(ia, 4);
  const double b11 = A1U_(ib, 1), b12 = A1U_(ib, 2), b13 = A1U_(ib, 3), b14 = A1U_(ib, 4);
  const double b21 = A2U_(ib, 1), b22 = A2U_(ib, 2), b23 = A2U_(ib, 3), b24 = A2U_(ib, 4);
  const double dlx = g.dltxc[e], dly = g.dltyc[e];
  const double arta = g.art[ia], artb = g.art[ib], cca = g.cc_hvc[ia], ccb = g.cc_hvc[ib];
  const double visc0 = g.fm1 * 0.5 * (cca + ccb);
  const double dtj1 = a.dt[j1], dtj2 = a.dt[j2];
  const double isbc_tmp = (double)g.isbc[e];
  const long es = (long)g.NE + 1;
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.n1;
    const double dij = 0.5 * (dtj1 * DZ(j1, k) + dtj2 * DZ(j2, k));
    const double uia = U[ko + ia], vIA = V[ko + ia], uib = U[ko + ib], vIB = V[ko + ib];
    const double u1 = U[ko + k1], u2 = U[ko + k2], u3 = U[ko + k3], u4 = U[ko + k4], u5 = U[ko + k5], u6 = U[ko + k6];
    const double v1 = V[ko + k1], v2 = V[ko + k2], v3 = V[ko + k3], v4 = V[ko + k4], v5 = V[ko + k5], v6 = V[ko + k6];
    const double cofa1 = a11 * uia + a12 * u1 + a13 * u2 + a14 * u3;
    const double cofa2 = a21 * uia + a22 * u1 + a23 * u2 + a24 * u3;
    const double cofa5 = a11 * vIA + a12 * v1 + a13 * v2 + a14 * v3;
    const double cofa6 = a21 * vIA + a22 * v1 + a23 * v2 + a24 * v3;
    const double uij1 = uia + cofa1 * xija + cofa2 * yija;
    const double vij1 = vIA + cofa5 * xija + cofa6 * yija;
    const double cofa3 = b11 * uib + b12 * u4 + b13 * u5 + b14 * u6;
    const double cofa4 = b21 * uib + b22 * u4 + b23 * u5 + b24 * u6;
    const double cofa7 = b11 * vIB + b12 * v4 + b13 * v5 + b14 * v6;
    const double cofa8 = b21 * vIB + b22 * v4 + b23 * v5 + b24 * v6;
    const double uij2 = uib + cofa3 * xijb + cofa4 * yijb;
    const double vij2 = vIB + cofa7 * xijb + cofa8 * yijb;
    const double uij = 0.5 * (uij1 + uij2), vij = 0.5 * (vij1 + vij2);
    const double viscof1 = arta * sqrt(sq(cofa1) + sq(cofa6) + 0.5 * sq(cofa2 + cofa5));
    const double viscof2 = artb * sqrt(sq(cofa3) + sq(cofa8) + 0.5 * sq(cofa4 + cofa7));
    const double viscof = (g.fact * 0.5 * (viscof1 * cca + viscof2 * ccb) + visc0) / g.hprnu;
    const double txxij = (cofa1 + cofa3) * viscof, tyyij = (cofa6 + cofa8) * viscof;
    const double txyij = 0.5 * (cofa2 + cofa4 + cofa5 + cofa7) * viscof;
    const double fxx = dij * (txxij * dly - txyij * dlx);
    const double fyy = dij * (txyij * dly - tyyij * dlx);
    EdgeFlux4 r;
    if (MODE == 0) {
      const double un_tmp = vij * dlx - uij * dly;
      const double sgn = fsign1(un_tmp);
      r.xadv = dij * un_tmp * ((1.0 - sgn) * uij2 + (1.0 + sgn) * uij1) * 0.5;
      r.yadv = dij * un_tmp * ((1.0 - sgn) * vij2 + (1.0 + sgn) * vij1) * 0.5;
    } else {
      const double exflux = dij * (-uij * dly + vij * dlx);
      const double sgn = fsign1(exflux);
      r.xadv = exflux * ((1.0 - sgn) * uij2 + (1.0 + sgn) * uij1) * 0.5;
      r.yadv = exflux * ((1.0 - sgn) * vij2 + (1.0 + sgn) * vij1) * 0.5;
      a.efvisc[(long)(k - 1) * es + e] = viscof;
    }
    r.fx = fxx + 3.0 * fxx * isbc_tmp;
    r.fy = fyy + 3.0 * fyy * isbc_tmp;
    a.ef3[(long)(k - 1) * es + e] = r;
  }
}

// ADVECTION_EDGE_GCN summed over the column into ADX2D/ADY2D
// (internal_step.F:374-384): per layer the three edge terms in edge order,
// OBC elements zeroed, river momentum (advection_edge_gcn.F:600-734).
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
    const EdgeFlux4* fk = a.ef3 + (long)(k - 1) * es;
    double xf = 0.0, yf = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const EdgeFlux4 r = fk[ee[q] >> 1];
      if ((ee[q] & 1) == 0) {
        xf = xf + r.xadv * tp[q] + r.fx * ep;
        yf = yf + r.yadv * tp[q] + r.fy * ep;
      } else {
        xf = xf - r.xadv * tp[q] - r.fx * ep;
        yf = yf - r.yadv * tp[q] - r.fy * ep;
      }
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
    const EdgeFlux4* fk = a.ef3 + (long)(k - 1) * es;
    const double* vk_ = a.efvisc + (long)(k - 1) * es;
    const double dzk = DZ1(i, k);
    double xf = 0.0, yf = 0.0, px = 0.0, py = 0.0, vm = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int e = ee[q] >> 1;
      const EdgeFlux4 r = fk[e];
      vm = vm + vk_[e];
      const double gd = ge * dt1 * dzk * elij[q];
      if ((ee[q] & 1) == 0) {
        xf = xf + r.xadv * tp[q] + r.fx * ep;
        yf = yf + r.yadv * tp[q] + r.fy * ep;
        px = px - gd * dly[q];
        py = py + gd * dlx[q];
      } else {
        xf = xf - r.xadv * tp[q] - r.fx * ep;
        yf = yf - r.yadv * tp[q] - r.fy * ep;
        px = px + gd * dly[q];
        py = py - gd * dlx[q];
      }
    }
    xf = xf + px;
    yf = yf + py;
    // vertical advection (adv_uv_edge_gcn.F:927-946)
    const double uk = U[ko + i], vk = V[ko + i];
    double xfluxv, yfluxv;
    if (k == 1) {
      const long kp = ko + g.n1; const double dzp = DZ1(i, k + 1);
      xfluxv = -W[kp + i] * (uk * dzp + U[kp + i] * dzk) / (dzk + dzp);
      yfluxv = -W[kp + i] * (vk * dzp + V[kp + i] * dzk) / (dzk + dzp);
    } else if (k == g.KBM1) {
      const long km = ko - g.n1; const double dzm = DZ1(i, k - 1);
      xfluxv = W[ko + i] * (uk * dzm + U[km + i] * dzk) / (dzk + dzm);
      yfluxv = W[ko + i] * (vk * dzm + V[km + i] * dzk) / (dzk + dzm);
    } else {
      const long kp = ko + g.n1, km = ko - g.n1;
      const double dzp = DZ1(i, k + 1), dzm = DZ1(i, k - 1);
      xfluxv = W[ko + i] * (uk * dzm + U[km + i] * dzk) / (dzk + dzm) - W[kp + i] * (uk * dzp + U[kp + i] * dzk) / (dzk + dzp);
      yfluxv = W[ko + i] * (vk * dzm + V[km + i] * dzk) / (dzk + dzm) - W[kp + i] * (vk * dzp + V[kp + i] * dzk) / (dzk + dzp);
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
