// fvgpu_kernels_tracer.h -- node-centred control-volume kernels: the MY2.5
// turbulence quantities (ADV_Q, FCT_Q2/Q2L, VDIF_Q) and the tracers
// (ADV_T/ADV_S, FCT_T/FCT_S, VDIF_TS, BCOND_TS), plus DENS2.
//
// The upwind face value of a control-volume sub-edge needs, at both of its
// nodes, the ring gradient (adv_t.F:219-376) and the ring min/max
// (adv_t.F:469-480).  k_ring_prep evaluates those once per node-layer into
// scratch arrays; the flux kernels then gather over each node's sub-edges in
// sub-edge order, every flux being evaluated identically from both sides.
#pragma once
#include "fvgpu_kernels_3d.h"

// Ring gradient and bounds of field f at one node-layer.  mean: per-node
// mean field (TMEAN1/SMEAN1) or nullptr with the constant cmean (QMEAN1).
__global__ void k_ring_prep(DevGrid g, IntArgs a, const double* f, const double* mean, double cmean,
                            int k0, int k1) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nk = k1 - k0 + 1;
  if (t >= (long)g.M * nk) return;
  const int i = (int)(t % g.M) + 1, k = (int)(t / g.M) + k0;
  const long ko = (long)(k - 1) * g.m1;
  const int o = g.nbsn_off[i], nt = g.ntsn[i];
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
  double lo = f[ko + g.nbsn[o]], hi = lo;
  for (int j = 0; j < nt - 1; ++j) {
    const int i1 = g.nbsn[o + j], i2 = g.nbsn[o + j + 1];
    const double f1 = f[ko + i1], f2 = f[ko + i2];
    const double ffd = mean ? 0.5 * (f1 + f2 - mean[ko + i1] - mean[ko + i2]) : 0.5 * (f1 + f2 - cmean - cmean);
    const double ff1 = 0.5 * (f1 + f2);
    s1 = s1 + ff1 * g.dltytrie[o + j];
    s2 = s2 + ff1 * g.dltxtrie[o + j];
    s3 = s3 + ffd * g.dltytrie[o + j];
    s4 = s4 + ffd * g.dltxtrie[o + j];
    if (j >= 1) { lo = fmin(lo, f1); hi = fmax(hi, f1); }
  }
  const double c = f[ko + i];
  const double art2 = g.art2[i];
  const long q = NI(g, i, k);
  a.gx[q] = s1 / art2;
  a.gy[q] = s2 / art2;
  a.gxd[q] = s3 / art2;
  a.gyd[q] = s4 / art2;
  a.fmin_[q] = fmin(lo, c);
  a.fmax_[q] = fmax(hi, c);
}

// ADV_Q(Q,QF) (adv_q.F:35-418) for one node column, layers 2..KBM1.
__global__ void k_adv_q(DevGrid g, IntArgs a, const double* q, double* qf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > g.M) return;
  if (i == 0) {
    for (int k = 1; k <= g.KB; ++k) qf[NI(g, 0, k)] = 0.0;
    return;
  }
  const int c0 = g.cv_off[i], c1 = g.cv_off[i + 1];
  const double art1 = g.art1[i];
  qf[NI(g, i, 1)] = 0.0;
  qf[NI(g, i, g.KB)] = 0.0;
  for (int k = 2; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.m1;
    const long eo = (long)(k - 1) * g.n1, eom = eo - g.n1;
    double xflux = 0.0;
    for (int p = c0; p < c1; ++p) {
      const int ent = g.cv_ent[p], c = ent >> 1;
      const int ia = NIEC_(c, 1), ib = NIEC_(c, 2), i1 = g.ntrg[c];
      const double dz1k = DZ1(i1, k), dz1m = DZ1(i1, k - 1);
      const double uq = (a.u[eo + i1] * dz1m + a.u[eom + i1] * dz1k) / (dz1k + dz1m);
      const double vq = (a.v[eo + i1] * dz1m + a.v[eom + i1] * dz1k) / (dz1k + dz1m);
      const double dtij = a.dt1[i1] * DZZ1(i1, k - 1);
      const double un = vq * g.dltxe[c] - uq * g.dltye[c];
      double fij1 = q[ko + ia] + DLTXNCVE_(c, 1) * a.gx[ko + ia] + DLTYNCVE_(c, 1) * a.gy[ko + ia];
      double fij2 = q[ko + ib] + DLTXNCVE_(c, 2) * a.gx[ko + ib] + DLTYNCVE_(c, 2) * a.gy[ko + ib];
      const double q1min = a.fmin_[ko + ia], q1max = a.fmax_[ko + ia];
      const double q2min = a.fmin_[ko + ib], q2max = a.fmax_[ko + ib];
      if (fij1 < q1min) fij1 = q1min;
      if (fij1 > q1max) fij1 = q1max;
      if (fij2 < q2min) fij2 = q2min;
      if (fij2 > q2max) fij2 = q2max;
      const double vfa = (a.viscofh[ko + ia] * DZ(ia, k - 1) + a.viscofh[ko - g.m1 + ia] * DZ(ia, k)) / (DZ(ia, k) + DZ(ia, k - 1));
      const double vfb = (a.viscofh[ko + ib] * DZ(ib, k - 1) + a.viscofh[ko - g.m1 + ib] * DZ(ib, k)) / (DZ(ib, k) + DZ(ib, k - 1));
      const double viscof = (g.fact * 0.5 * (vfa * g.nn_hvc[ia] + vfb * g.nn_hvc[ib]) + g.fm1 * 0.5 * (g.nn_hvc[ia] + g.nn_hvc[ib])) / g.hprnu;
      const double txx = 0.5 * (a.gxd[ko + ia] + a.gxd[ko + ib]) * viscof;
      const double tyy = 0.5 * (a.gyd[ko + ia] + a.gyd[ko + ib]) * viscof;
      const double fxx = -dtij * txx * g.dltye[c];
      const double fyy = dtij * tyy * g.dltxe[c];
      const double sgn = fsign1(un);
      const double exflux = -un * dtij * ((1.0 + sgn) * fij2 + (1.0 - sgn) * fij1) * 0.5 + fxx + fyy;
      if (ent & 1) xflux = xflux - exflux;
      else xflux = xflux + exflux;
    }
    const double temp = a.wts[ko - g.m1 + i] * q[ko - g.m1 + i] - a.wts[ko + g.m1 + i] * q[ko + g.m1 + i];
    xflux = xflux + temp * art1 * DZZ(i, k - 1) / (DZ(i, k - 1) + DZ(i, k));
    qf[ko + i] = (q[ko + i] - xflux / art1 * (g.dti / (a.dt[i] * DZZ(i, k - 1)))) * (a.dt[i] / a.d[i]);
  }
}

// FCT_Q2 / FCT_Q2L (fct_q2.F:31-104).
__global__ void k_fct_q(DevGrid g, const double* q, double* qf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > g.M || g.fct_skip[i]) return;
  const int o = g.nbsn_off[i], nt = g.ntsn[i];
  for (int k = 2; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.m1;
    double qmax = q[ko + g.nbsn[o]], qmin = qmax;
    for (int j = 1; j < nt; ++j) { const double x = q[ko + g.nbsn[o + j]]; qmax = fmax(qmax, x); qmin = fmin(qmin, x); }
    const double qk = q[ko + i];
    if (k == 2) {
      const double v = (qk * DZZ(i, k) + q[ko + g.m1 + i] * DZZ(i, k - 1)) / (DZZ(i, k) + DZZ(i, k - 1));
      qmax = fmax(qmax, v); qmin = fmin(qmin, v);
    } else if (k == g.KBM1) {
      const double v = (qk * DZZ(i, k - 2) + q[ko - g.m1 + i] * DZZ(i, k - 1)) / (DZZ(i, k - 1) + DZZ(i, k - 2));
      qmax = fmax(qmax, v); qmin = fmin(qmin, v);
    } else {
      const double v1 = (qk * DZZ(i, k - 2) + q[ko - g.m1 + i] * DZZ(i, k - 1)) / (DZZ(i, k - 1) + DZZ(i, k - 2));
      const double v2 = (qk * DZZ(i, k) + q[ko + g.m1 + i] * DZZ(i, k - 1)) / (DZZ(i, k) + DZZ(i, k - 1));
      qmax = fmax(qmax, fmax(v1, v2)); qmin = fmin(qmin, fmin(v1, v2));
    }
    double x = qf[ko + i];
    if (qmin - x > 0.0) x = qmin;
    if (x - qmax > 0.0) x = qmax;
    qf[ko + i] = x;
  }
}

// ADV_T / ADV_S (adv_t.F:34-1077, adv_s.F:34-1029) for one node column.
// which = 1 temperature, 2 salinity; both set PFPXB/PFPYB/AH_BOTTOM
// (adv_t.F:377-388, adv_s.F:71-84).
__global__ void k_adv_ts(DevGrid g, IntArgs a, const double* f1, double* ff, const double* fdis, int which) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > g.M) return;
  const int c0 = g.cv_off[i], c1 = g.cv_off[i + 1];
  const double art1 = g.art1[i];
  const int obc = g.node_obc[i];
  const bool isobc = g.isonb[i] == 2;
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.m1;
    const long eo = (long)(k - 1) * g.n1;
    double xflux = 0.0, xadv = 0.0;
    for (int p = c0; p < c1; ++p) {
      const int ent = g.cv_ent[p], c = ent >> 1;
      const int ia = NIEC_(c, 1), ib = NIEC_(c, 2), i1 = g.ntrg[c];
      const double dtij = a.dt1[i1] * DZ1(i1, k);
      const double un = a.v[eo + i1] * g.dltxe[c] - a.u[eo + i1] * g.dltye[c];
      double fij1 = f1[ko + ia] + DLTXNCVE_(c, 1) * a.gx[ko + ia] + DLTYNCVE_(c, 1) * a.gy[ko + ia];
      double fij2 = f1[ko + ib] + DLTXNCVE_(c, 2) * a.gx[ko + ib] + DLTYNCVE_(c, 2) * a.gy[ko + ib];
      const double t1min = a.fmin_[ko + ia], t1max = a.fmax_[ko + ia];
      const double t2min = a.fmin_[ko + ib], t2max = a.fmax_[ko + ib];
      if (fij1 < t1min) fij1 = t1min;
      if (fij1 > t1max) fij1 = t1max;
      if (fij2 < t2min) fij2 = t2min;
      if (fij2 > t2max) fij2 = t2max;
      const double viscof = g.fact * 0.5 * (a.viscofh[ko + ia] * g.nn_hvc[ia] + a.viscofh[ko + ib] * g.nn_hvc[ib]) +
                            g.fm1 * 0.5 * (g.nn_hvc[ia] + g.nn_hvc[ib]);
      const double txx = 0.5 * (a.gxd[ko + ia] + a.gxd[ko + ib]) * viscof;
      const double tyy = 0.5 * (a.gyd[ko + ia] + a.gyd[ko + ib]) * viscof;
      const double fxx = -dtij * txx * g.dltye[c];
      const double fyy = dtij * tyy * g.dltxe[c];
      const double sgn = fsign1(un);
      const double exflux = -un * dtij * ((1.0 + sgn) * fij2 + (1.0 - sgn) * fij1) * 0.5 + fxx + fyy;
      if (ent & 1) { xflux = xflux - exflux; xadv = xadv - (exflux - fxx - fyy); }
      else { xflux = xflux + exflux; xadv = xadv + (exflux - fxx - fyy); }
    }
    if (obc >= 0) a.xflux_obc[(long)(k - 1) * g.IOBCN + obc] = xadv;
    if (k == g.KBM1) {  // ADV_T and ADV_S both set the bottom-gradient terms VDIF_TS reads
      a.pfpxb[i] = a.gx[ko + i];
      a.pfpyb[i] = a.gy[ko + i];
      a.ah_bottom[i] = (g.fact * a.viscofh[ko + i] + g.fm1) * g.nn_hvc[i];
    }
    const double fk = f1[ko + i], dzk = DZ(i, k);
    double temp;
    if (k == 1) {
      const double dzp = DZ(i, k + 1);
      temp = -a.wts[ko + g.m1 + i] * (fk * dzp + f1[ko + g.m1 + i] * dzk) / (dzk + dzp);
    } else if (k == g.KBM1) {
      const double dzm = DZ(i, k - 1);
      temp = a.wts[ko + i] * (fk * dzm + f1[ko - g.m1 + i] * dzk) / (dzk + dzm);
    } else {
      const double dzp = DZ(i, k + 1), dzm = DZ(i, k - 1);
      temp = a.wts[ko + i] * (fk * dzm + f1[ko - g.m1 + i] * dzk) / (dzk + dzm) -
             a.wts[ko + g.m1 + i] * (fk * dzp + f1[ko + g.m1 + i] * dzk) / (dzk + dzp);
    }
    if (isobc) xflux = temp * art1;
    else xflux = xflux + temp * art1;
    if (!g.river_ts_specified) {
      const int rv = g.node_riv[i];
      if (rv >= 0)
        for (int j = rv; j < g.NUMQBC; ++j)
          if (g.inodeq[j] == i) xflux = xflux - a.qdis[j] * g.vqdist[(long)(k - 1) * g.NUMQBC + j] * fdis[j];
    }
    if (which == 1) xflux = xflux - 0.0 * art1;
    ff[ko + i] = (fk - xflux / art1 * (g.dti / (a.dt[i] * dzk))) * (a.dt[i] / a.dtfa[i]);
  }
}

// FCT_T / FCT_S (fct_t.F:33-108).
__global__ void k_fct_ts(DevGrid g, const double* f1, double* ff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > g.M || g.fct_skip[i]) return;
  const int o = g.nbsn_off[i], nt = g.ntsn[i];
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.m1;
    double tmax = f1[ko + g.nbsn[o]], tmin = tmax;
    for (int j = 1; j < nt; ++j) { const double x = f1[ko + g.nbsn[o + j]]; tmax = fmax(tmax, x); tmin = fmin(tmin, x); }
    const double fk = f1[ko + i];
    if (k == 1) {
      const double v = (fk * DZ(i, k + 1) + f1[ko + g.m1 + i] * DZ(i, k)) / (DZ(i, k) + DZ(i, k + 1));
      tmax = fmax(tmax, v); tmin = fmin(tmin, v);
    } else if (k == g.KBM1) {
      const double v = (fk * DZ(i, k - 1) + f1[ko - g.m1 + i] * DZ(i, k)) / (DZ(i, k) + DZ(i, k - 1));
      tmax = fmax(tmax, v); tmin = fmin(tmin, v);
    } else {
      const double v1 = (fk * DZ(i, k - 1) + f1[ko - g.m1 + i] * DZ(i, k)) / (DZ(i, k) + DZ(i, k - 1));
      const double v2 = (fk * DZ(i, k + 1) + f1[ko + g.m1 + i] * DZ(i, k)) / (DZ(i, k) + DZ(i, k + 1));
      tmax = fmax(tmax, fmax(v1, v2)); tmin = fmin(tmin, fmin(v1, v2));
    }
    double x = ff[ko + i];
    if (tmin - x > 0.0) x = tmin;
    if (x - tmax > 0.0) x = tmax;
    ff[ko + i] = x;
  }
}

// BCOND_TS(NCON) (bcond_ts.F:28-268), one thread per OBC node; thread
// IOBCN.. handles the specified river values first.
__global__ void k_bcond_ts(DevGrid g, IntArgs a, int ncon) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double* ff = ncon == 1 ? a.tf1 : a.sf1;
  const double* f1 = ncon == 1 ? a.t1 : a.s1;
  if (t == 0 && g.river_ts_specified) {
    for (int r = 0; r < g.NUMQBC; ++r) {
      const int j11 = g.inodeq[r];
      for (int k = 1; k <= g.KBM1; ++k) { a.tf1[NI(g, j11, k)] = g.tdis[r]; a.sf1[NI(g, j11, k)] = g.sdis[r]; }
    }
  }
  if (!g.obc_on || t >= g.IOBCN) return;
  const int j = g.i_obc_n[t], j1 = g.next_obc[t];
  const int KBM1 = g.KBM1;
  double t2d = 0.0, t2d_next = 0.0, xflux2d = 0.0;
  for (int k = 1; k <= KBM1; ++k) {
    t2d = t2d + f1[NI(g, j, k)] * DZ(j, k);
    t2d_next = t2d_next + ff[NI(g, j1, k)] * DZ(j1, k);
    xflux2d = xflux2d + a.xflux_obc[(long)(k - 1) * g.IOBCN + t];
  }
  if (a.uard_obcn[t] > 0.0) {
    const double tmp = xflux2d + t2d * a.uard_obcn[t];
    const double t2d_obc = (t2d * a.dt[j] - tmp * g.dti / g.art1[j]) / a.d[j];
    const double cc = sqrt(g.grav_n[j] * g.h[j]) * g.dti / g.dltn_obc[t], cp = cc + 1.0;
    double ttmp[FV_MAXKB + 1];
    for (int k = 1; k <= KBM1; ++k) {
      const double pert_next = ff[NI(g, j1, k)] - t2d_next, pert = f1[NI(g, j, k)] - t2d;
      if (g.type_tsobc == 1) ttmp[k] = ff[NI(g, j1, k)] - t2d_next;
      else if (g.type_tsobc == 2) ttmp[k] = (cc * pert_next + pert) / cp;
      else ttmp[k] = (cc * pert_next + pert * (1.0 - g.dti / 10800.0)) / cp;
    }
    const int o = g.nbsn_off[j], nt = g.ntsn[j];
    for (int k = 1; k <= KBM1; ++k) {
      double x = t2d_obc + ttmp[k];
      const long ko = (long)(k - 1) * g.m1;
      double tmax = f1[ko + g.nbsn[o]], tmin = tmax;
      for (int jj = 1; jj < nt; ++jj) { const double y = f1[ko + g.nbsn[o + jj]]; tmax = fmax(tmax, y); tmin = fmin(tmin, y); }
      const double fk = f1[ko + j];
      if (k == 1) {
        const double v = (fk * DZ(j, k + 1) + f1[ko + g.m1 + j] * DZ(j, k)) / (DZ(j, k) + DZ(j, k + 1));
        tmax = fmax(tmax, v); tmin = fmin(tmin, v);
      } else if (k == KBM1) {
        const double v = (fk * DZ(j, k - 1) + f1[ko - g.m1 + j] * DZ(j, k)) / (DZ(j, k) + DZ(j, k - 1));
        tmax = fmax(tmax, v); tmin = fmin(tmin, v);
      } else {
        const double v1 = (fk * DZ(j, k - 1) + f1[ko - g.m1 + j] * DZ(j, k)) / (DZ(j, k) + DZ(j, k - 1));
        const double v2 = (fk * DZ(j, k + 1) + f1[ko + g.m1 + j] * DZ(j, k)) / (DZ(j, k) + DZ(j, k + 1));
        tmax = fmax(tmax, fmax(v1, v2)); tmin = fmin(tmin, fmin(v1, v2));
      }
      if (tmin - x > 0.0) x = tmin;
      if (x - tmax > 0.0) x = tmax;
      ff[ko + j] = x;
    }
  } else {
    for (int k = 1; k <= KBM1; ++k) ff[NI(g, j, k)] = f1[NI(g, j, k)];
  }
}

// DENS2 (eqs_of_state.F:254-403): RHO1 at every node layer, then zero the
// bottom level and node 0.
__global__ void k_dens2(DevGrid g, IntArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.m1 * g.KB) return;
  const int i = (int)(t % g.m1), k = (int)(t / g.m1) + 1;
  if (k == g.KB || i == 0) { a.rho1[t] = 0.0; return; }
  const double s = a.s1[t], tt = a.t1[t];
  double r = s * s * s * 6.76786136E-6 - s * s * 4.8249614E-4 + s * 8.14876577E-1 - 0.22584586E0;
  r = r * (tt * tt * tt * 1.667E-8 - tt * tt * 8.164E-7 + tt * 1.803E-5);
  r = r + 1.0 - tt * tt * tt * 1.0843E-6 + tt * tt * 9.8185E-5 - tt * 4.786E-3;
  r = r * (s * s * s * 6.76786136E-6 - s * s * 4.8249614E-4 + s * 8.14876577E-1 + 3.895414E-2);
  r = r - (tt - 3.98) * (tt - 3.98) * (tt + 283.0) / (503.57 * (tt + 67.26));
  a.rho1[t] = r * 1.e-3;
}

// End of INTERNAL_STEP (internal_step.F:1541-1590): ET=EL, DT=D, ET1=EL1, DT1=D1.
__global__ void k_end_2d(DevGrid g, IntArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= g.M) { a.et[i] = a.el[i]; a.dt[i] = a.d[i]; }
  if (i <= g.N) { a.et1[i] = a.el1[i]; a.dt1[i] = a.d1[i]; }
}
