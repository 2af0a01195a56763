// fvgpu_kernels_tracer.h -- node-centred kernels of the MY2.5 turbulence
// quantities (FCT_Q2/Q2L, VDIF_Q) and the tracers (FCT_T/FCT_S, VDIF_TS,
// BCOND_TS), plus DENS2.  Their horizontal advection is in
// fvgpu_kernels_cv.h.
#pragma once
#include "fvgpu_kernels_3d.h"

// ADJUST_TS (adjust_ts.F:150-172), run when RIVER_TS_SETTING is
// 'calculated' (internal_step.F:1374-1376), in one workgroup: T1/S1 of
// every river node raised to its river's TDIS/SDIS on layers 1..KBM1, one
// thread per layer walking the rivers in order (two rivers at one node see
// each other's result, as in the reference's loop); then N2E3D of T1 and S1
// for the elements around the river nodes, the only elements whose three
// nodes changed (the others' T/S already hold the same values).
__global__ void k_adjust_ts(DevGrid g, IntArgs a) {
  // a river's nodes: its node, or the two nodes of its solid edge
  auto river_nodes = [&](int r, int (&nd)[2]) {
    if (g.icellq[r] > 0) { nd[0] = g.n_icellq[r]; nd[1] = g.n_icellq[g.NUMQBC + r]; return 2; }
    nd[0] = g.inodeq[r];
    return 1;
  };
  for (int k = threadIdx.x + 1; k <= g.KBM1; k += blockDim.x)
    for (int r = 0; r < g.NUMQBC; ++r) {
      int nd[2];
      const int c = river_nodes(r, nd);
      for (int q = 0; q < c; ++q) {
        const long p = NI(g, nd[q], k);
        a.t1[p] = fmax(a.t1[p], g.tdis[r]);
        a.s1[p] = fmax(a.s1[p], g.sdis[r]);
      }
    }
  __syncthreads();
  const double one_third = 1.0 / 3.0;
  for (int r = 0; r < g.NUMQBC; ++r) {
    int nd[2];
    const int c = river_nodes(r, nd);
    for (int q0 = 0; q0 < c; ++q0) {
      const int j = nd[q0], off = g.nbve_off[j], nt = g.ntve[j];
      for (int q = threadIdx.x; q < nt * g.KB; q += blockDim.x) {
        const int i = g.nbve[off + q % nt], k = q / nt + 1;
        const int j1 = NV_(i, 1), j2 = NV_(i, 2), j3 = NV_(i, 3);
        a.t[EI(g, i, k)] = one_third * (a.t1[NI(g, j1, k)] + a.t1[NI(g, j2, k)] + a.t1[NI(g, j3, k)]);
        a.s[EI(g, i, k)] = one_third * (a.s1[NI(g, j1, k)] + a.s1[NI(g, j2, k)] + a.s1[NI(g, j3, k)]);
      }
    }
  }
}

// BCOND_TS(NCON) (bcond_ts.F:28-268) in two kernels.  Part 1, one thread
// per OBC node: the column sums of the node (T2D, the interior neighbour's
// T2D_NEXT, the OBC advective flux XFLUX2D) in layer order, into
// sums[3*t .. 3*t+2]; thread 0 also sets the specified river values first.
// The reference sets TF1 and SF1 at the river nodes in both calls; each call
// here sets its own tracer only, which leaves T1 and S1 unchanged (each copy
// F1 = FF follows its own call) and lets the T and S passes run
// concurrently.
__global__ void k_bcond_ts_sums(DevGrid g, IntArgs a, int ncon, double* sums) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const double* ff = ncon == 1 ? a.tf1 : a.sf1;
  const double* f1 = ncon == 1 ? a.t1 : a.s1;
  if (t == 0 && g.river_ts_specified) {
    double* fw = ncon == 1 ? a.tf1 : a.sf1;
    const double* fdis = ncon == 1 ? g.tdis : g.sdis;
    for (int r = 0; r < g.NUMQBC; ++r) {
      if (g.icellq[r] > 0) {   // both nodes of an edge river's solid edge
        const int j11 = g.n_icellq[r], j22 = g.n_icellq[g.NUMQBC + r];
        for (int k = 1; k <= g.KBM1; ++k) { fw[NI(g, j11, k)] = fdis[r]; fw[NI(g, j22, k)] = fdis[r]; }
        continue;
      }
      const int j11 = g.inodeq[r];
      for (int k = 1; k <= g.KBM1; ++k) fw[NI(g, j11, k)] = fdis[r];
    }
  }
  if (!g.obc_on || t >= g.IOBCN) return;
  const int j = g.i_obc_n[t], j1 = g.next_obc[t];
  double t2d = 0.0, t2d_next = 0.0, xflux2d = 0.0;
  // eight layers' loads issued together, then summed in layer order
  for (int k0 = 1; k0 <= g.KBM1; k0 += 8) {
    double fa[8], za[8], fb[8], zb[8], xf[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = min(k0 + q, g.KBM1);
      fa[q] = f1[NI(g, j, k)];
      za[q] = DZ(j, k);
      fb[q] = ff[NI(g, j1, k)];
      zb[q] = DZ(j1, k);
      xf[q] = a.xflux_obc[(long)(k - 1) * g.IOBCN + t];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (k0 + q > g.KBM1) break;
      t2d = t2d + fa[q] * za[q];
      t2d_next = t2d_next + fb[q] * zb[q];
      xflux2d = xflux2d + xf[q];
    }
  }
  sums[5 * t] = t2d;
  sums[5 * t + 1] = t2d_next;
  sums[5 * t + 2] = xflux2d;
  if (g.type_tsobc == 4) {
    // the Orlanski form's column sums of F1 at NEXT_OBC and NEXT_OBC2 (its
    // F1M1 is F1 here: see k_bcond_ts_layers)
    const int j2 = g.next_obc2[t];
    double s1 = 0.0, s2 = 0.0;
    for (int k = 1; k <= g.KBM1; ++k) {
      s1 = s1 + f1[NI(g, j1, k)] * DZ(j1, k);
      s2 = s2 + f1[NI(g, j2, k)] * DZ(j2, k);
    }
    sums[5 * t + 3] = s1;
    sums[5 * t + 4] = s2;
  }
}

// Part 2, one thread per OBC node and layer: the radiated value (TYPE_TSOBC
// 1-3) limited to the ring and vertical-neighbour bounds, or F1 itself when
// the node has no inflow.  NEXT_OBC is an interior node, so no thread reads
// what another writes.
__global__ void k_bcond_ts_layers(DevGrid g, IntArgs a, int ncon, const double* sums) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!g.obc_on || idx >= (long)g.IOBCN * g.KBM1) return;
  const int t = (int)(idx % g.IOBCN), k = (int)(idx / g.IOBCN) + 1;
  double* ff = ncon == 1 ? a.tf1 : a.sf1;
  const double* f1 = ncon == 1 ? a.t1 : a.s1;
  const int j = g.i_obc_n[t], j1 = g.next_obc[t];
  const int KBM1 = g.KBM1;
  if (!(a.uard_obcn[t] > 0.0)) {
    ff[NI(g, j, k)] = f1[NI(g, j, k)];
    return;
  }
  const double t2d = sums[5 * t], t2d_next = sums[5 * t + 1], xflux2d = sums[5 * t + 2];
  const double tmp = xflux2d + t2d * a.uard_obcn[t];
  const double t2d_obc = (t2d * a.dt[j] - tmp * g.dti / g.art1[j]) / a.d[j];
  const double cc = sqrt(g.grav_n[j] * g.h[j]) * g.dti / g.dltn_obc[t], cp = cc + 1.0;
  const double pert_next = ff[NI(g, j1, k)] - t2d_next, pert = f1[NI(g, j, k)] - t2d;
  double ttmp;
  if (g.type_tsobc == 1) ttmp = ff[NI(g, j1, k)] - t2d_next;
  else if (g.type_tsobc == 2) ttmp = (cc * pert_next + pert) / cp;
  else if (g.type_tsobc == 3) ttmp = (cc * pert_next + pert * (1.0 - g.dti / 10800.0)) / cp;
  else {
    // Orlanski (mod_obcs.F:1258-1285, 1336-1363).  Its F1M1/F1M2 are F1
    // one and two external RK stages back (ASSIGN_ELM1_TO_ELM2, called at
    // every stage, mod_obcs.F:190-196); F1 does not change during the
    // external steps, which run before BCOND_TS, so both equal F1 here, and
    // their column sums are the same sums of F1 (T2D at the OBC node,
    // T2D_NEXT1 at NEXT_OBC) formed in the same order.
    const int j2 = g.next_obc2[t];
    const double t2d_next1 = sums[5 * t + 3], tm12d_next2 = sums[5 * t + 4];
    const double tm12d_next1 = t2d, tm22d_next1 = t2d_next1;
    const double f1j1 = f1[NI(g, j1, k)], m2j1 = f1j1, m1j2 = f1[NI(g, j2, k)], m1j = f1[NI(g, j, k)];
    const double cl = ((m2j1 - tm22d_next1) - (f1j1 - t2d_next1)) /
                      ((f1j1 - t2d_next1) + (m2j1 - tm22d_next1) - 2.0 * (m1j2 - tm12d_next2));
    double mu;
    if (cl >= 1.0) mu = 1.0;
    else if (cl > 0.0 && cl < 1.0) mu = cl;
    else mu = 0.0;
    ttmp = ((m1j - tm12d_next1) * (1.0 - mu) + 2.0 * mu * (f1j1 - t2d_next1)) / (1.0 + mu);
  }
  double x = t2d_obc + ttmp;
  const long ko = (long)(k - 1) * g.m1;
  const int o = g.nbsn_off[j], nt = g.ntsn[j];
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
