// fvgpu_kernels_3d.h -- the baroclinic (internal) mode on the device.
//
// Element kernels own one element column each (layers in a loop, so the
// column sums ADX2D, DRX2D and the prefix sums of BAROPG are formed in the
// reference's k order); node kernels own one node column.  Horizontal fluxes
// are gathered through the same sorted edge / sub-edge lists as the 2-D mode.
#pragma once
#include "fvgpu_kernels_2d.h"

// Per node-layer inputs of a control-volume face flux, packed so that one
// gather of a neighbour node reads one record (64 bytes in FP64, 32 in the
// mixed form's FP32): the field, its ring gradient (adv_t.F:219-376), ring
// bounds (adv_t.F:469-480), the gradient of the anomaly, and the node's
// horizontal viscosity times NN_HVC.
template <typename T>
struct __align__(16) CvRecT {
  T f, gx, gy, fmin, fmax, gxd, gyd, vhn;
};

struct IntArgs {
  // elements, 3-D
  double *u, *v, *w, *ww, *uf, *vf, *drhox, *drhoy, *rho, *rmean, *t, *s, *km1, *viscofm;
  // nodes, 3-D
  double *t1, *s1, *tf1, *sf1, *rho1, *rmean1, *tmean1, *smean1, *q2, *q2l, *q2f, *q2lf;
  double *l, *km, *kh, *kq, *wts, *viscofh;
  // node-layer scratch: ring gradients and bounds
  double* gx;            // node-layer scratch (PROD of VDIF_Q)
  double *elm1, *elm2;   // the elevation one and two RK stages back (Orlanski OBC)
  // flux buffers, of the flux type FT (double, or float in the mixed form)
  void* ef2;             // per-edge ADVAVE terms (2-D), EdgeFlux4T<FT>
  double* elij2;         // per-edge mean elevation (2-D)
  void* rec;             // node-layer records of the field being advected, CvRecT<FT>
  double2* vhpair;       // node-layer pair scratch: VDIF_Q's elimination coefficients
  double* obcsum;        // BCOND_TS column sums, 5 per OBC node
  void* cvflux;          // per sub-edge, per layer face flux (NCV+1) x KB FT
  // 2-D
  double *adx2d, *ady2d, *drx2d, *dry2d, *cbc, *wubot, *wvbot, *tps, *uard, *vard, *fluxobn;
  double *el, *d, *et, *dt, *egf, *dtfa, *elf;          // nodes (elf = EL after the ext. loop)
  double *ua, *va, *el1, *d1, *et1, *dt1;               // elements
  double *pfpxb, *pfpyb, *ah_bottom;
  double *qdis, *xflux_obc, *uard_obcn;
  DevClock* clk;
};

#define GRID1(n) (((n) + 255) / 256), 256

// Start of INTERNAL_STEP (internal_step.F:115-120) and BCOND_GCN(8,0) rivers.
__global__ void k_int_clock(DevGrid g, DevClock* c, double* qdis) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  c->iint += 1;
  if (c->iramp != 0) c->ramp_int = fv_tanh((double)c->iint / (double)c->iramp);
  else c->ramp_int = 1.0;
  for (int j = 0; j < g.NUMQBC; ++j) qdis[j] = g.qdis0[j] * c->ramp_int;
}

__global__ void k_zero(double* p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// VERTICAL_MIXING_TYPE 'constant' (internal_step.F:1042-1045): KM = UMOL
// and KH = UMOL*VPRNU (formed on the host) at every node and level.
__global__ void k_const_mixing(double* km, double* kh, double umol, double kh_value, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    km[i] = umol;
    kh[i] = kh_value;
  }
}

__global__ void k_zero_int(int* p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_copy(double* dst, const double* src, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ADJUST2D3D(1) (adjust2d3d.F:52-63), BCOND_GCN(5,0) (bcond_gcn.F:537-616)
// and BOTTOM_ROUGHNESS (brough.F:88-151) for one element column.
template <int KMAX>
__global__ __launch_bounds__(256) void k_adjust1_brough(DevGrid g, IntArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > g.N) return;
  const int KBM1 = g.KBM1;
  // the column is read once into registers (KMAX >= KBM1) and written once
  double uk[KMAX], vk[KMAX];
  static_for<KMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value + 1;
    if (k <= KBM1) { uk[k - 1] = a.u[EI(g, i, k)]; vk[k - 1] = a.v[EI(g, i, k)]; }
  });
  double uac = 0.0, vac = 0.0;
  static_for<KMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value + 1;
    if (k <= KBM1) {
      uac = uac + uk[k - 1] * DZ1(i, k);
      vac = vac + vk[k - 1] * DZ1(i, k);
    }
  });
  const double du = a.ua[i] - uac, dv = a.va[i] - vac;
  const WallOp wop = wall_prep(g, i);
  double ub = 0.0, vb = 0.0;
  static_for<KMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value + 1;
    if (k <= KBM1) {
      double uu = uk[k - 1] + 1.0 * du;
      double vv = vk[k - 1] + 1.0 * dv;
      wall_apply(wop, uu, vv);
      a.u[EI(g, i, k)] = uu;
      a.v[EI(g, i, k)] = vv;
      ub = uu; vb = vv;
    }
  });
  const double vk2 = .160;
  const double zb = ZZ1(i, g.KBM1) - Z1(i, g.KB);
  double cbc;
  if (a.dt1[i] > 3.0) cbc = vk2 / sq(fv_log(zb * a.dt1[i] / g.cc_z0b[i]));
  else cbc = vk2 / sq(fv_log(zb * 3.0 / g.cc_z0b[i]));
  if (cbc < g.cbcmin) cbc = g.cbcmin;
  a.cbc[i] = cbc;
  const double btps = cbc * sqrt(ub * ub + vb * vb);
  a.wubot[i] = -btps * ub;
  a.wvbot[i] = -btps * vb;
}

// BAROPG (baropg.F:33-166) with the RMEAN round trip of RHO applied to the
// element's own column; the node round trip is k_rho1_roundtrip.
__global__ void k_baropg(DevGrid g, IntArgs a, int barotropic) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > g.N) return;
  if (i == 0) {
    for (int k = 1; k <= g.KB; ++k) { a.rmean[EI(g, 0, k)] = 0.0; a.rho[EI(g, 0, k)] = 0.0; }
    return;
  }
  if (barotropic) { a.drx2d[i] = 0.0; a.dry2d[i] = 0.0; return; }
  int nn[3][2];
  double dyv[3], dxv[3], dij[3];
  for (int j = 1; j <= 3; ++j) {
    const int jj1 = j + 1 - ((j + 1) / 4) * 3, jj2 = j + 2 - ((j + 2) / 4) * 3;
    nn[j - 1][0] = NV_(i, jj1); nn[j - 1][1] = NV_(i, jj2);
    dyv[j - 1] = g.vy[nn[j - 1][0]] - g.vy[nn[j - 1][1]];
    dxv[j - 1] = g.vx[nn[j - 1][1]] - g.vx[nn[j - 1][0]];
    dij[j - 1] = 0.5 * (a.dt[nn[j - 1][0]] + a.dt[nn[j - 1][1]]);
  }
  const double dt1 = a.dt1[i], ge = g.grav_e[i], ramp = a.clk->ramp_int;
  const int nv[3] = {NV_(i, 1), NV_(i, 2), NV_(i, 3)};
  // edge j joins nodes NV(J1) and NV(J2) with (J1, J2) = (2,3), (3,1), (1,2)
  constexpr int e0[3] = {1, 2, 0}, e1[3] = {2, 0, 1};
  double rprev[3] = {0, 0, 0}, d1c[3] = {0, 0, 0}, d2c = 0.0, rhoprev = 0.0, zzprev = 0.0;
  double sx = 0.0, sy = 0.0;
  for (int k = 1; k <= g.KBM1; ++k) {
    const double zzk = ZZ1(i, k);
    const long q = EI(g, i, k);
    const double rho_k = a.rho[q], rmean_k = a.rmean[q];
    const double rhok = rho_k - rmean_k;
    a.rho[q] = rhok + rmean_k;   // the RMEAN round trip of RHO (baropg.F:57, 161)
    // RHO1 - RMEAN1 at the element's three nodes, each loaded once
    double rn3[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) rn3[m] = a.rho1[NI(g, nv[m], k)] - a.rmean1[NI(g, nv[m], k)];
    double rij[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double r0 = rn3[e0[j]];
      const double r1 = rn3[e1[j]];
      rij[j] = 0.5 * (r0 + r1);
      if (k == 1) d1c[j] = rij[j] * (-zzk);
      else {
        const double t = 0.5 * (rprev[j] + rij[j]) * (zzprev - zzk);
        d1c[j] = t + d1c[j];
      }
    }
    if (k == 1) d2c = 0.0;
    else {
      const double t = 0.5 * (zzprev + zzk) * (rhok - rhoprev);
      d2c = d2c + t;
    }
    double dx = 0.0, dy = 0.0;
    for (int j = 0; j < 3; ++j) {
      double drho1 = dyv[j] * d1c[j] * dt1;
      double drho2 = dyv[j] * dij[j] * d2c;
      dx = dx + drho1 + drho2;
      drho1 = dxv[j] * d1c[j] * dt1;
      drho2 = dxv[j] * dij[j] * d2c;
      dy = dy + drho1 + drho2;
    }
    dx = dx * dt1 * DZ1(i, k) * ge * ramp;
    dy = dy * dt1 * DZ1(i, k) * ge * ramp;
    a.drhox[EI(g, i, k)] = dx;
    a.drhoy[EI(g, i, k)] = dy;
    sx = sx + dx;
    sy = sy + dy;
    for (int j = 0; j < 3; ++j) rprev[j] = rij[j];
    rhoprev = rhok;
    zzprev = zzk;
  }
  a.drx2d[i] = sx;
  a.dry2d[i] = sy;
  const long qb = EI(g, i, g.KB);
  a.rho[qb] = (a.rho[qb] - a.rmean[qb]) + a.rmean[qb];
}

// The RMEAN1 round trip of RHO1 (baropg.F:57, 161) at every node and layer
// 1..KBM1, one thread per node-layer.
__global__ void k_rho1_roundtrip(DevGrid g, IntArgs a) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)g.KBM1 * g.m1) return;
  a.rho1[q] = (a.rho1[q] - a.rmean1[q]) + a.rmean1[q];
}

// Ext-mode ADVAVE of INTERNAL_STEP and GX/GY (internal_step.F:387-392),
// plus the zeroing of UARD/VARD (internal_step.F:395-397).
template <typename FT>
__global__ void k_advave_int(DevGrid g, IntArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > g.N) return;
  a.uard[i] = 0.0;
  a.vard[i] = 0.0;
  if (i < 1) return;
  double xf, yf, px, py;
  advave_gather<FT>(g, i, a.ef2, a.elij2, a.ua, a.va, a.d1, a.fluxobn, a.qdis, xf, yf, px, py);
  a.adx2d[i] = a.adx2d[i] - xf;
  a.ady2d[i] = a.ady2d[i] - yf;
}

// ADJUST2D3D(2) (adjust2d3d.F:65-110).
template <int KMAX>
__global__ __launch_bounds__(256) void k_adjust2(DevGrid g, IntArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > g.N) return;
  const double uard = a.uard[i] / (double)g.isplit;
  const double vard = a.vard[i] / (double)g.isplit;
  a.uard[i] = uard;
  a.vard[i] = vard;
  if (i < 1) return;
  const int KBM1 = g.KBM1;
  // the column is read once into registers (KMAX >= KBM1) and written once
  double uk[KMAX], vk[KMAX];
  static_for<KMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value + 1;
    if (k <= KBM1) { uk[k - 1] = a.u[EI(g, i, k)]; vk[k - 1] = a.v[EI(g, i, k)]; }
  });
  double utmp = 0.0, vtmp = 0.0;
  static_for<KMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value + 1;
    if (k <= KBM1) {
      utmp = utmp + uk[k - 1] * DZ1(i, k);
      vtmp = vtmp + vk[k - 1] * DZ1(i, k);
    }
  });
  const double dt1 = a.dt1[i];
  utmp = utmp * dt1;
  vtmp = vtmp * dt1;
  static_for<KMAX>([&](auto kc) {
    constexpr int k = decltype(kc)::value + 1;
    if (k <= KBM1) {
      a.u[EI(g, i, k)] = uk[k - 1] - (utmp - uard) / dt1;
      a.v[EI(g, i, k)] = vk[k - 1] - (vtmp - vard) / dt1;
    }
  });
}

// N2E3D (mod_main.F:1633-1645) over layers 1..kmax: one thread per element
// and group of eight layers (blockIdx.y), the group's node loads issued
// together.  The grid's x extent is a multiple of 8, so every layer group
// of an element range runs on the same XCD and shares its L2.
__global__ __launch_bounds__(256) void k_n2e3d(DevGrid g, const double* nvar, double* evar, int kmax) {
  const int i = (int)xcd_thread() + 1;
  if (i > g.N) return;
  const int j1 = NV_(i, 1), j2 = NV_(i, 2), j3 = NV_(i, 3);
  const double one_third = 1.0 / 3.0;
  const int k0 = (int)blockIdx.y * 8 + 1;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = k0 + q;
    if (k <= kmax) evar[EI(g, i, k)] = one_third * (nvar[NI(g, j1, k)] + nvar[NI(g, j2, k)] + nvar[NI(g, j3, k)]);
  }
}

// WREAL (wreal.F:32-140): one thread per element column, its layers taken
// eight at a time with the eight layers' loads issued before their
// arithmetic.
__global__ __launch_bounds__(256) void k_wreal(DevGrid g, IntArgs a) {
  const int i = (int)xcd_thread() + 1;
  if (i > g.N) return;
  const int j1 = NV_(i, 1), j2 = NV_(i, 2), j3 = NV_(i, 3);
  const double dddx = AWX_(i, 1) * a.d[j1] + AWX_(i, 2) * a.d[j2] + AWX_(i, 3) * a.d[j3];
  const double dddy = AWY_(i, 1) * a.d[j1] + AWY_(i, 2) * a.d[j2] + AWY_(i, 3) * a.d[j3];
  const double dedx = AWX_(i, 1) * a.elf[j1] + AWX_(i, 2) * a.elf[j2] + AWX_(i, 3) * a.elf[j3];
  const double dedy = AWY_(i, 1) * a.elf[j1] + AWY_(i, 2) * a.elf[j2] + AWY_(i, 3) * a.elf[j3];
  const double etf1aa = (1.0 / 3.0) * (a.el[j1] + a.el[j2] + a.el[j3]);
  const double et1 = a.et1[i];
  const int KBM1 = g.KBM1;
  for (int k0 = 1; k0 <= KBM1; k0 += 8) {
    double w0[8], w1[8], uk[8], vk[8], zz[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = min(k0 + q, KBM1);
      const long p = EI(g, i, k);
      w0[q] = a.w[p];
      w1[q] = a.w[p + g.n1];
      uk[q] = a.u[p];
      vk[q] = a.v[p];
      zz[q] = ZZ1(i, k);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + q;
      if (k > KBM1) break;
      const double ww1 = 0.5 * (w0[q] + w1[q]) + uk[q] * (zz[q] * dddx + dedx) + vk[q] * (zz[q] * dddy + dedy);
      const double ww2 = (zz[q] + 1.) * (etf1aa - et1) / g.dti;
      a.ww[EI(g, i, k)] = ww1 + ww2;
    }
  }
}
