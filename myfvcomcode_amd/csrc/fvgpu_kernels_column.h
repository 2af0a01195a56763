// fvgpu_kernels_column.h -- the implicit vertical solves with their Thomas
// coefficients in registers.
//
// One thread owns one column.  The recursion coefficients VH/VHP of the
// forward sweep are the only per-level state the back substitution needs;
// they live in register arrays of a compile-time bound KMAX >= KB (the loops
// are fully unrolled, so every index is a constant), and everything else
// (A, C, BOY, L, GH, STF, DTEF) is recomputed where it is used with the same
// operations, so the results are bit-identical to the straightforward form.
// The "last level" values are carried in scalars because KBM1/KBM2 are
// runtime values.
#pragma once
#include "fvgpu_kernels_tracer.h"

// VDIF_UV (vdif_uv.F:34-189): one thread per (element column, component);
// blockIdx.y = 0 solves U, 1 solves V.  The coefficient VHU = VHV is
// evaluated by both with the same operations.  The sponge and the wall
// projection BCOND_GCN(3,0) follow in k_uv_finish.
template <int KMAX>
__global__ __launch_bounds__(256) void k_vdif_uv_r(DevGrid g, IntArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > g.N) return;
  const int f = blockIdx.y;
  const int KBM1 = g.KBM1, KBM2 = g.KBM2;
  const double dti = g.dti, umol = g.umol, d1 = a.d1[i];
  if (!(d1 > 0.0)) return;
  const long n1 = g.n1;
  double* X = f ? a.vf : a.uf;
  const int it1 = g.isonb[NV_(i, 1)], it2 = g.isonb[NV_(i, 2)], it3 = g.isonb[NV_(i, 3)];
  const double tmp = (it1 == 2 || it2 == 2 || it3 == 2) ? 0.0 : 1.0;
  const double ramp = a.clk->ramp_int;
  const double ub = a.u[EI(g, i, KBM1)], vb = a.v[EI(g, i, KBM1)];
  const double tps = a.cbc[i] * sqrt(ub * ub + vb * vb);
  if (f == 0) {
    a.tps[i] = tps;
    a.wubot[i] = -tps * ub;
    a.wvbot[i] = -tps * vb;
  }
  const double ws = g.wind_on ? -(f ? g.wsy[i] : g.wsx[i]) * ramp * 0.001 : 0.0;
  double VH[KMAX + 1], VHP[KMAX + 1];
  const double a1 = -dti * (a.km1[EI(g, i, 2)] + umol) / (DZ1(i, 1) * DZZ1(i, 1) * d1 * d1);
  VH[1] = a1 / (a1 - 1.);
  VHP[1] = (-dti * ws * tmp / (-DZ1(i, 1) * d1) - X[i]) / (a1 - 1.);
  double lastvh = VH[1], lastvhp = VHP[1];
  static_for<KMAX - 1>([&](auto kc) {
    constexpr int k = decltype(kc)::value + 2;
    if (k <= KBM2) {
      const double ak = -dti * (a.km1[(long)k * n1 + i] + umol) / (DZ1(i, k) * DZZ1(i, k) * d1 * d1);
      const double ck = -dti * (a.km1[(long)(k - 1) * n1 + i] + umol) / (DZ1(i, k) * DZZ1(i, k - 1) * d1 * d1);
      const double t = 1.0 / (ak + ck * (1. - VH[k - 1]) - 1.);
      VH[k] = ak * t;
      VHP[k] = (ck * VHP[k - 1] - X[(long)(k - 1) * n1 + i]) * t;
      lastvh = VH[k];
      lastvhp = VHP[k];
    }
  });
  const double cb = -dti * (a.km1[(long)(KBM1 - 1) * n1 + i] + umol) / (DZ1(i, KBM1) * DZZ1(i, KBM1 - 1) * d1 * d1);
  double x = (cb * lastvhp - X[(long)(KBM1 - 1) * n1 + i]) /
             (tps * dti / (-DZ1(i, KBM1) * d1) - 1. - (lastvh - 1.) * cb);
  X[(long)(KBM1 - 1) * n1 + i] = x;
  static_for<KMAX>([&](auto kc) {
    constexpr int k = KMAX - decltype(kc)::value;
    if (k <= KBM2) {
      x = (VH[k] * x + VHP[k]);
      X[(long)(k - 1) * n1 + i] = x;
    }
  });
}

// Sponge of VDIF_UV (vdif_uv.F:155-161) and BCOND_GCN(3,0) (bcond_gcn.F:424-530)
// per element-layer.
__global__ void k_uv_finish(DevGrid g, IntArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.N * g.KBM1) return;
  const int i = (int)(t % g.N) + 1, k = (int)(t / g.N) + 1;
  const long q = EI(g, i, k);
  const double sp = g.cc_sponge[i];
  double uu = a.uf[q], vv = a.vf[q];
  uu = uu / (1.0 + sp * uu * uu);
  vv = vv / (1.0 + sp * vv * vv);
  wall_project(g, i, uu, vv);
  a.uf[q] = uu;
  a.vf[q] = vv;
}

// VDIF_TS(NCON,F) (vdif_ts.F:42-385), FP64, no surface flux.
template <int KMAX>
__global__ __launch_bounds__(256) void k_vdif_ts_r(DevGrid g, IntArgs a, double* f, int ncon) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > g.M || g.isonb[i] == 2) return;   // OBC columns keep F (vdif_ts.F:259,313)
  const int KBM1 = g.KBM1, KBM2 = g.KBM2;
  const double dti = g.dti, umolpr = g.umol * 1.E0, d = a.d[i];
  const long m1 = g.m1;
  double VHF[KMAX + 1], VHPF[KMAX + 1];
  const double af1 = -dti * (a.kh[NI(g, i, 2)] + umolpr) / (DZ(i, 1) * DZZ(i, 1) * d * d);
  VHF[1] = af1 / (af1 - 1.);
  VHPF[1] = -dti * (0.0 + 0.0 - 0.0 + 0.0 - 0.0) / (-DZ(i, 1) * d) - f[i];
  VHPF[1] = VHPF[1] / (af1 - 1.);
  double lastvh = VHF[1], lastvhp = VHPF[1];
  static_for<KMAX - 1>([&](auto kc) {
    constexpr int k = decltype(kc)::value + 2;
    if (k <= KBM2) {
      const double afk = -dti * (a.kh[(long)k * m1 + i] + umolpr) / (DZ(i, k) * DZZ(i, k) * d * d);
      const double cfk = -dti * (a.kh[(long)(k - 1) * m1 + i] + umolpr) / (DZ(i, k) * DZZ(i, k - 1) * d * d);
      const double t = 1. / (afk + cfk * (1. - VHF[k - 1]) - 1.);
      VHF[k] = afk * t;
      VHPF[k] = (cfk * VHPF[k - 1] - f[(long)(k - 1) * m1 + i] + dti * (0.0 - 0.0) / (d * DZ(i, k))) * t;
      lastvh = VHF[k];
      lastvhp = VHPF[k];
    }
  });
  const double khbottom = a.kh[NI(g, i, KBM1)];
  const double cfb = -dti * (khbottom + umolpr) / (DZ(i, KBM1) * DZZ(i, KBM1 - 1) * d * d);
  const double tmp1 = a.pfpxb[i] * g.cos_sita[i] + a.pfpyb[i] * g.sin_sita[i];
  const double tmp2 = a.ah_bottom[i] * g.phpn[i];
  const double tmp3 = khbottom + umolpr + a.ah_bottom[i] * g.phpn[i] * g.phpn[i];
  double tmp = tmp1 * tmp2 / tmp3 * (khbottom + umolpr);
  if (g.noflux_bot) {
    if (ncon == 1) { if (tmp1 > 0.0) tmp = 0.0; }
    else { if (tmp1 < 0.0) tmp = 0.0; tmp = -tmp; }
  } else {
    tmp = 0.0;
  }
  double x = (cfb * lastvhp - f[NI(g, i, KBM1)] - 0.0 + dti * (0.0 - 0.0 - tmp) / (d * DZ(i, KBM1))) /
             (cfb * (1. - lastvh) - 1.);
  f[NI(g, i, KBM1)] = x;
  static_for<KMAX>([&](auto kc) {
    constexpr int k = KMAX - decltype(kc)::value;
    if (k <= KBM2) {
      x = (VHF[k] * x + VHPF[k]);
      f[(long)(k - 1) * m1 + i] = x;
    }
  });
}

// Per-level quantities of VDIF_Q (vdif_q.F:255-368): from Q2, Q2L and RHO1
// at levels k-1 and k of node i (the values passed in).
struct Q25Level { double q2, boy, ll, gh, stf, dtef; };

__device__ __forceinline__ Q25Level q25_eval(const DevGrid& g, int i, int k, double d, double l0, double q2raw,
                                             double q2l, double rho_km1, double rho_k) {
  const double SMALL = 1.E-8, KAPPA = 0.40, GH_MAX = .0233, GHC = -6.0, GH_MIN = -.281, B1 = 16.6;
  const double p9 = (double)0.9f, p1 = (double)0.1f;
  Q25Level r;
  r.q2 = fabs(q2raw + SMALL);
  r.boy = 0.0;
  if (k >= 2 && k <= g.KBM1) r.boy = g.grav_n[i] * (rho_km1 - rho_k) / (DZZ(i, k - 1) * d);
  if (k == 1) r.ll = KAPPA * l0;
  else if (k == g.KB) r.ll = 0.0;
  else {
    double ll = fabs(q2l) / r.q2;
    if (ZN(i, k) > -0.5) ll = fmax(ll, KAPPA * l0);
    if (r.boy < 0.0) {
      const double lmax = sqrt(fabs(GH_MIN) * r.q2 / (-r.boy + SMALL));
      ll = fmin(ll, lmax);
    }
    r.ll = ll;
  }
  r.gh = 0.0;
  if (k >= 2 && k <= g.KBM1) r.gh = fmin((sq(r.ll) / r.q2) * r.boy, GH_MAX);
  r.stf = 1.0;
  if (r.gh < 0.0) r.stf = 1.0 - p9 * fv_pow15(r.gh / GHC);
  if (r.gh < GHC) r.stf = p1;
  r.dtef = sqrt(r.q2) / (B1 * r.ll + SMALL) * r.stf;
  return r;
}

// VDIF_Q, MY2.5 (vdif_q.F:34-551), one thread per node column in two
// sweeps.  The downward sweep k = 1..KB reads every node-layer value once
// (sliding windows over RHO1 and KQ, the ring averages of U/V): it forms the
// level quantities (vdif_q.F:255-368) and PROD (vdif_q.F:259-302), runs the
// forward elimination of both the Q2 solve (vdif_q.F:372-378) and the Q2L
// solve (vdif_q.F:398-417), and updates KQ/KM/KH and L, which depend on the
// old Q2/Q2L only (vdif_q.F:466-487; L(1)=KAPPA*L0, L(KB)=0).  The
// elimination coefficients VH of both solves go to registers (VHMODE 1,
// loops unrolled) or to the node-layer pair scratch `vhs` (VHMODE 0 with
// unrolled loops, 2 with rolled ones); VHP(k) is parked in Q2F(k)/Q2LF(k),
// which the elimination has just consumed.  The upward
// sweep does both back-substitutions (vdif_q.F:380-391, 420-431) and the
// threshold (vdif_q.F:440-448), and stores each final value both to
// Q2F/Q2LF and to Q2/Q2L, the copy that follows VDIF_Q in INTERNAL_STEP
// (internal_step.F:1034-1035); column 0 is copied by its thread.  Every value is formed by the reference's
// operations in its order; only the loops are fused.
template <int KMAX, int RMAX, int VHMODE>
__global__ __launch_bounds__(256) void k_vdif_q_r(DevGrid g, IntArgs a, double2* vhs, const int* only) {
  const int i = (int)xcd_thread();
  if (i > g.M) return;
  if (only && !only[i]) return;   // column subset (the learned closure's fallback)
  const int KB = g.KB, KBM1 = g.KBM1;
  const double A1 = 0.92, B1 = 16.6, A2 = 0.74, B2 = 10.1, C1 = 0.08, E1 = 1.80, E2 = 1.33, SEF = 1.00;
  const double KAPPA = 0.40, SMALL = 1.E-8, SURFL = 2.E+5;
  double* Q2F = a.q2f; double* Q2LF = a.q2lf; double* L = a.l; double* KM = a.km; double* KH = a.kh; double* KQ = a.kq;
  const long m1 = g.m1;
  if (i == 0) {
    L[NI(g, 0, 1)] = 0.0;
    L[NI(g, 0, KB)] = 0.0;
    for (int k = 1; k <= KB; ++k) {
      const double kn0 = L[NI(g, 0, k)] * sqrt(fabs(a.q2[NI(g, 0, k)] + SMALL));
      KQ[NI(g, 0, k)] = 0.5 * (kn0 * KAPPA * 0.0 + KQ[NI(g, 0, k)]);
      KM[NI(g, 0, k)] = 0.5 * (kn0 * 0.0 + KM[NI(g, 0, k)]);
      KH[NI(g, 0, k)] = 0.5 * (kn0 * 0.0 * g.vprnu + KH[NI(g, 0, k)]);
    }
    for (int k = 1; k <= KB; ++k) {
      a.q2[NI(g, 0, k)] = Q2F[NI(g, 0, k)];
      a.q2l[NI(g, 0, k)] = Q2LF[NI(g, 0, k)];
    }
    return;
  }
  const double dti = g.dti, umol = g.umol, d = a.d[i];
  const int o = g.nbve_off[i], nt = g.ntve[i];
  const double ramp = a.clk->ramp_int;
  double wus = 0.0, wvs = 0.0, wub = 0.0, wvb = 0.0;
  for (int j = 0; j < nt; ++j) {
    const int e = g.nbve[o + j];
    wus = wus + (g.wind_on ? -g.wsx[e] * ramp * 0.001 : 0.0);
    wvs = wvs + (g.wind_on ? -g.wsy[e] * ramp * 0.001 : 0.0);
  }
  const double ntd = (double)nt;
  wus = wus / ntd; wvs = wvs / ntd;
  const double utau2 = sqrt(sq(wus) + sq(wvs));
  const double l0 = SURFL * utau2 / g.grav_n[i];
  for (int j = 0; j < nt; ++j) {
    const int e = g.nbve[o + j];
    wub = wub + a.wubot[e];
    wvb = wvb + a.wvbot[e];
  }
  wub = wub / ntd; wvb = wvb / ntd;
  const double q2f_kb = sqrt(sq(wub) + sq(wvb)) * g.const1;
  Q2F[NI(g, i, KB)] = q2f_kb;
  int re[RMAX];
  static_for<RMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    re[j] = j < nt ? g.nbve[o + j] : 0;
  });
  // ring values loaded unconditionally (slots past the ring read element
  // 0) so that the loads issue together; the sum takes the ring in order
  auto ring_avg = [&](const double* f, long eo) {
    double fv[RMAX];
    static_for<RMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      fv[j] = f[eo + re[j]];
    });
    double x = 0.0;
    static_for<RMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      if (j < nt) x = x + fv[j];
    });
    return x / ntd;
  };
  const double coef4 = 18.0 * A1 * A1 + 9.0 * A1 * A2, coef5 = 9.0 * A1 * A2;
  // stability functions and the under-relaxed mixing coefficients at level k
  auto mix = [&](long q, const Q25Level& lv, double kq_old) {
    const double coef1 = A2 * (1.0 - 6.0 * A1 / B1 * lv.stf);
    const double coef2 = 3.0 * A2 * B2 / lv.stf + 18.0 * A1 * A2;
    const double coef3 = A1 * (1.0 - 3.0 * C1 - 6.0 * A1 / B1 * lv.stf);
    const double sh = coef1 / (1.0 - coef2 * lv.gh);
    const double sm = (coef3 + sh * coef4 * lv.gh) / (1. - coef5 * lv.gh);
    const double kn = lv.ll * sqrt(lv.q2);
    KQ[q] = 0.5 * (kn * KAPPA * sm + kq_old);
    KM[q] = 0.5 * (kn * sm + KM[q]);
    KH[q] = 0.5 * (kn * sh * g.vprnu + KH[q]);
    L[q] = lv.ll;
  };
  const double z1 = ZN(i, 1), zkb = ZN(i, KB);
  constexpr int NVH = VHMODE == 1 ? KMAX + 1 : 1;
  double VH1[NVH], VH2[NVH];
  double vh1p = 0.0, vh2p = 0.0;             // VH(k-1) of the two solves
  const double vhp1_1 = g.cvhp1 * utau2;
  double vhp1 = vhp1_1, vhp2 = 0.0;          // VHP(k-1) of the two solves
  // level 1
  double rho_km1 = a.rho1[NI(g, i, 1)];
  double kq_km1 = KQ[NI(g, i, 1)];           // old KQ(k-1)
  double kq_k = KQ[NI(g, i, 2)];             // old KQ(k)
  {
    const long q = NI(g, i, 1);
    const Q25Level lv = q25_eval(g, i, 1, d, l0, a.q2[q], a.q2l[q], 0.0, 0.0);
    mix(q, lv, kq_km1);
  }
  double ukm1 = 0.0, vkm1 = 0.0;
  if (d > 0.0) { ukm1 = ring_avg(a.u, 0); vkm1 = ring_avg(a.v, 0); }
  auto fwd = [&](const int k) __attribute__((always_inline)) {
    {
      const long q = (long)(k - 1) * m1 + i;
      const double rho_k = a.rho1[q];
      const double kq_kp1 = KQ[q + m1];
      const Q25Level lv = q25_eval(g, i, k, d, l0, a.q2[q], a.q2l[q], rho_km1, rho_k);
      // PROD (vdif_q.F:259-302)
      const double boy = lv.boy;
      const double kmq = KM[q];
      double pr = kmq * 0.0 * fv_pow15(.5 * (-boy + fabs(boy)));
      if (d > 0.0) {
        const long eo = (long)(k - 1) * g.n1;
        const double uk = ring_avg(a.u, eo), vk = ring_avg(a.v, eo);
        pr = pr + kmq * SEF * (sq(uk - ukm1) + sq(vk - vkm1)) / sq(DZZ(i, k - 1) * d);
        pr = pr + KH[q] * boy;
        ukm1 = uk;
        vkm1 = vk;
      }
      // the threshold before the solves (vdif_q.F:136-145): both to SMALL
      // if either is at or below it
      double q2f = Q2F[q], q2lf = Q2LF[q];
      if (q2f <= SMALL || q2lf <= SMALL) { q2f = SMALL; q2lf = SMALL; }
      const double ak = -0.5 * dti * (kq_kp1 + kq_k + 2.0 * umol) / (DZZ(i, k - 1) * DZ(i, k) * d * d);
      const double ck = -0.5 * dti * (kq_km1 + kq_k + 2.0 * umol) / (DZZ(i, k - 1) * DZ(i, k - 1) * d * d);
      const double t1 = 1. / (ak + ck * (1. - vh1p) - (2. * dti * lv.dtef + 1.));
      const double vh1 = ak * t1;
      const double p1 = (-2. * dti * pr + ck * vhp1 - q2f) * t1;
      double vh2, p2;
      if (d > 0.0) {
        const double zk = ZN(i, k);
        const double dtef = lv.dtef * (1. + E2 * sq((1. / fabs(zk - z1) + 1. / fabs(zk - zkb)) * lv.ll / (d * KAPPA)));
        const double t2 = 1. / (ak + ck * (1. - vh2p) - (dti * dtef + 1.));
        vh2 = ak * t2;
        p2 = (dti * (-pr * lv.ll * E1) + ck * vhp2 - q2lf) * t2;
      } else {              // D <= 0: VH/VHP keep the first solve's values
        vh2 = vh1;
        p2 = p1;
      }
      if constexpr (VHMODE == 1) { VH1[k] = vh1; VH2[k] = vh2; }
      else vhs[q] = make_double2(vh1, vh2);
      Q2F[q] = p1;
      Q2LF[q] = p2;
      vh1p = vh1;
      vh2p = vh2;
      vhp1 = p1;
      vhp2 = p2;
      mix(q, lv, kq_k);
      rho_km1 = rho_k;
      kq_km1 = kq_k;
      kq_k = kq_kp1;
    }
  };
  if constexpr (VHMODE != 2) {
    static_for<KMAX - 1>([&](auto kc) {
      constexpr int k = decltype(kc)::value + 2;
      if (k <= KBM1) fwd(k);
    });
  } else {
    for (int k = 2; k <= KBM1; ++k) fwd(k);
  }
  {
    // level KB (kq_k is the old KQ(KB) here)
    const long q = NI(g, i, KB);
    const Q25Level lv = q25_eval(g, i, KB, d, l0, a.q2[q], a.q2l[q], 0.0, 0.0);
    mix(q, lv, kq_k);
  }
  // upward sweep: both back-substitutions, then the threshold on each level
  double x1 = q2f_kb, x2 = Q2LF[NI(g, i, KB)];
  a.q2[NI(g, i, KB)] = x1;
  a.q2l[NI(g, i, KB)] = x2;
  auto bwd = [&](const int k) __attribute__((always_inline)) {
    {
      const long q = (long)(k - 1) * m1 + i;
      double v1 = 0.0, v2 = 0.0, p1 = vhp1_1, p2 = 0.0;   // VH(1) = 0, VHP(1) of the two solves
      if (k >= 2) {
        if constexpr (VHMODE == 1) { v1 = VH1[k]; v2 = VH2[k]; }
        else { const double2 w = vhs[q]; v1 = w.x; v2 = w.y; }
        p1 = Q2F[q];
        p2 = Q2LF[q];
      }
      x1 = v1 * x1 + p1;
      x2 = v2 * x2 + p2;
      double y1 = x1, y2 = x2;
      if (k >= 2 && (x1 <= SMALL || x2 <= SMALL)) { y1 = SMALL; y2 = SMALL; }
      Q2F[q] = y1;
      Q2LF[q] = y2;
      a.q2[q] = y1;
      a.q2l[q] = y2;
    }
  };
  if constexpr (VHMODE != 2) {
    static_for<KMAX>([&](auto kc) {
      constexpr int k = KMAX - decltype(kc)::value;
      if (k <= KBM1) bwd(k);
    });
  } else {
    for (int k = KBM1; k >= 1; --k) bwd(k);
  }
}
