// fvgpu_kernels_edge.h -- the 3-D momentum edge loops.
//
// ADVECTION_EDGE_GCN (advection_edge_gcn.F:397-734) and ADV_UV_EDGE_GCN
// (adv_uv_edge_gcn.F:512-848) loop over edges and scatter each edge's terms
// to its two elements; ADX2D/ADY2D are their column sums
// (internal_step.F:374-384), and ADV_UV finishes each element with the
// internal barotropic pressure gradient, vertical advection, the baroclinic
// gradient, Coriolis, OBC zeroing and rivers (adv_uv_edge_gcn.F:776-1194).
// The two differ only in how the upwind flux is grouped, which is kept as
// each reference file writes it.
#pragma once
#include "fvgpu_kernels_cv.h"

// One workgroup per block of at most 128 own elements (the block tables
// UvBlocks of DevGrid: for ADVECTION the compact element tiles of the
// recursive coordinate bisection, for ADV_UV runs of consecutive
// elements).  Per layer, lane t
//   1. forms the least-squares gradient record of the block's t-th element
//      (its own elements first, then the other element of each of their
//      edges) into LDS: U, V, the four gradient terms COFA1/2/5/6 and the
//      element's viscosity factor ART*sqrt(..)*CC_HVC;
//   2. evaluates the block's t-th edge once from the records of its IA and
//      IB, each in its reference role, into LDS;
//   3. finishes one component of one own element: lanes 0..127 the x
//      sums of the block's own elements, lanes 128..255 the y sums, each
//      summing its three edges in increasing edge index -- the reference's
//      scatter order -- with the reference's operations (a component's
//      operands are selected, never combined), so the sums are
//      bit-identical.
// Each element's gradient is formed in its own block and in each block it
// borders, each edge on a block boundary in both blocks, from the same
// operands, so every copy has the same bits; no record or edge term
// crosses HBM.  The velocities a lane's record needs are loaded a layer
// ahead, and the edge and element parts' inputs before the record
// arithmetic, so that their latency overlaps the barriers.
template <typename FT>
struct UvRecsLds {   // one array per quantity: a wave's reads of one spread over the banks
  FT u[256], v[256], c1[256], c2[256], c5[256], c6[256], p[256];
};
template <typename FT>
struct UvFluxLds {
  FT xadv[256], yadv[256], fx[256], fy[256], visc[256];
};

template <int MODE, typename FT>
__global__ __launch_bounds__(256) void k_uv_blk(DevGrid g, IntArgs a) {
  __shared__ UvRecsLds<FT> sr;
  __shared__ UvFluxLds<FT> sf;
  const int b = xcd_block(), t = threadIdx.x;
  const FT half = (FT)0.5, one = (FT)1.0;
  const double* U = a.u;
  const double* V = a.v;
  if (MODE == 1 && b == 0 && t == 0)
    for (int k = 1; k <= g.KB; ++k) a.viscofm[EI(g, 0, k)] = 0.0;
  // 1. record part: element x
  const UvBlocks& B = MODE == 0 ? g.ub_tile : g.ub_run;
  const bool ract = t < B.ns[b];
  const int x = ract ? B.elems[256L * b + t] : 0;
  const int x1 = NBE_(x, 1), x2 = NBE_(x, 2), x3 = NBE_(x, 3);
  FT ra1[4], ra2[4];
  lsq_coef<FT>(g, x, ra1, ra2);
  const FT rart = (FT)g.art[x], rcc = (FT)g.cc_hvc[x];
  // 2. edge part: edge c
  const bool eact = t < B.ne[b];
  const int c = eact ? B.edges[256L * b + t] : 1;
  const int pl = eact ? B.eloc[256L * b + t] : 0;
  const int sa = pl & 255, sb = (pl >> 8) & 255;
  const int ia = IEC_(c, 1), ib = IEC_(c, 2), j1 = IENODE_(c, 1), j2 = IENODE_(c, 2);
  const FT xija = (FT)(g.xijc[c] - g.xc[ia]), yija = (FT)(g.yijc[c] - g.yc[ia]);
  const FT xijb = (FT)(g.xijc[c] - g.xc[ib]), yijb = (FT)(g.yijc[c] - g.yc[ib]);
  const FT dlx = (FT)g.dltxc[c], dly = (FT)g.dltyc[c];
  const FT visc0 = (FT)g.fm1 * half * ((FT)g.cc_hvc[ia] + (FT)g.cc_hvc[ib]);
  const FT fact = (FT)g.fact, hprnu = (FT)g.hprnu;
  const double dtj1 = a.dt[j1], dtj2 = a.dt[j2];
  const FT isbc_tmp = (FT)g.isbc[c];
  // 3. element part: component cy (0: x, 1: y) of own element i
  const bool cy = t >= 128;
  const int tl = cy ? t - 128 : t;
  const bool oact = tl < B.nown[b];
  const int i = oact ? B.elems[256L * b + tl] : 1;   // the block's own elements are its first records
  const int io = i;
  const int ep3 = B.epos[io];
  int slot[3];
  bool flip[3];       // the element is IB of the edge
  double tp[3], elij[3], dlp[3];
  const double ep = g.epor[io];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int ent = g.ed_ent[3 * io + q], e = ent >> 1;
    slot[q] = (ep3 >> (8 * q)) & 255;
    flip[q] = (ent & 1) != 0;
    tp[q] = (double)(1 - g.isbc[e]) * ep;
    dlp[q] = cy ? g.dltxc[e] : g.dltyc[e];   // the pressure term's edge component
    elij[q] = MODE == 1 ? 0.5 * (a.egf[IENODE_(e, 1)] + a.egf[IENODE_(e, 2)]) : 0.0;
  }
  const bool obc = g.isbce[io] == 2;
  const int r0 = g.re_off[io], r1 = g.re_off[io + 1];
  const double art = g.art[io], dt1 = a.dt1[io], d1 = a.d1[io], cor = g.cor[io], ge = g.grav_e[io];
  const double* C = cy ? V : U;      // the component's velocity
  const double* O = cy ? U : V;      // the other component (Coriolis)
  const double* R = cy ? a.drhoy : a.drhox;
  double sc = 0.0;
  // the record's velocities of the coming layer, loaded a layer ahead
  double pu[4], pv[4];
  const int xs[4] = {x, x1, x2, x3};
#pragma unroll
  for (int q = 0; q < 4; ++q) { pu[q] = U[xs[q]]; pv[q] = V[xs[q]]; }
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.n1;
    // this layer's inputs of the edge and element parts, issued before the
    // record arithmetic and the barriers (every index is valid for idle lanes)
    const double dzj1 = DZ(j1, k), dzj2 = DZ(j2, k);
    const long kp = ko + g.n1, km = k != 1 ? ko - g.n1 : ko;
    double ck = 0.0, ok = 0.0, wk = 0.0, wp = 0.0, cp = 0.0, cm = 0.0, rk = 0.0;
    if (MODE == 1) {
      ck = C[ko + io]; ok = O[ko + io]; wk = a.w[ko + io]; wp = a.w[kp + io];
      cp = C[kp + io]; cm = C[km + io]; rk = R[ko + io];
    }
    const double dzk = DZ1(io, k), dzp = DZ1(io, k + 1), dzm = DZ1(io, k != 1 ? k - 1 : k);
    if (ract) {
      const FT us = (FT)pu[0], vs = (FT)pv[0];
      const FT u1 = (FT)pu[1], u2 = (FT)pu[2], u3 = (FT)pu[3];
      const FT v1 = (FT)pv[1], v2 = (FT)pv[2], v3 = (FT)pv[3];
      const FT c1 = ra1[0] * us + ra1[1] * u1 + ra1[2] * u2 + ra1[3] * u3;
      const FT c2 = ra2[0] * us + ra2[1] * u1 + ra2[2] * u2 + ra2[3] * u3;
      const FT c5 = ra1[0] * vs + ra1[1] * v1 + ra1[2] * v2 + ra1[3] * v3;
      const FT c6 = ra2[0] * vs + ra2[1] * v1 + ra2[2] * v2 + ra2[3] * v3;
      const FT vis = rart * sqrt(sq(c1) + sq(c6) + half * sq(c2 + c5));
      sr.u[t] = us; sr.v[t] = vs;
      sr.c1[t] = c1; sr.c2[t] = c2; sr.c5[t] = c5; sr.c6[t] = c6;
      sr.p[t] = vis * rcc;
    }
    if (k < g.KBM1) {
#pragma unroll
      for (int q = 0; q < 4; ++q) { pu[q] = U[kp + xs[q]]; pv[q] = V[kp + xs[q]]; }
    }
    __syncthreads();
    if (eact) {
      const FT dij = (FT)(0.5 * (dtj1 * dzj1 + dtj2 * dzj2));
      const FT uia = sr.u[sa], vIA = sr.v[sa], uib = sr.u[sb], vIB = sr.v[sb];
      const FT cofa1 = sr.c1[sa], cofa2 = sr.c2[sa], cofa5 = sr.c5[sa], cofa6 = sr.c6[sa];
      const FT cofa3 = sr.c1[sb], cofa4 = sr.c2[sb], cofa7 = sr.c5[sb], cofa8 = sr.c6[sb];
      const FT pa = sr.p[sa], pb = sr.p[sb];
      const FT uij1 = uia + cofa1 * xija + cofa2 * yija;
      const FT vij1 = vIA + cofa5 * xija + cofa6 * yija;
      const FT uij2 = uib + cofa3 * xijb + cofa4 * yijb;
      const FT vij2 = vIB + cofa7 * xijb + cofa8 * yijb;
      const FT uij = half * (uij1 + uij2), vij = half * (vij1 + vij2);
      const FT viscof = (fact * half * (pa + pb) + visc0) / hprnu;
      const FT txxij = (cofa1 + cofa3) * viscof, tyyij = (cofa6 + cofa8) * viscof;
      const FT txyij = half * (cofa2 + cofa4 + cofa5 + cofa7) * viscof;
      const FT fxx = dij * (txxij * dly - txyij * dlx);
      const FT fyy = dij * (txyij * dly - tyyij * dlx);
      FT xadv_, yadv_;
      if (MODE == 0) {
        const FT un_tmp = vij * dlx - uij * dly;
        const FT sgn = fsign1(un_tmp);
        xadv_ = dij * un_tmp * ((one - sgn) * uij2 + (one + sgn) * uij1) * half;
        yadv_ = dij * un_tmp * ((one - sgn) * vij2 + (one + sgn) * vij1) * half;
      } else {
        const FT exflux = dij * (-uij * dly + vij * dlx);
        const FT sgn = fsign1(exflux);
        xadv_ = exflux * ((one - sgn) * uij2 + (one + sgn) * uij1) * half;
        yadv_ = exflux * ((one - sgn) * vij2 + (one + sgn) * vij1) * half;
        sf.visc[t] = viscof;
      }
      sf.xadv[t] = xadv_;
      sf.yadv[t] = yadv_;
      sf.fx[t] = fxx + (FT)3.0 * fxx * isbc_tmp;
      sf.fy[t] = fyy + (FT)3.0 * fyy * isbc_tmp;
    }
    __syncthreads();
    if (oact) {
      // the gathers' sums of one component: x as (xf, px), y as (yf, py)
      double f = 0.0, pr = 0.0, vm = 0.0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int s = slot[q];
        const double adv = (double)(cy ? sf.yadv[s] : sf.xadv[s]);
        const double fl = (double)(cy ? sf.fy[s] : sf.fx[s]);
        const double av = adv * tp[q], bv = fl * ep;
        f = flip[q] ? (f - av) - bv : (f + av) + bv;
        if (MODE == 1) {
          if (!cy) vm = vm + (double)sf.visc[s];
          const double gd = ge * dt1 * dzk * elij[q];
          const double gz = gd * dlp[q];
          // x: IA subtracts G*DLTYC, IB adds; y: IA adds G*DLTXC, IB subtracts
          pr = (flip[q] != cy) ? pr + gz : pr - gz;
        }
      }
      if (MODE == 0) {
        if (obc) f = 0.0;
        for (int p = r0; p < r1; ++p) {
          const int r = g.re_riv[p];
          const int j = g.inodeq[r];
          const bool edge = g.icellq[r] > 0;   // an edge river's momentum enters its cell whole
          const double vlctyq = a.qdis[r] / g.qarea[r];
          const double vq = g.vqdist[(long)(k - 1) * g.NUMQBC + r];
          const double temp = (edge ? 1.0 : 0.5) * a.qdis[r] * vq * vq * vlctyq / (edge ? DZ1(i, k) : DZ(j, k));
          f = f - temp * (cy ? g.sin_angleq[r] : g.cos_angleq[r]);
        }
        sc = sc + f;
      } else {
        f = f + pr;
        // vertical advection (adv_uv_edge_gcn.F:927-946)
        double fluxv;
        if (k == 1) {
          fluxv = -wp * (ck * dzp + cp * dzk) / (dzk + dzp);
        } else if (k == g.KBM1) {
          fluxv = wk * (ck * dzm + cm * dzk) / (dzk + dzm);
        } else {
          fluxv = wk * (ck * dzm + cm * dzk) / (dzk + dzm) - wp * (ck * dzp + cp * dzk) / (dzk + dzp);
        }
        // Coriolis: x subtracts COR*V*.., y adds COR*U*..
        const double fa = f + fluxv * art + rk;
        const double co = cor * ok * dt1 * dzk * art;
        f = cy ? fa + co : fa - co;
        if (obc) f = 0.0;
        for (int p = r0; p < r1; ++p) {
          const int r = g.re_riv[p];
          const int j = g.inodeq[r];
          const bool edge = g.icellq[r] > 0;   // an edge river's momentum enters its cell whole
          const double vlctyq = a.qdis[r] / g.qarea[r];
          const double vq = g.vqdist[(long)(k - 1) * g.NUMQBC + r];
          const double temp = (edge ? 1.0 : 0.5) * a.qdis[r] * vq * vq * vlctyq / (edge ? DZ1(i, k) : DZ(j, k));
          f = f - temp * (cy ? g.sin_angleq[r] : g.cos_angleq[r]);
        }
        (cy ? a.vf : a.uf)[ko + i] = ck * dt1 / d1 - g.dti * f / art / (d1 * dzk);
        if (!cy) a.viscofm[ko + i] = vm / art;
      }
    }
  }
  if (oact) {
    if (MODE == 0) {
      (cy ? a.ady2d : a.adx2d)[i] = sc;
    } else if (!cy) {
      a.viscofm[EI(g, i, g.KB)] = 0.0 / art;
    }
  }
}
