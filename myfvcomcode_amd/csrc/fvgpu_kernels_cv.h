// fvgpu_kernels_cv.h -- control-volume flux kernels of the internal mode in
// two phases, and VERTVL with its sub-edge sums in registers.
//
// The reference's scalar advection (ADV_T/ADV_S, ADV_Q) loops over the
// control-volume sub-edges and scatters each face flux to its two nodes.
// Here:
//   1. k_cvrec     one thread per node: ring gradient, ring bounds and the
//                  viscosity factor of every layer, packed per node-layer
//                  (CvRec), the node's ring kept in registers across layers;
//   2. k_cvflux_*  one thread per element: the face fluxes of its three
//                  sub-edges at every layer, each evaluated once from the
//                  element's three node records (gathered once per layer
//                  for the three faces) and its velocity, written to the
//                  element-major positions 3*(e-1)+j+1 (coalesced);
//   3. k_*_gather  one thread per node: the node's sub-edge fluxes summed in
//                  increasing sub-edge order (the reference's scatter order,
//                  so the sum is bit-identical), then the vertical flux, the
//                  update and FCT, whose ring bounds come from the field
//                  itself (the ring's values are in cache beside the node's
//                  own column) rather than through HBM from phase 1.
// Node lists are read slot-major (cvl_ent) so consecutive nodes read
// consecutive addresses, and workgroups are placed XCD-contiguously.
#pragma once
#include "fvgpu_kernels_column.h"


// Ring gradient / bounds of field f at node i for the layers [k0, k1]
// (adv_t.F:219-376 and :469-480; adv_q.F the same on layer interfaces).
// RMAX >= NTSN bounds the ring held in registers.  The sums are FP64; the
// record is stored in the flux type FT.
// kind 0: tracer, VHN = VISCOFH*NN_HVC; kind 1: turbulence, VHN = the
// DZ-weighted interface viscosity (adv_q.F) times NN_HVC.
// Ring of node i (RMAX >= NTSN): node indices and the DLTXTRIE/DLTYTRIE
// factors of its ring segments, in registers across the layers.
template <int RMAX>
struct RingGeo {
  int rn[RMAX];
  double rdx[RMAX], rdy[RMAX];
  int nt;
};

template <int RMAX>
__device__ __forceinline__ void ring_geo_load(const DevGrid& g, int i, RingGeo<RMAX>& r) {
  const int o = g.nbsn_off[i], nt = g.ntsn[i];
  r.nt = nt;
  static_for<RMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    r.rn[j] = j < nt ? g.nbsn[o + j] : 0;
    r.rdx[j] = j < nt - 1 ? g.dltxtrie[o + j] : 0.0;
    r.rdy[j] = j < nt - 1 ? g.dltytrie[o + j] : 0.0;
  });
}

// The record of node i at layer k: ring gradient of f and of f - mean
// (adv_t.F:219-376; adv_q.F:185-230 with the constant mean QMEAN1), the
// ring bounds with the node (adv_t.F:469-480) and the viscosity factor.
// kind 0: tracer, VHN = VISCOFH*NN_HVC; kind 1: turbulence, VHN = the
// DZ-weighted interface viscosity (adv_q.F) times NN_HVC.  The sums are
// FP64; the record is stored in the flux type FT.
//
// The values the record reads -- the field at the node and its ring, and
// for the tracer the mean at the ring and VISCOFH at the layer -- are
// loaded by cv_load, so that a caller can issue them a layer ahead.
// (ADV_Q's record keeps its mean and its VISCOFH loads, of two levels, in
// the record: loaded ahead they measured slower.)
template <int RMAX>
struct CvLoads {
  double fv[RMAX], mv[RMAX];
  double c, vh;
};

template <int KIND, int RMAX>
__device__ __forceinline__ void cv_load(const DevGrid& g, const IntArgs& a, int i, const RingGeo<RMAX>& rg,
                                        const double* f, const double* mean, double cmean, int k, CvLoads<RMAX>& L) {
  const long ko = (long)(k - 1) * g.m1;
  // the ring's values are loaded unconditionally (slots past the ring
  // read node 0) so that the loads issue together
  static_for<RMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    L.fv[j] = f[ko + rg.rn[j]];
    if (KIND == 0) L.mv[j] = mean ? mean[ko + rg.rn[j]] : cmean;
  });
  L.c = f[ko + i];
  if (KIND == 0) L.vh = a.viscofh[ko + i];
}

template <int KIND, int RMAX, typename FT>
__device__ __forceinline__ CvRecT<FT> cv_record_of(const DevGrid& g, const IntArgs& a, int i, const RingGeo<RMAX>& rg,
                                                   double art2, double nn, const CvLoads<RMAX>& L,
                                                   const double* mean, double cmean, int k) {
  const long ko = (long)(k - 1) * g.m1;
  const int nt = rg.nt;
  const double* fv = L.fv;
  // the sums take the ring's own terms in ring order
  double mv[RMAX];
  static_for<RMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    mv[j] = KIND == 0 ? L.mv[j] : (mean ? mean[ko + rg.rn[j]] : cmean);
  });
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
  double f1 = fv[0];
  double m1v = mv[0];
  double lo = f1, hi = f1;
  static_for<RMAX - 1>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < nt - 1) {
      const double f2 = fv[j + 1];
      const double m2v = mv[j + 1];
      const double ffd = 0.5 * (f1 + f2 - m1v - m2v);
      const double ff1 = 0.5 * (f1 + f2);
      s1 = s1 + ff1 * rg.rdy[j];
      s2 = s2 + ff1 * rg.rdx[j];
      s3 = s3 + ffd * rg.rdy[j];
      s4 = s4 + ffd * rg.rdx[j];
      if (j >= 1) { lo = fmin(lo, f1); hi = fmax(hi, f1); }
      f1 = f2;
      m1v = m2v;
    }
  });
  const double c = L.c;
  double vhn;
  if (KIND == 0) {
    vhn = L.vh * nn;
  } else {
    const double vfa = (a.viscofh[ko + i] * DZ(i, k - 1) + a.viscofh[ko - g.m1 + i] * DZ(i, k)) /
                       (DZ(i, k) + DZ(i, k - 1));
    vhn = vfa * nn;
  }
  CvRecT<FT> r;
  r.f = (FT)c;
  r.gx = (FT)(s1 / art2);
  r.gy = (FT)(s2 / art2);
  r.fmin = (FT)fmin(lo, c);
  r.fmax = (FT)fmax(hi, c);
  r.gxd = (FT)(s3 / art2);
  r.gyd = (FT)(s4 / art2);
  r.vhn = (FT)vhn;
  return r;
}

template <int KIND, int RMAX, typename FT>
__device__ __forceinline__ CvRecT<FT> cv_record(const DevGrid& g, const IntArgs& a, int i, const RingGeo<RMAX>& rg,
                                                double art2, double nn, const double* f, const double* mean,
                                                double cmean, int k) {
  CvLoads<RMAX> L;
  cv_load<KIND, RMAX>(g, a, i, rg, f, mean, cmean, k, L);
  return cv_record_of<KIND, RMAX, FT>(g, a, i, rg, art2, nn, L, mean, cmean, k);
}

// Ring gradient / bounds of field f at node i for the layers [k0, k1], one
// thread per node with its ring in registers.
template <int KIND, int RMAX, typename FT>
__global__ __launch_bounds__(256) void k_cvrec(DevGrid g, IntArgs a, const double* f, const double* mean,
                                               double cmean, int k0, int k1) {
  const int i = (int)xcd_thread();
  if (i < 1 || i > g.M) return;
  CvRecT<FT>* rec = (CvRecT<FT>*)a.rec;
  RingGeo<RMAX> rg;
  ring_geo_load<RMAX>(g, i, rg);
  const double art2 = g.art2[i], nn = g.nn_hvc[i];
  for (int k = k0; k <= k1; ++k)
    rec[(long)(k - 1) * g.m1 + i] = cv_record<KIND, RMAX, FT>(g, a, i, rg, art2, nn, f, mean, cmean, k);
}

// Min and max of field f over the NBSN ring of node i at layer offset ko,
// the ring alone (entry NTSN repeats entry 1), in the order k_cvrec forms
// them: the FCT bounds (fct_t.F:74-80, fct_q2.F:60-66).
template <int RMAX>
__device__ __forceinline__ void ring_bounds(const int (&rn)[RMAX], int nt, const double* f, long ko, double& lo,
                                            double& hi) {
  lo = f[ko + rn[0]];
  hi = lo;
  // Slots past the ring repeat its first node (min and max are unchanged),
  // so every load is unconditional and all of them issue before the first
  // wait.
  static_for<RMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j >= 1) {
      const double v = f[ko + (j < nt - 1 ? rn[j] : rn[0])];
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
  });
}

template <int RMAX>
__device__ __forceinline__ int load_ring(const DevGrid& g, int i, int (&rn)[RMAX]) {
  const int o = g.nbsn_off[i], nt = g.ntsn[i];
  static_for<RMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    rn[j] = j < nt ? g.nbsn[o + j] : 0;
  });
  return nt;
}

// Face flux of ADV_T/ADV_S through sub-edge c at layer k (adv_t.F:440-500),
// in the flux type FT (FP64: bit-identical to the reference's expression).
template <typename FT>
struct CvFace { FT exflux, fxx, fyy; };

template <typename FT>
__device__ __forceinline__ CvFace<FT> cv_face(FT un, FT dtij, FT dltx, FT dlty, FT x1, FT y1, FT x2, FT y2,
                                              FT fact, FT visc0, const CvRecT<FT>& ra, const CvRecT<FT>& rb) {
  const FT half = (FT)0.5, one = (FT)1.0;
  FT fij1 = ra.f + x1 * ra.gx + y1 * ra.gy;
  FT fij2 = rb.f + x2 * rb.gx + y2 * rb.gy;
  if (fij1 < ra.fmin) fij1 = ra.fmin;
  if (fij1 > ra.fmax) fij1 = ra.fmax;
  if (fij2 < rb.fmin) fij2 = rb.fmin;
  if (fij2 > rb.fmax) fij2 = rb.fmax;
  const FT viscof = fact * half * (ra.vhn + rb.vhn) + visc0;
  const FT txx = half * (ra.gxd + rb.gxd) * viscof;
  const FT tyy = half * (ra.gyd + rb.gyd) * viscof;
  CvFace<FT> r;
  r.fxx = -dtij * txx * dlty;
  r.fyy = dtij * tyy * dltx;
  const FT sgn = fsign1(un);
  r.exflux = -un * dtij * ((one + sgn) * fij2 + (one - sgn) * fij1) * half + r.fxx + r.fyy;
  return r;
}

// The three sub-edges of element e in edge slots (slot j joins the
// element's nodes j and j+1 mod 3, called A and B): geometry in the flux
// type, with the interpolation offsets of A and B (DLTXNCVE/DLTYNCVE of the
// side that node is on) and whether IA is B.
template <typename FT>
struct ElemFaces {
  FT dltx[3], dlty[3], xa[3], ya[3], xb[3], yb[3], visc0[3];
  bool swap[3];
};

template <typename FT>
__device__ __forceinline__ ElemFaces<FT> elem_faces(const DevGrid& g, int e) {
  ElemFaces<FT> f;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int c = g.ecv_c[(long)j * g.n1 + e];
    const bool sw = g.ecv_swap[(long)j * g.n1 + e] != 0;
    f.swap[j] = sw;
    f.dltx[j] = (FT)g.dltxe[c];
    f.dlty[j] = (FT)g.dltye[c];
    const FT x1 = (FT)DLTXNCVE_(c, 1), y1 = (FT)DLTYNCVE_(c, 1), x2 = (FT)DLTXNCVE_(c, 2), y2 = (FT)DLTYNCVE_(c, 2);
    f.xa[j] = sw ? x2 : x1;
    f.ya[j] = sw ? y2 : y1;
    f.xb[j] = sw ? x1 : x2;
    f.yb[j] = sw ? y1 : y2;
    f.visc0[j] = (FT)(g.fm1 * 0.5 * (g.nn_hvc[NIEC_(c, 1)] + g.nn_hvc[NIEC_(c, 2)]));
  }
  return f;
}

// Face value of one side, limited to that node's ring bounds
// (adv_t.F:447-466).
template <typename FT>
__device__ __forceinline__ FT face_value(const CvRecT<FT>& r, FT x, FT y) {
  FT v = r.f + x * r.gx + y * r.gy;
  if (v < r.fmin) v = r.fmin;
  if (v > r.fmax) v = r.fmax;
  return v;
}

// The three face fluxes of one element at one layer (adv_t.F:440-500,
// adv_q.F:200-330) from the records r[3] of its nodes NV(e,1..3), the
// normal velocity (un = V*DLTXE - U*DLTYE) from the element velocity (uk,
// vk) and the layer thickness dtij; viscosity divided by HPRNU when DIVP.
template <typename FT, bool DIVP>
__device__ __forceinline__ void cv_faces(const ElemFaces<FT>& f, const CvRecT<FT> (&r)[3], FT uk, FT vk, FT dtij,
                                         FT fact, FT rhprnu, FT* out) {
  const FT half = (FT)0.5, one = (FT)1.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const CvRecT<FT>& ra = r[j];
    const CvRecT<FT>& rb = r[(j + 1) % 3];
    const FT fa = face_value(ra, f.xa[j], f.ya[j]), fb = face_value(rb, f.xb[j], f.yb[j]);
    const FT fij1 = f.swap[j] ? fb : fa, fij2 = f.swap[j] ? fa : fb;
    const FT un = vk * f.dltx[j] - uk * f.dlty[j];
    FT viscof = fact * half * (ra.vhn + rb.vhn) + f.visc0[j];
    if (DIVP && rhprnu != one) viscof = viscof / rhprnu;   // x / 1 is x: skipped when HPRNU = 1
    const FT txx = half * (ra.gxd + rb.gxd) * viscof;
    const FT tyy = half * (ra.gyd + rb.gyd) * viscof;
    const FT fxx = -dtij * txx * f.dlty[j];
    const FT fyy = dtij * tyy * f.dltx[j];
    const FT sgn = fsign1(un);
    out[j] = -un * dtij * ((one + sgn) * fij2 + (one - sgn) * fij1) * half + fxx + fyy;
  }
}

// Phase 2 of ADV_T/ADV_S: the three face fluxes of one element per layer
// (adv_t.F:440-500), with the geometry held in registers across layers.
// The two nodes of a face enter the diffusive terms through sums, which are
// the same in either order, and the upwind value through FIJ1 (side IA) and
// FIJ2 (side IB), exchanged when IA is node B.
template <typename FT>
__global__ __launch_bounds__(256) void k_cvflux_ts(DevGrid g, IntArgs a) {
  const int e = (int)xcd_thread();
  if (e < 1 || e > g.N) return;
  const CvRecT<FT>* rec = (const CvRecT<FT>*)a.rec;
  FT* cvflux = (FT*)a.cvflux;
  const ElemFaces<FT> f = elem_faces<FT>(g, e);
  const int nd[3] = {NV_(e, 1), NV_(e, 2), NV_(e, 3)};
  const FT fact = (FT)g.fact, half = (FT)0.5, one = (FT)1.0;
  const double dt1 = a.dt1[e];
  const long cs = (long)g.NCV + 1, p0 = 3L * (e - 1) + 1;
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.m1, eo = (long)(k - 1) * g.n1;
    const FT dtij = (FT)(dt1 * DZ1(e, k));
    const FT vk = (FT)a.v[eo + e], uk = (FT)a.u[eo + e];
    CvRecT<FT> r[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) r[q] = rec[ko + nd[q]];
    cv_faces<FT, false>(f, r, uk, vk, dtij, fact, one, cvflux + (long)(k - 1) * cs + p0);
  }
}

// Advective part of the horizontal flux at the open-boundary nodes
// (XFLUX_OBC of adv_t.F:507-520).  LP lanes per OBC node and layer (LP >=
// the node's sub-edge count): lane j evaluates the node's j-th sub-edge face
// into LDS, signed by the node's side (x - y and x + (-y) are the same IEEE
// operation), then the first lane sums them in sub-edge order.  Every
// face's loads issue at once instead of one face after another.
template <int LP, typename FT>
__global__ __launch_bounds__(256) void k_adv_ts_obc(DevGrid g, IntArgs a) {
  __shared__ double sadv[256];
  const int lane = threadIdx.x % LP;
  const long t = (long)blockIdx.x * (256 / LP) + threadIdx.x / LP;
  const bool act = t < (long)g.IOBCN * g.KBM1;
  const CvRecT<FT>* rec = (const CvRecT<FT>*)a.rec;
  const int obc = act ? (int)(t % g.IOBCN) : 0, k = act ? (int)(t / g.IOBCN) + 1 : 1;
  const int i = g.i_obc_n[obc];
  const int n = act ? g.cv_cnt[i] : 0;
  const long ko = (long)(k - 1) * g.m1, eo = (long)(k - 1) * g.n1;
  if (lane < n) {
    const int e = g.cv_ent[g.cv_off[i] + lane], c = e >> 1;
    const int ia = NIEC_(c, 1), ib = NIEC_(c, 2), i1 = g.ntrg[c];
    const FT dltx = (FT)g.dltxe[c], dlty = (FT)g.dltye[c];
    const FT dtij = (FT)(a.dt1[i1] * DZ1(i1, k));
    const FT un = (FT)a.v[eo + i1] * dltx - (FT)a.u[eo + i1] * dlty;
    const FT visc0 = (FT)(g.fm1 * 0.5 * (g.nn_hvc[ia] + g.nn_hvc[ib]));
    const CvFace<FT> r = cv_face<FT>(un, dtij, dltx, dlty, (FT)DLTXNCVE_(c, 1), (FT)DLTYNCVE_(c, 1),
                                     (FT)DLTXNCVE_(c, 2), (FT)DLTYNCVE_(c, 2), (FT)g.fact, visc0,
                                     rec[ko + ia], rec[ko + ib]);
    const double adv = (double)(r.exflux - r.fxx - r.fyy);
    sadv[threadIdx.x] = (e & 1) ? -adv : adv;
  }
  __syncthreads();
  if (!act || lane != 0) return;
  double xadv = 0.0;
  for (int j = 0; j < n; ++j) xadv = xadv + sadv[threadIdx.x + j];
  a.xflux_obc[(long)(k - 1) * g.IOBCN + obc] = xadv;
}

// Phase 3 of ADV_T/ADV_S (adv_t.F:507-1052): ordered sub-edge sums (in FP64
// whatever the flux type), bottom gradients for VDIF_TS, vertical flux,
// rivers, the update TF1 = (T1 - XFLUX/ART1*DTI/(DT*DZ))*DT/DTFA and FCT.
template <int SMAX, int RMAX, typename FT>
__global__ __launch_bounds__(256) void k_adv_ts_gather(DevGrid g, IntArgs a, const double* f1, double* ff,
                                                        const double* fdis, int which, int fct) {
  const int i = (int)xcd_thread();
  if (i < 1 || i > g.M) return;
  const bool clip = fct && !g.fct_skip[i];
  const int n = g.cv_cnt[i];
  int ce[SMAX];
  static_for<SMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    ce[j] = j < n ? g.cvl_ent[(long)j * g.m1 + i] : 0;
  });
  const double art1 = g.art1[i];
  const bool isobc = g.isonb[i] == 2;
  const long cs = (long)g.NCV + 1;
  int rn[RMAX];
  const int nt = clip ? load_ring<RMAX>(g, i, rn) : 0;
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.m1;
    const FT* fl = (const FT*)a.cvflux + (long)(k - 1) * cs;
    // all SMAX loads issue together (slots past the node's count read the
    // unused entry 0); the sum takes the node's own terms in sub-edge order
    FT xv[SMAX];
    static_for<SMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      xv[j] = fl[ce[j] >> 1];
    });
    double xflux = 0.0;
    static_for<SMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      if (j < n) {
        const double x = (double)xv[j];
        if (ce[j] & 1) xflux = xflux - x;
        else xflux = xflux + x;
      }
    });
    if (k == g.KBM1) {  // ADV_T and ADV_S both set the bottom-gradient terms VDIF_TS reads
      const CvRecT<FT> r = ((const CvRecT<FT>*)a.rec)[ko + i];
      a.pfpxb[i] = (double)r.gx;
      a.pfpyb[i] = (double)r.gy;
      a.ah_bottom[i] = (g.fact * a.viscofh[ko + i] + g.fm1) * g.nn_hvc[i];
    }
    const double fk = f1[ko + i], dzk = DZ(i, k);
    // DZ-weighted values at the layer's upper (nu/du) and lower (nd/dd)
    // interfaces: the vertical flux (adv_t.F:873-881) and the FCT bounds
    // (fct_t.F:82-101) use the same expressions.  The neighbours are loaded
    // unconditionally (layer 1 reads itself as its upper neighbour, level
    // KB exists below KBM1) so that the loads issue together.
    const long kup = k != 1 ? ko - g.m1 : ko;
    const double fup = f1[kup + i], dzm = DZ(i, k != 1 ? k - 1 : k);
    const double fdn = f1[ko + g.m1 + i], dzp = DZ(i, k + 1);
    const double wk = a.wts[ko + i], wkp = a.wts[ko + g.m1 + i];
    double temp, nu = 0.0, du = 1.0, nd = 0.0, dd = 1.0;
    const bool has_up = k != 1, has_dn = k == 1 || k != g.KBM1;   // the reference's branch order
    if (has_up) {
      nu = fk * dzm + fup * dzk;
      du = dzk + dzm;
    }
    if (has_dn) {
      nd = fk * dzp + fdn * dzk;
      dd = dzk + dzp;
    }
    if (k == 1) temp = -wkp * nd / dd;
    else if (k == g.KBM1) temp = wk * nu / du;
    else temp = wk * nu / du - wkp * nd / dd;
    if (isobc) xflux = temp * art1;
    else xflux = xflux + temp * art1;
    if (!g.river_ts_specified) {
      const int rv = g.node_riv[i];
      if (rv >= 0)
        for (int j = rv; j < g.NUMQBC; ++j) {
          double w;
          if (river_share(g, j, i, w))
            xflux = xflux - a.qdis[j] * w * g.vqdist[(long)(k - 1) * g.NUMQBC + j] * fdis[j];
        }
    }
    if (which == 1) xflux = xflux - 0.0 * art1;
    double x = (fk - xflux / art1 * (g.dti / (a.dt[i] * dzk))) * (a.dt[i] / a.dtfa[i]);
    if (clip) {  // FCT_T / FCT_S (fct_t.F:74-101)
      double tmin, tmax;
      ring_bounds<RMAX>(rn, nt, f1, ko, tmin, tmax);
      if (has_up) { const double v = nu / du; tmax = fmax(tmax, v); tmin = fmin(tmin, v); }
      if (has_dn) { const double v = nd / dd; tmax = fmax(tmax, v); tmin = fmin(tmin, v); }
      if (tmin - x > 0.0) x = tmin;
      if (x - tmax > 0.0) x = tmax;
    }
    ff[ko + i] = x;
  }
}

// Phase 2 of ADV_Q (adv_q.F:200-330): the three face fluxes of one element
// on layer interfaces k = 2..KBM1, the element velocity interpolated to the
// interface with DZ1 weights (once for the three faces); faces as in
// k_cvflux_ts.
template <typename FT>
__global__ __launch_bounds__(256) void k_cvflux_q(DevGrid g, IntArgs a) {
  const int e = (int)xcd_thread();
  if (e < 1 || e > g.N) return;
  const CvRecT<FT>* rec = (const CvRecT<FT>*)a.rec;
  FT* cvflux = (FT*)a.cvflux;
  const ElemFaces<FT> f = elem_faces<FT>(g, e);
  const int nd[3] = {NV_(e, 1), NV_(e, 2), NV_(e, 3)};
  const FT fact = (FT)g.fact, rhprnu = (FT)g.hprnu, half = (FT)0.5, one = (FT)1.0;
  const double dt1 = a.dt1[e];
  const long cs = (long)g.NCV + 1, p0 = 3L * (e - 1) + 1;
  double um = a.u[e], vm = a.v[e];   // layer 1
  for (int k = 2; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.m1, eo = (long)(k - 1) * g.n1;
    const double uk = a.u[eo + e], vk = a.v[eo + e];
    const double dz1k = DZ1(e, k), dz1m = DZ1(e, k - 1), dzz1 = DZZ1(e, k - 1);
    CvRecT<FT> r[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) r[q] = rec[ko + nd[q]];
    // Every load of the layer is issued before the two divisions: without
    // the barrier the scheduler trades them for registers and waits on each
    // record part in turn.
    __builtin_amdgcn_sched_barrier(0);
    const FT uq = (FT)((uk * dz1m + um * dz1k) / (dz1k + dz1m));
    const FT vq = (FT)((vk * dz1m + vm * dz1k) / (dz1k + dz1m));
    um = uk;
    vm = vk;
    const FT dtij = (FT)(dt1 * dzz1);
    cv_faces<FT, true>(f, r, uq, vq, dtij, fact, rhprnu, cvflux + (long)(k - 1) * cs + p0);
  }
}

// The block's records of one layer in LDS, one array per quantity.
template <typename FT>
struct CvRecLds {
  FT f[FV_CVB], gx[FV_CVB], gy[FV_CVB], fmin[FV_CVB], fmax[FV_CVB], gxd[FV_CVB], gyd[FV_CVB], vhn[FV_CVB];
  __device__ __forceinline__ void put(int s, const CvRecT<FT>& r) {
    f[s] = r.f; gx[s] = r.gx; gy[s] = r.gy; fmin[s] = r.fmin;
    fmax[s] = r.fmax; gxd[s] = r.gxd; gyd[s] = r.gyd; vhn[s] = r.vhn;
  }
  __device__ __forceinline__ CvRecT<FT> get(int s) const {
    CvRecT<FT> r;
    r.f = f[s]; r.gx = gx[s]; r.gy = gy[s]; r.fmin = fmin[s];
    r.fmax = fmax[s]; r.gxd = gxd[s]; r.gyd = gyd[s]; r.vhn = vhn[s];
    return r;
  }
};

// ADV_T/ADV_S (KIND 0, layers 1..KBM1) or ADV_Q (KIND 1, interfaces
// 2..KBM1) phases 1 and 2 in one launch.  A workgroup of 256 lanes takes a
// block of at most FV_CVB = 128 elements and the at most 128 distinct nodes
// they use (cvb_*: compact tiles of the mesh).  Its first two waves form
// the nodes' records (one lane per node, ring in registers: cv_record, the
// same arithmetic as k_cvrec) of layer k into one of two LDS buffers while
// its last two waves evaluate the elements' three face fluxes (one lane per
// element: cv_faces, as k_cvflux_*) of layer k-1 from the other buffer; one
// barrier per layer hands the buffers over.  The two roles hold only their
// own registers and run side by side.  A node shared by two blocks is
// formed by both; the records cross HBM only for the nodes whose records
// other kernels read (rec_keep: OBC nodes and their sub-edge neighbours)
// and, for the tracer, at the bottom layer (the bottom gradients of
// VDIF_TS).
template <int KIND, int RMAX, typename FT>
__global__ __launch_bounds__(256) void k_cvrecflux(DevGrid g, IntArgs a, const double* fld, const double* mean,
                                                   double cmean) {
  __shared__ CvRecLds<FT> srec[2];
  const int b = xcd_block(), t = threadIdx.x;
  const int k0 = KIND == 0 ? 1 : 2;
  const int nk = g.KBM1 - k0 + 1;   // layers (interfaces) of the pass
  // The role is taken from the wave's first lane, so that the branch is
  // uniform over each wave and both roles pass nk + 1 barriers.
  if (__builtin_amdgcn_readfirstlane(t) < FV_CVB) {
    // records: local node t of the block
    CvRecT<FT>* rec = (CvRecT<FT>*)a.rec;
    const bool nact = t < g.cvb_nn[b];
    const int i = nact ? g.cvb_nodes[(long)FV_CVB * b + t] : 1;
    RingGeo<RMAX> rg;
    ring_geo_load<RMAX>(g, i, rg);
    const double art2 = g.art2[i], nnv = g.nn_hvc[i];
    const bool keep = nact && g.rec_keep[i] != 0;
    // the field's values are loaded a layer ahead (rings of up to 10
    // nodes; the 16-node form has no registers to spare)
    constexpr bool ahead = RMAX <= 10;
    CvLoads<RMAX> L;
    if (ahead) cv_load<KIND, RMAX>(g, a, i, rg, fld, mean, cmean, k0, L);
    for (int s = 0; s <= nk; ++s) {
      const int k = k0 + s;
      if (s < nk) {
        if (nact) {
          if (!ahead) cv_load<KIND, RMAX>(g, a, i, rg, fld, mean, cmean, k, L);
          const CvRecT<FT> r = cv_record_of<KIND, RMAX, FT>(g, a, i, rg, art2, nnv, L, mean, cmean, k);
          srec[s & 1].put(t, r);
          if (keep || (KIND == 0 && k == g.KBM1)) rec[(long)(k - 1) * g.m1 + i] = r;
        }
        if (ahead && s + 1 < nk) cv_load<KIND, RMAX>(g, a, i, rg, fld, mean, cmean, k + 1, L);
      }
      __syncthreads();
    }
  } else {
    // faces: local element t - FV_CVB of the block
    FT* cvflux = (FT*)a.cvflux;
    const int te = t - FV_CVB;
    const bool eact = te < g.cvb_ne[b];
    const int e = eact ? g.cvb_elems[(long)FV_CVB * b + te] : 1;
    const ElemFaces<FT> f = elem_faces<FT>(g, e);
    const int pk = g.cvb_loc[e];
    const int l0 = pk & 255, l1 = (pk >> 8) & 255, l2 = (pk >> 16) & 255;
    const FT fact = (FT)g.fact, rhprnu = (FT)g.hprnu, one = (FT)1.0;
    const double dt1 = a.dt1[e];
    const long cs = (long)g.NCV + 1, p0 = 3L * (e - 1) + 1;
    double um = a.u[e], vm = a.v[e];   // layer 1 (ADV_Q's interface interpolation)
    for (int s = 0; s <= nk; ++s) {
      if (s > 0) {
        const int k = k0 + s - 1;
        const long eo = (long)(k - 1) * g.n1;
        const double uk = a.u[eo + e], vk = a.v[eo + e];
        const double dz1k = DZ1(e, k);
        if (eact) {
          const CvRecLds<FT>& sr = srec[(s - 1) & 1];
          const CvRecT<FT> r[3] = {sr.get(l0), sr.get(l1), sr.get(l2)};
          FT* out = cvflux + (long)(k - 1) * cs + p0;
          if (KIND == 0) {
            const FT dtij = (FT)(dt1 * dz1k);
            cv_faces<FT, false>(f, r, (FT)uk, (FT)vk, dtij, fact, one, out);
          } else {
            const double dz1m = DZ1(e, k - 1), dzz1 = DZZ1(e, k - 1);
            const FT uq = (FT)((uk * dz1m + um * dz1k) / (dz1k + dz1m));
            const FT vq = (FT)((vk * dz1m + vm * dz1k) / (dz1k + dz1m));
            um = uk;
            vm = vk;
            cv_faces<FT, true>(f, r, uq, vq, (FT)(dt1 * dzz1), fact, rhprnu, out);
          }
        }
      }
      __syncthreads();
    }
  }
}

// Phase 3 of ADV_Q (adv_q.F:340-390): ordered sums, vertical flux, update.
template <int SMAX, int RMAX, typename FT>
__global__ __launch_bounds__(256) void k_adv_q_gather(DevGrid g, IntArgs a, const double* q, double* qf, int fct) {
  const int i = (int)xcd_thread();
  if (i > g.M) return;
  if (i == 0) {
    for (int k = 1; k <= g.KB; ++k) qf[NI(g, 0, k)] = 0.0;
    return;
  }
  const int n = g.cv_cnt[i];
  int ce[SMAX];
  static_for<SMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    ce[j] = j < n ? g.cvl_ent[(long)j * g.m1 + i] : 0;
  });
  const double art1 = g.art1[i], dti = g.dti, dtn = a.dt[i], dn = a.d[i];
  const long cs = (long)g.NCV + 1;
  const bool clip = fct && !g.fct_skip[i];
  int rn[RMAX];
  const int nt = clip ? load_ring<RMAX>(g, i, rn) : 0;
  qf[NI(g, i, 1)] = 0.0;
  qf[NI(g, i, g.KB)] = 0.0;
  for (int k = 2; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.m1;
    const FT* fl = (const FT*)a.cvflux + (long)(k - 1) * cs;
    // all SMAX loads issue together (slots past the node's count read the
    // unused entry 0); the sum takes the node's own terms in sub-edge order
    FT xv[SMAX];
    static_for<SMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      xv[j] = fl[ce[j] >> 1];
    });
    double xflux = 0.0;
    static_for<SMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      if (j < n) {
        const double x = (double)xv[j];
        if (ce[j] & 1) xflux = xflux - x;
        else xflux = xflux + x;
      }
    });
    const double temp = a.wts[ko - g.m1 + i] * q[ko - g.m1 + i] - a.wts[ko + g.m1 + i] * q[ko + g.m1 + i];
    xflux = xflux + temp * art1 * DZZ(i, k - 1) / (DZ(i, k - 1) + DZ(i, k));
    const double qk = q[ko + i];
    double x = (qk - xflux / art1 * (dti / (dtn * DZZ(i, k - 1)))) * (dtn / dn);
    if (clip) {  // FCT_Q2 / FCT_Q2L (fct_q2.F:60-95)
      double qmin, qmax;
      ring_bounds<RMAX>(rn, nt, q, ko, qmin, qmax);
      if (k != 2) {
        const double v = (qk * DZZ(i, k - 2) + q[ko - g.m1 + i] * DZZ(i, k - 1)) / (DZZ(i, k - 1) + DZZ(i, k - 2));
        qmax = fmax(qmax, v); qmin = fmin(qmin, v);
      }
      if (k == 2 || k != g.KBM1) {
        const double v = (qk * DZZ(i, k) + q[ko + g.m1 + i] * DZZ(i, k - 1)) / (DZZ(i, k) + DZZ(i, k - 1));
        qmax = fmax(qmax, v); qmin = fmin(qmin, v);
      }
      if (qmin - x > 0.0) x = qmin;
      if (x - qmax > 0.0) x = qmax;
    }
    qf[ko + i] = x;
  }
}

// VERTVL_EDGE (vertvl_edge.F:73-553) by node tile: a workgroup of 256 lanes
// takes a tile of at most FV_NTN = 64 nodes and the at most FV_NTE = 192
// elements around them (nt_*).  Its first three waves form each element's
// three sub-edge terms of layer k, DT1*DZ1*(-U*DLTYE + V*DLTXE) as seen from
// the sub-edge's IB node, into one of two LDS buffers; its last wave sums
// the owned nodes' sub-edges of layer k-1 from the other buffer in sub-edge
// order -- adding a term on side IB, subtracting it on side IA, which is the
// reference's term with the signed DLTXE/DLTYE bit for bit, since negation
// is exact -- then adds the rivers and integrates omega down the column into
// WTS; after the last layer the linear mass-error correction of WTS and
// DTFA (vertvl_edge.F:500-533).  One barrier per layer hands the buffers
// over; the role is read from the wave's first lane, so that the branch is
// uniform over each wave and both roles pass KBM1 + 1 barriers.  Each
// element-layer velocity is loaded once per tile, and the per-layer sums
// read LDS instead of gathering the elements' velocities again for each of
// their nodes.
template <int SMAX>
__global__ __launch_bounds__(256) void k_vertvl_tile(DevGrid g, IntArgs a) {
  __shared__ double sterm[2][3 * FV_NTE];
  const int b = xcd_block(), t = threadIdx.x;
  const int KBM1 = g.KBM1;
  if (b == 0 && t == 0) {
    for (int k = 1; k <= g.KB; ++k) a.wts[NI(g, 0, k)] = 0.0;
    a.dtfa[0] = a.d[0];
  }
  if (__builtin_amdgcn_readfirstlane(t) < FV_NTE) {
    // element terms: local element t of the tile
    const bool eact = t < g.nt_nel[b];
    const int e = eact ? g.nt_elems[(long)FV_NTE * b + t] : 1;
    const double dt1 = a.dt1[e];
    double dx[3], dy[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      dx[j] = g.ecv_dx[(long)j * g.n1 + e];
      dy[j] = g.ecv_dy[(long)j * g.n1 + e];
    }
    // the velocities and DZ1 of layer k are loaded two layers ahead (every
    // index is valid for idle lanes), so that their latency overlaps the
    // barriers of the layers between
    double u0 = a.u[e], v0 = a.v[e], z0 = DZ1(e, 1);
    const long k2 = KBM1 > 1 ? g.n1 : 0;
    double u1 = a.u[k2 + e], v1 = a.v[k2 + e], z1 = DZ1(e, KBM1 > 1 ? 2 : 1);
    for (int s = 0; s <= KBM1; ++s) {
      if (s < KBM1) {
        const double u = u0, v = v0, dz = z0;
        u0 = u1; v0 = v1; z0 = z1;
        if (s + 2 < KBM1) {
          const long kn = (long)(s + 2) * g.n1;
          u1 = a.u[kn + e]; v1 = a.v[kn + e]; z1 = DZ1(e, s + 3);
        }
        if (eact) {
          const double dij = dt1 * dz;
#pragma unroll
          for (int j = 0; j < 3; ++j) sterm[s & 1][3 * t + j] = dij * (-u * dy[j] + v * dx[j]);
        }
      }
      __syncthreads();
    }
  } else {
    // column integration: owned node t - FV_NTE of the tile
    const int tn = t - FV_NTE;
    const bool nact = tn < g.nt_nown[b];
    const int i = nact ? g.nt_nodes[(long)FV_NTN * b + tn] : 1;
    const int n = g.cv_cnt[i];
    int lp[SMAX];
    static_for<SMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      lp[j] = j < n ? g.nt_lpos[(long)j * g.m1 + i] : 0;
    });
    const bool obc = g.isonb[i] == 2;
    const int rv = g.node_riv[i];
    const double art1 = g.art1[i], dd = a.d[i] - a.dt[i];
    double w = 0.0;
    if (nact) a.wts[NI(g, i, 1)] = 0.0;
    for (int s = 0; s <= KBM1; ++s) {
      if (s > 0 && nact) {
        const int k = s;
        const double* st = sterm[(s - 1) & 1];
        double x = 0.0;
        static_for<SMAX>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if (j < n) {
            const double tm = st[lp[j] >> 1];
            x = (lp[j] & 1) ? x + tm : x - tm;
          }
        });
        double xflux = obc ? 0.0 : x;
        if (rv >= 0)
          for (int j = rv; j < g.NUMQBC; ++j) {
            double w;
            if (river_share(g, j, i, w)) xflux = xflux - a.qdis[j] * w * g.vqdist[(long)(k - 1) * g.NUMQBC + j];
          }
        w = w + xflux / art1 + DZ(i, k) * dd / g.dti;
        a.wts[NI(g, i, k + 1)] = w;
      }
      __syncthreads();
    }
    if (nact) {
      const double wbottom = 0.0;
      double dtfa = a.d[i];
      const bool corr = fabs(w - wbottom) > 1.0E-8 && !obc;
      if (corr) {
        double tmp1 = a.elf[i] * (double)KBM1 - (w - wbottom) * g.dti / DZ(i, 1);
        tmp1 = tmp1 / (double)KBM1;
        dtfa = tmp1 + g.h[i];
        // eight levels' loads, then their stores, so that the loads of a
        // batch are in flight together
        const double dw = w - wbottom;
        for (int k0 = 2; k0 <= KBM1 + 1; k0 += 8) {
          double wv[8];
#pragma unroll
          for (int q = 0; q < 8; ++q)
            if (k0 + q <= KBM1 + 1) wv[q] = a.wts[NI(g, i, k0 + q)];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int k = k0 + q;
            if (k <= KBM1 + 1) {
              const double ratio = (double)((float)(k - 1) / (float)KBM1);
              a.wts[NI(g, i, k)] = wv[q] - ratio * dw;
            }
          }
        }
      }
      a.dtfa[i] = dtfa;
    }
  }
}

// VISCOF_H (viscofh.F:33-127): nodal strain-rate magnitude from the NBVE
// element ring, one thread per node with the ring (RMAX >= NTVE) held in
// registers across the layers.  The first and last ring elements carry the
// boundary terms of nodes on the boundary; with NTVE = 1 the single element
// enters twice, as in the reference's three-part loop.
template <int RMAX>
__global__ __launch_bounds__(256) void k_viscofh_r(DevGrid g, IntArgs a) {
  const int i = (int)xcd_thread();
  if (i < 1 || i > g.M) return;
  const int o = g.nbve_off[i], nt = g.ntve[i];
  const bool bnd = g.isonb[i] != 0;
  int re[RMAX];
  double rx[RMAX], ry[RMAX];
  static_for<RMAX>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    re[j] = j < nt ? g.nbve[o + j] : 0;
    rx[j] = j < nt ? g.dltxecec[o + j] : 0.0;
    ry[j] = j < nt ? g.dltyecec[o + j] : 0.0;
  });
  const int jl = nt - 1;
  const int el = g.nbve[o + jl];
  const double rxl = g.dltxecec[o + jl], ryl = g.dltyecec[o + jl];
  const double nx0 = bnd ? g.dltxnec[o] : 0.0, ny0 = bnd ? g.dltynec[o] : 0.0;
  const double nxl = bnd ? g.dltxnec[o + jl] : 0.0, nyl = bnd ? g.dltynec[o + jl] : 0.0;
  const double art1 = g.art1[i];
  for (int k = 1; k <= g.KBM1; ++k) {
    const long ko = (long)(k - 1) * g.n1;
    // the ring's velocities are loaded unconditionally (slots past the ring
    // read element 0) so that the loads issue together
    double ur[RMAX], vr[RMAX];
    static_for<RMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      ur[j] = a.u[ko + re[j]];
      vr[j] = a.v[ko + re[j]];
    });
    double pupx = 0.0, pupy = 0.0, pvpx = 0.0, pvpy = 0.0;
    {
      const double uu = ur[0], vv = vr[0];
      pupx = pupx + uu * ry[0]; pupy = pupy + uu * rx[0];
      pvpx = pvpx + vv * ry[0]; pvpy = pvpy + vv * rx[0];
      if (bnd) {
        pupx = pupx + uu * ny0; pupy = pupy + uu * nx0;
        pvpx = pvpx + vv * ny0; pvpy = pvpy + vv * nx0;
      }
    }
    static_for<RMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      if (j >= 1 && j <= nt - 2) {
        const double uu = ur[j], vv = vr[j];
        pupx = pupx + uu * ry[j]; pupy = pupy + uu * rx[j];
        pvpx = pvpx + vv * ry[j]; pvpy = pvpy + vv * rx[j];
      }
    });
    {
      const double uu = a.u[ko + el], vv = a.v[ko + el];
      pupx = pupx + uu * ryl; pupy = pupy + uu * rxl;
      pvpx = pvpx + vv * ryl; pvpy = pvpy + vv * rxl;
      if (bnd) {
        pupx = pupx + uu * (-nyl); pupy = pupy + uu * (-nxl);
        pvpx = pvpx + vv * (-nyl); pvpy = pvpy + vv * (-nxl);
      }
    }
    pupx = pupx / art1; pupy = pupy / art1; pvpx = pvpx / art1; pvpy = pvpy / art1;
    const double tmp1 = sq(pupx) + sq(pvpy);
    const double tmp2 = 0.5 * sq(pupy + pvpx);
    a.viscofh[NI(g, i, k)] = sqrt(tmp1 + tmp2) * art1;
  }
}
