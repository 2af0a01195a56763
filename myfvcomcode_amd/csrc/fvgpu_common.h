// fvgpu_common.h -- device-side layout shared by the HIP kernels.
//
// All arrays keep the reference's dummy entry 0 and its column-major 3-D
// layout A(0:MT,1:KB): entity index fastest, so a wavefront of consecutive
// nodes/elements at one layer reads consecutive addresses.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#define FV_MAXKB 64
// VDIF_Q keeps the elimination coefficients of its two solves in registers
// (1) or in a node-layer scratch pair (0: unrolled level loops, 2: rolled),
// trading VGPRs for occupancy.
#ifndef FV_VDIFQ_VHREG
#define FV_VDIFQ_VHREG 0
#endif
#define FV_MAX_CV 32   // control-volume sub-edges per node (twice the node degree)
#define FV_MAX_RING 16 // NTSN / NTVE bound of the node rings held in registers
#define FV_CVB 128     // elements and nodes per block of the fused CV kernels
#define FV_NTN 64      // owned nodes per node tile
#define FV_NTE 192     // elements per node tile

// Element blocks of the blocked ADVECTION / ADV_UV kernels: block b owns
// nown[b] (<= 128) elements, its records 0 .. nown[b]-1, followed by the
// other elements of their edges: ns[b] (<= 256) records, element
// elems[256*b + t]; its ne[b] (<= 256) edges edges[256*b + t], with the
// record slots of IA and IB in bits 0-7 and 8-15 of eloc[256*b + t];
// epos[e] packs the edge slots of element e's three edges (in ed_ent
// order) in bits 0-7, 8-15, 16-23.
struct UvBlocks {
  int nblk;
  const int *nown, *ns, *elems, *ne, *edges, *eloc, *epos;
};

struct DevGrid {
  int M, N, NE, NCV, KB, KBM1, KBM2, IOBCN, NUMQBC, NTIDE;
  long m1, n1;
  int sig_uniform;          // all node (and all element) sigma columns identical
  double dte, dti, umol, vprnu, hprnu, cbcmin, min_depth;
  double fact, fm1;         // horizontal mixing closure / constant switches
  double const1, cvhp1;     // 16.6**.6666667*SEF and (15.8*CBCNST)**0.6666667
  int isplit, iramp, noflux_bot, river_ts_specified, type_tsobc, obc_on, obc_elev, wind_on;
  // nodes
  const double *vx, *vy, *h, *art1, *art2, *grav_n, *nn_hvc, *cos_sita, *sin_sita, *phpn;
  const int *isonb, *ntsn, *nbsn_off, *nbsn, *ntve, *nbve_off, *nbve;
  const double *dltxtrie, *dltytrie, *dltxecec, *dltyecec, *dltxnec, *dltynec;
  // elements
  const double *xc, *yc, *art, *h1, *cor, *cc_hvc, *cc_sponge, *cc_z0b, *grav_e, *epor;
  const double *sin_alpha, *cos_alpha, *wsx, *wsy;
  const int *isbce, *iucp, *nv, *nbe;
  const double *a1u, *a2u, *awx, *awy;
  const double* a12;               // A1U(e,1..4), A2U(e,1..4) of element e at 8e..8e+7 (64-byte aligned)
  // edges and CV sub-edges
  const int *iec, *ienode, *isbc, *niec, *ntrg;
  const double *xijc, *yijc, *dltxc, *dltyc, *dltxe, *dltye, *dltxncve, *dltyncve;
  // sigma (per column, or one column when sig_uniform)
  const double *z, *zz, *dz, *dzz, *z1, *zz1, *dz1, *dzz1;
  // gather lists
  const int *cv_off, *cv_ent;      // node -> (cv<<1 | side), sorted by cv
  // the same lists slot-major (entry j of node i at j*m1+i, j < cv_cnt[i]):
  // consecutive nodes read consecutive addresses for each slot.  cvl_ent
  // holds (p<<1 | side), p the element-major position of the sub-edge's
  // flux in the per-layer flux buffers (below).
  int cv_max;
  int ring_max, ve_max;            // max NTSN, max NTVE
  const int *cv_cnt, *cvl_ent;
  // element -> its 3 sub-edges, slot-major (slot j of element e at
  // j*n1+e), slot j being the sub-edge on the element edge from NV(e,j+1)
  // to NV(e,(j+1)%3+1): the sub-edge index, and 1 if its IA is the second
  // of those nodes.  The flux of slot j of element e is stored at position
  // 3*(e-1)+j+1.
  const int *ecv_c, *ecv_swap;
  const double *ecv_dx, *ecv_dy;   // DLTXE/DLTYE of slot j's sub-edge, same layout
  const int *ed_ent;               // element -> 3 x (edge<<1 | side), sorted by edge
  const int *fo_off, *fo_obc;      // element -> FLUXOBN contributions (obc<<1 | slot)
  const int *node_obc;             // node -> OBC index or -1
  const int *node_riv;             // node -> river index or -1
  const int *re_off, *re_riv;      // element -> rivers whose NBVE(1) / NBVE(NTVE) it is
  const int *fct_skip;             // node skipped by FCT (OBC or river node)
  // Element blocks of the fused record + face-flux kernels: block b holds
  // the cvb_ne[b] (<= FV_CVB) elements cvb_elems[FV_CVB*b + t] and the
  // cvb_nn[b] (<= FV_CVB) distinct nodes they use, cvb_nodes[FV_CVB*b + t];
  // cvb_loc[e] packs the local indices of NV(e,1..3) in bits 0-7, 8-15,
  // 16-23.  rec_keep[i] = 1 for the nodes whose records other kernels read
  // at every layer (the OBC nodes and the other ends of their sub-edges).
  int cvb_nblk;
  const int *cvb_nn, *cvb_nodes, *cvb_ne, *cvb_elems, *cvb_loc, *rec_keep;
  // Node tiles (VERTVL): tile b owns the nt_nown[b] (<= FV_NTN) nodes
  // nt_nodes[FV_NTN*b + t]; nt_elems[FV_NTE*b + t] are the nt_nel[b]
  // (<= FV_NTE) elements that have one of them as a vertex.  Slot j of
  // owned node i: nt_lpos[j*m1 + i] = (3*local element + element slot) << 1
  // | side, in the order of cvl_ent.
  int nt_nblk;
  const int *nt_nown, *nt_nodes, *nt_nel, *nt_elems, *nt_lpos;
  // Element blocks of the blocked ADVECTION / ADV_UV kernels (UvBlocks):
  // compact tiles for ADVECTION, runs of consecutive elements for ADV_UV.
  UvBlocks ub_tile, ub_run;
  // OBC / tide / rivers
  const int *i_obc_n, *next_obc, *nadjc_obc, *adjc_obc, *type_obc;
  const int *next_obc2;
  int obc_radiate;                 // some OBC node has TYPE_OBC 5..10
  int obc_ore;                     // some OBC node has TYPE_OBC 9/10 (Orlanski)
  const double *fluxf_obc, *dltn_obc, *emean, *apt, *phai, *period;
  const int *inodeq;               // node river: its node, else 0
  const int *icellq, *n_icellq;    // edge river: its cell and the two nodes of its solid edge, else 0
  const double *rdisq;             // edge river: the two nodes' shares of its discharge
  const double *qarea, *cos_angleq, *sin_angleq, *qdis0, *tdis, *sdis, *vqdist;
};

// Workgroup index in XCD-contiguous order.  The dispatcher deals consecutive
// workgroups round-robin to the 8 XCDs, each with its own L2; this remap gives
// XCD x one contiguous range of indices, so workgroups over neighbouring mesh
// entities (which gather the same node/element data) share an L2.  It is a
// bijection on [0, gridDim.x).
__device__ __forceinline__ int xcd_block() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int per = nb >> 3, rem = nb & 7, x = b & 7;
  return x * per + (x < rem ? x : rem) + (b >> 3);
}
__device__ __forceinline__ long xcd_thread() { return (long)xcd_block() * blockDim.x + threadIdx.x; }

// 1-based layer accessors
__device__ __forceinline__ long NI(const DevGrid& g, int i, int k) { return (long)(k - 1) * g.m1 + i; }
__device__ __forceinline__ long EI(const DevGrid& g, int i, int k) { return (long)(k - 1) * g.n1 + i; }
__device__ __forceinline__ double SIGN_(const DevGrid& g, const double* a, int i, int k) {
  return g.sig_uniform ? a[k - 1] : a[NI(g, i, k)];
}
__device__ __forceinline__ double SIGE_(const DevGrid& g, const double* a, int i, int k) {
  return g.sig_uniform ? a[k - 1] : a[EI(g, i, k)];
}
#define DZ(i, k) SIGN_(g, g.dz, i, k)
#define DZZ(i, k) SIGN_(g, g.dzz, i, k)
#define ZN(i, k) SIGN_(g, g.z, i, k)
#define ZZN(i, k) SIGN_(g, g.zz, i, k)
#define DZ1(i, k) SIGE_(g, g.dz1, i, k)
#define DZZ1(i, k) SIGE_(g, g.dzz1, i, k)
#define Z1(i, k) SIGE_(g, g.z1, i, k)
#define ZZ1(i, k) SIGE_(g, g.zz1, i, k)
#define NV_(i, j) g.nv[(long)((j)-1) * g.n1 + (i)]
#define NBE_(i, j) g.nbe[(long)((j)-1) * g.n1 + (i)]
#define A1U_(i, j) g.a1u[(long)((j)-1) * g.n1 + (i)]
#define A2U_(i, j) g.a2u[(long)((j)-1) * g.n1 + (i)]
#define AWX_(i, j) g.awx[(long)((j)-1) * g.n1 + (i)]
#define AWY_(i, j) g.awy[(long)((j)-1) * g.n1 + (i)]
#define IEC_(e, j) g.iec[(long)((j)-1) * (g.NE + 1) + (e)]
#define IENODE_(e, j) g.ienode[(long)((j)-1) * (g.NE + 1) + (e)]
#define NIEC_(c, j) g.niec[(long)((j)-1) * (g.NCV + 1) + (c)]
#define DLTXNCVE_(c, j) g.dltxncve[(long)((j)-1) * (g.NCV + 1) + (c)]
#define DLTYNCVE_(c, j) g.dltyncve[(long)((j)-1) * (g.NCV + 1) + (c)]

// The least-squares coefficients A1U(e,1..4) and A2U(e,1..4) of element
// e, in the flux type, from the interleaved copy (two 32-byte loads each).
template <typename FT>
__device__ __forceinline__ void lsq_coef(const DevGrid& g, int e, FT (&a1)[4], FT (&a2)[4]) {
  const double4* p = reinterpret_cast<const double4*>(g.a12) + 2L * e;
  const double4 u = p[0], w = p[1];
  a1[0] = (FT)u.x; a1[1] = (FT)u.y; a1[2] = (FT)u.z; a1[3] = (FT)u.w;
  a2[0] = (FT)w.x; a2[1] = (FT)w.y; a2[2] = (FT)w.z; a2[3] = (FT)w.w;
}

__device__ __forceinline__ double fsign1(double b) { return b >= 0.0 ? 1.0 : -1.0; }
__device__ __forceinline__ float fsign1(float b) { return b >= 0.0f ? 1.0f : -1.0f; }
__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ float sq(float x) { return x * x; }

// Whether river j's discharge enters node i, and its share there: all of
// it at a node river's node (INODEQ), RDISQ(J,1) / RDISQ(J,2) at the two
// nodes of an edge river's solid edge (N_ICELLQ).  A node river's share is
// exactly 1, so QDIS*share is QDIS bit for bit.
__device__ __forceinline__ bool river_share(const DevGrid& g, int j, int i, double& w) {
  if (g.inodeq[j] == i) { w = 1.0; return true; }
  if (g.icellq[j] > 0) {
    if (g.n_icellq[j] == i) { w = g.rdisq[j]; return true; }
    if (g.n_icellq[g.NUMQBC + j] == i) { w = g.rdisq[g.NUMQBC + j]; return true; }
  }
  return false;
}

// Device clock: the time-dependent scalars of the current step, written by
// tiny kernels so that a captured graph replays correctly.
struct DevClock {
  long long iint;
  int iext;
  double ramp_int;      // RAMP of INTERNAL_STEP
  double ramp_ext;      // RAMP of the current EXTERNAL_STEP
  double ext_time;      // ExtTime (s since SpecTime)
  double time0, dti, dte;
  int isplit, iramp;
  int depth_fail;
  int pad;
};

// Compile-time loop: f(std::integral_constant<int, K>) for K = 0..N-1, fully
// expanded so that register arrays are indexed by constants.
template <typename F, int... Ks>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Ks...>) {
  (f(std::integral_constant<int, Ks>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Host-side choice of the register-array bound KMAX >= KBM1 for the
// column kernels templated on it.
#define FV_KMAX_DISPATCH(KBM1, CALL)                 \
  do {                                               \
    if ((KBM1) <= 8) { CALL(8); }                    \
    else if ((KBM1) <= 16) { CALL(16); }             \
    else if ((KBM1) <= 24) { CALL(24); }             \
    else if ((KBM1) <= 32) { CALL(32); }             \
    else if ((KBM1) <= 40) { CALL(40); }             \
    else if ((KBM1) <= 48) { CALL(48); }             \
    else { CALL(64); }                               \
  } while (0)

// Host-side choice of the per-node sub-edge bound SMAX >= cv_max.
#define FV_SMAX_DISPATCH(SMAX_, CALL)                \
  do {                                               \
    if ((SMAX_) <= 16) { CALL(16); }                 \
    else if ((SMAX_) <= 24) { CALL(24); }            \
    else { CALL(32); }                               \
  } while (0)

// Host-side choice of the node-ring bound RMAX >= ring_max (NTSN or NTVE).
#define FV_RMAX_DISPATCH(RM_, CALL)                  \
  do {                                               \
    if ((RM_) <= 10) { CALL(10); }                   \
    else { CALL(16); }                               \
  } while (0)
