// fvgpu.hip -- MI355X (gfx950) implementation of FVCOM's mode-split time
// step behind the C-ABI of include/fvgpu.h.
//
// fvgpu_internal_step replaces INTERNAL_STEP (FVCOM_source/internal_step.F:29)
// and runs ISPLIT external steps (EXTERNAL_STEP, external_step.F:30) inside
// it.  The state stays resident in HBM between calls; the host only
// uploads at fvgpu_create and copies arrays on request.
//
// Design points:
//  * every scatter loop of the reference becomes a gather through sorted
//    CSR lists (node -> control-volume sub-edges, element -> edges), so no
//    atomics are used and sums are formed in the reference's order;
//  * the RK stages of the 2-D mode ping-pong between three buffer sets with
//    the fourth stage written in place over the RK base (pointwise reads
//    only), so the state returns to fixed buffers after every external step
//    and one internal step can be captured as a replayable hipGraph;
//  * time-dependent scalars (RAMP, ExtTime) live in a device clock updated
//    by single-thread kernels, never by host writes between launches.
#include "fvgpu_closure.h"
#include "../../include/fvgpu.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include <algorithm>
#include <utility>

namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return 1; }

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct KernelStat { double total_ms = 0; int64_t count = 0; };

}  // namespace

// Staging of fvgpu_ensemble_update_noalloc: X5 (FV_ENS_MAX x FV_ENS_MAX)
// and the member offsets (FV_ENS_MAX), 33,280 bytes held on every GPU the
// library runs on until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: host code uses only the device
// address of the current GPU's copy, looked up once per device by
// fvgpu_create (scratch_of).  Every call on one GPU shares its copy, so
// overlapping calls on one GPU (from different handles' streams or host
// threads) race.
static __device__ __align__(256) double scratch[FV_ENS_MAX * FV_ENS_MAX + FV_ENS_MAX];
#define FV_SCRATCH_DEVICES 64

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[FV_SCRATCH_DEVICES] = {};

// Device address of the current GPU's `scratch` (looked up on first use
// for that device, which must be current).  Null with the error set on
// failure.
double* scratch_of(int dev) {
  if (dev < 0 || dev >= FV_SCRATCH_DEVICES) { fail("device index beyond the staging table"); return nullptr; }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace


struct fvgpu_handle {
  const fv_model* md = nullptr;
  int precision = 0;
  int device = 0;
  int cu_count = 0;       // compute units of the device
  hipStream_t stream = nullptr;
  // Independent chains of the step (the 2-D elevation and edge terms, the
  // T and S passes, the Q2 and Q2L advection) fork onto stream2 and join
  // back; both streams have the default priority.  `cur` is the stream the
  // launch macros enqueue on.
  hipStream_t stream2 = nullptr, cur = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int concurrent = 3;     // bit 0: 3-D chain pairs, bit 1: the 2-D stage pair on stream2
  // CV records then face fluxes in two launches (1) or one fused launch
  // (0), for ADV_T/ADV_S and for ADV_Q
  int two_phase_ts = 0, two_phase_q = 0;
  // wind speeds at the start (wu0, wv0) and end (wu1, wv1) of the coming
  // internal step; once they differ, every external step forms its own
  // stress (ws2x, ws2y), otherwise it uses the end stress g.wsx/g.wsy
  double *wu0 = nullptr, *wv0 = nullptr, *wu1 = nullptr, *wv1 = nullptr, *ws2x = nullptr, *ws2y = nullptr;
  int wind_varies = 0;
  DevGrid g{};
  std::vector<void*> allocs;
  // device state
  DevClock* clk = nullptr;
  Ext2DState ext[3];      // ext[0] = state / RK base, ext[1], ext[2] stage buffers
  IntArgs ia{};
  IntArgs ib{};           // ia with its own scratch, for the chain on stream2
  double *xflux_obcn = nullptr, *fobn = nullptr;
  double* phi = nullptr;  // per sub-edge EXTEL term of the 2-D stage input, element-major
  double* ens_buf = nullptr;  // X5 and member offsets of fvgpu_ensemble_update
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // profiling
  bool profiling = false;
  std::string profile_only;    // empty: every kernel family
  int profile_every = 1;       // time one in this many launches of the timed families
  long profile_seen = 0;       // launches of the timed families since profiling began
  std::map<std::string, KernelStat> stats;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;
  size_t event_next = 0;
  void* diag_part = nullptr;   // 256 partial sums of the diagnostics reduction
  // learned closure (mode 1), allocated by fvgpu_set_closure
  int closure_mode = 0;
  float* cl_x = nullptr;       // features, M*(KBM1-1) x 8
  float* cl_yp = nullptr;      // predicted (KM, KH)
  MlpPacked* cl_w = nullptr;   // BF16 fragments
  int* cl_flag = nullptr;      // per node: 1 = MY2.5 fallback
  int* cl_nfall = nullptr;     // fallback columns of the last step
  std::map<std::string, int> tune;   // workgroup size per kernel family (autotuner)
  std::vector<std::pair<void*, size_t>> rw;   // state buffers (saved around autotuning)
};

namespace {

template <typename T>
int dalloc(fvgpu_handle* h, T** p, size_t n) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
  HIPCHK(hipMemset(q, 0, (n ? n : 1) * sizeof(T)));
  h->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

template <typename T>
int dupload(fvgpu_handle* h, const T** p, const T* src, size_t n) {
  T* q = nullptr;
  if (dalloc(h, &q, n)) return 1;
  if (n) HIPCHK(hipMemcpy(q, src, n * sizeof(T), hipMemcpyHostToDevice));
  *p = q;
  return 0;
}

int dupload_rw(fvgpu_handle* h, double** p, const double* src, size_t n) {
  if (dalloc(h, p, n)) return 1;
  h->rw.push_back({(void*)*p, (n ? n : 1) * sizeof(double)});
  if (n && src) HIPCHK(hipMemcpy(*p, src, n * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

// profiling helpers: events around launches when enabled (eager mode only).
// The enqueue helpers below leave their HIP status to the runtime's last
// error, which the entry points check with hipGetLastError after enqueuing.
hipEvent_t next_event(fvgpu_handle* h) {
  if (h->event_next >= h->event_pool.size()) {
    hipEvent_t e;
    (void)hipEventCreate(&e);
    h->event_pool.push_back(e);
  }
  return h->event_pool[h->event_next++];
}

struct Launch {
  fvgpu_handle* h;
  const char* name;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool on;
  Launch(fvgpu_handle* h_, const char* n) : h(h_), name(n) {
    on = h->profiling && (h->profile_only.empty() || h->profile_only == n);
    if (on && h->profile_every > 1) on = h->profile_seen++ % h->profile_every == 0;
    if (on) { e0 = next_event(h); (void)hipEventRecord(e0, h->cur); }
  }
  ~Launch() {
    if (on) {
      e1 = next_event(h);
      (void)hipEventRecord(e1, h->cur);
      h->pending.push_back({name, {e0, e1}});
    }
  }
};

int flush_profile(fvgpu_handle* h) {
  if (h->pending.empty()) return 0;
  HIPCHK(hipStreamSynchronize(h->stream));
  for (auto& p : h->pending) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    KernelStat& s = h->stats[p.first];
    s.total_ms += ms;
    s.count += 1;
  }
  h->pending.clear();
  h->event_next = 0;
  return 0;
}

// The captured steps are stale (the step's launches or their options
// changed): the next graph run captures again.
void drop_graphs(fvgpu_handle* h) {
  if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
  if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
}

#define LAUNCH(NAME, KERNEL, GRID, BLOCK, ...)                                   \
  do {                                                                            \
    Launch l_(h, NAME);                                                           \
    hipLaunchKernelGGL(KERNEL, dim3(GRID), dim3(BLOCK), 0, h->cur, __VA_ARGS__);    \
  } while (0)

inline unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

// N2E3D of a node field over layers 1..KB (k_n2e3d's grid: elements by
// groups of eight layers, x extent a multiple of 8).
#define N2E3D(NVAR, EVAR)                                                                        \
  LAUNCH("n2e3d", k_n2e3d, dim3((nb(g.N) + 7) / 8 * 8, (unsigned)((g.KB + 7) / 8)), 256, g, NVAR, EVAR, \
         g.KB)

// Launch of a 1-D kernel family over n work items with the family's
// workgroup size (256 unless the autotuner chose 64 or 128).
int block_of(const fvgpu_handle* h, const char* name) {
  auto it = h->tune.find(name);
  return it == h->tune.end() ? 256 : it->second;
}
#define TLAUNCH(NAME, KERNEL, NITEMS, ...)                                                         \
  do {                                                                                            \
    const int bs_ = block_of(h, NAME);                                                            \
    Launch l_(h, NAME);                                                                           \
    hipLaunchKernelGGL(KERNEL, dim3((unsigned)(((long)(NITEMS) + bs_ - 1) / bs_)), dim3(bs_), 0, \
                       h->cur, __VA_ARGS__);                                                      \
  } while (0)

// Fork: work enqueued on stream2 after this waits for everything enqueued
// on the main stream so far.  Join: the main stream waits for stream2.
// Under graph capture the events become dependency edges of the graph.
// `what` is the concurrency bit of the pair (FV_CONC_3D or FV_CONC_2D);
// with that bit clear both chains go to the main stream in turn.
enum { FV_CONC_3D = 1, FV_CONC_2D = 2 };
void fork_streams(fvgpu_handle* h, int what) {
  if (!(h->concurrent & what)) return;
  (void)hipEventRecord(h->ev_fork, h->stream);
  (void)hipStreamWaitEvent(h->stream2, h->ev_fork, 0);
}
void join_streams(fvgpu_handle* h, int what) {
  if (!(h->concurrent & what)) return;
  (void)hipEventRecord(h->ev_join, h->stream2);
  (void)hipStreamWaitEvent(h->stream, h->ev_join, 0);
}
// enqueue on the second chain's stream / back on the main stream
void on_stream2(fvgpu_handle* h, int what) { h->cur = (h->concurrent & what) ? h->stream2 : h->stream; }
void on_stream1(fvgpu_handle* h) { h->cur = h->stream; }

const double ALPHA_RK[5] = {0.0, 0.2500, 1.0 / 3.0, 0.5000, 1.0};

// ------------------------------------------------------------- the steps
ExtArgs make_ext_args(fvgpu_handle* h, int K, const Ext2DState& cur, const Ext2DState& nxt) {
  ExtArgs a{};
  a.cur = cur;
  a.base = h->ext[0];
  a.nxt = nxt;
  a.K = K;
  a.alpha = ALPHA_RK[K];
  a.xflux_obcn = h->xflux_obcn;
  a.fobn = h->fobn;
  a.uard_obcn = h->ia.uard_obcn;
  a.egf = h->ia.egf;
  a.fluxobn = h->ia.fluxobn;
  a.uard = h->ia.uard;
  a.vard = h->ia.vard;
  a.adx2d = h->ia.adx2d; a.ady2d = h->ia.ady2d; a.drx2d = h->ia.drx2d; a.dry2d = h->ia.dry2d;
  a.wubot = h->ia.wubot; a.wvbot = h->ia.wvbot;
  a.qdis = h->ia.qdis;
  a.ef = h->ia.ef2;
  a.elij = h->ia.elij2;
  a.phi = h->phi;
  a.ws2x = h->wind_varies ? h->ws2x : h->g.wsx;
  a.ws2y = h->wind_varies ? h->ws2y : h->g.wsy;
  a.clk = h->clk;
  a.elm1 = h->ia.elm1;
  a.elm2 = h->ia.elm2;
  return a;
}

// `fresh`: the 2-D state may have changed since the last stage (first
// external step of an internal step, or a direct EXTERNAL_STEP call), so
// the EXTEL terms are formed from it first; otherwise the last stage's
// k_ext_momentum left them.
template <typename FT>
void enqueue_external_step_t(fvgpu_handle* h, int iext, bool fresh) {
  const DevGrid& g = h->g;
  LAUNCH("ext_clock", k_ext_clock, 1, 64, h->clk, iext);
  if (fresh) TLAUNCH("ext_phi", k_ext_phi, g.N, g, h->ext[0].ua, h->ext[0].va, h->ext[0].d1, h->phi);
  if (h->wind_varies)
    TLAUNCH("ext_wind", k_ext_wind, g.N, g, h->wu0, h->wv0, h->wu1, h->wv1, iext, h->ws2x, h->ws2y);
  // stage buffers: 1: A->B, 2: B->C, 3: C->B, 4: B->A (in place on the RK base)
  const Ext2DState* cur[5] = {nullptr, &h->ext[0], &h->ext[1], &h->ext[2], &h->ext[1]};
  const Ext2DState* nxt[5] = {nullptr, &h->ext[1], &h->ext[2], &h->ext[1], &h->ext[0]};
  for (int K = 1; K <= 4; ++K) {
    ExtArgs a = make_ext_args(h, K, *cur[K], *nxt[K]);
    // the edge terms and the elevation both read only the stage input
    fork_streams(h, FV_CONC_2D);
    on_stream2(h, FV_CONC_2D);
    TLAUNCH("advave_edge", k_advave_edge<FT>, g.NE + 1, g, a.cur.ua, a.cur.va, a.cur.d, a.cur.el, a.ef, a.elij);
    on_stream1(h);
#define EXTEL_CALL(SM) TLAUNCH("ext_elevation", k_ext_elevation<SM>, g.M + 1, g, a)
    FV_SMAX_DISPATCH(g.cv_max, EXTEL_CALL);
    if (g.obc_radiate) LAUNCH("ext_obc_radiate", k_ext_obc_radiate, nb(g.IOBCN), 256, g, a);
    if (g.obc_ore) LAUNCH("elm_shift", k_elm_shift, nb(g.M + 1), 256, g, a);
    join_streams(h, FV_CONC_2D);
    TLAUNCH("ext_momentum", k_ext_momentum<FT>, g.N + 1, g, a);
  }
}

// ADV_T or ADV_S with FCT, VDIF_TS, BCOND_TS, F1 = FF and N2E3D of one
// tracer (internal_step.F), on h->cur with the scratch of `a`.
template <typename FT>
void enqueue_ts_pass(fvgpu_handle* h, IntArgs& a, int which, int fct) {
  const DevGrid& g = h->g;
  const long mk = g.m1 * g.KB;
  double* f1 = which == 1 ? a.t1 : a.s1;
  double* ff = which == 1 ? a.tf1 : a.sf1;
  const double* fm = which == 1 ? a.tmean1 : a.smean1;
  const double* fdis = which == 1 ? g.tdis : g.sdis;
  if (h->two_phase_ts) {
#define CVREC_TS_CALL(RM) TLAUNCH("cvrec_ts", (k_cvrec<0, RM, FT>), g.M + 1, g, a, f1, fm, 0.0, 1, g.KBM1)
    FV_RMAX_DISPATCH(g.ring_max, CVREC_TS_CALL);
    TLAUNCH("cvflux_ts", k_cvflux_ts<FT>, g.N + 1, g, a);
  } else {
#define CVRF_TS_CALL(RM) LAUNCH("cvrecflux_ts", (k_cvrecflux<0, RM, FT>), g.cvb_nblk, 256, g, a, f1, fm, 0.0)
    FV_RMAX_DISPATCH(g.ring_max, CVRF_TS_CALL);
  }
#define ADV_TS_CALL10(SM) TLAUNCH("adv_ts", (k_adv_ts_gather<SM, 10, FT>), g.M + 1, g, a, f1, ff, fdis, which, fct)
#define ADV_TS_CALL16(SM) TLAUNCH("adv_ts", (k_adv_ts_gather<SM, 16, FT>), g.M + 1, g, a, f1, ff, fdis, which, fct)
  if (g.ring_max <= 10) FV_SMAX_DISPATCH(g.cv_max, ADV_TS_CALL10);
  else FV_SMAX_DISPATCH(g.cv_max, ADV_TS_CALL16);
  if (g.IOBCN > 0) {
    const long np = (long)g.IOBCN * g.KBM1;
    if (g.cv_max <= 16) LAUNCH("adv_ts_obc", (k_adv_ts_obc<16, FT>), (unsigned)((np + 15) / 16), 256, g, a);
    else LAUNCH("adv_ts_obc", (k_adv_ts_obc<32, FT>), (unsigned)((np + 7) / 8), 256, g, a);
  }
#define VDIF_TS_CALL(KM) TLAUNCH("vdif_ts", k_vdif_ts_r<KM>, g.M + 1, g, a, ff, which)
  FV_KMAX_DISPATCH(g.KBM1, VDIF_TS_CALL);
  TLAUNCH("bcond_ts", k_bcond_ts_sums, g.IOBCN + 1, g, a, which, a.obcsum);
  if (g.IOBCN > 0) TLAUNCH("bcond_ts", k_bcond_ts_layers, (long)g.IOBCN * g.KBM1, g, a, which, (const double*)a.obcsum);
  TLAUNCH("copy", k_copy, mk, f1, ff, mk);
  N2E3D(f1, which == 1 ? a.t : a.s);
}

// ADV_Q(Q, QF) with FCT_Q2 (adv_q.F, fct_q2.F) on h->cur with the scratch
// of `a`.
template <typename FT>
void enqueue_adv_q(fvgpu_handle* h, IntArgs& a, double* q, double* qf, int fct) {
  const DevGrid& g = h->g;
  const double qmean1 = (double)1.E-8f;
  if (h->two_phase_q) {
#define CVREC_Q_CALL(RM) TLAUNCH("cvrec_q", (k_cvrec<1, RM, FT>), g.M + 1, g, a, q, nullptr, qmean1, 2, g.KBM1)
    FV_RMAX_DISPATCH(g.ring_max, CVREC_Q_CALL);
    TLAUNCH("cvflux_q", k_cvflux_q<FT>, g.N + 1, g, a);
  } else {
#define CVRF_Q_CALL(RM) LAUNCH("cvrecflux_q", (k_cvrecflux<1, RM, FT>), g.cvb_nblk, 256, g, a, q, nullptr, qmean1)
    FV_RMAX_DISPATCH(g.ring_max, CVRF_Q_CALL);
  }
#define ADV_Q_CALL10(SM) TLAUNCH("adv_q", (k_adv_q_gather<SM, 10, FT>), g.M + 1, g, a, q, qf, fct)
#define ADV_Q_CALL16(SM) TLAUNCH("adv_q", (k_adv_q_gather<SM, 16, FT>), g.M + 1, g, a, q, qf, fct)
  if (g.ring_max <= 10) FV_SMAX_DISPATCH(g.cv_max, ADV_Q_CALL10);
  else FV_SMAX_DISPATCH(g.cv_max, ADV_Q_CALL16);
}

template <typename FT>
void enqueue_internal_step_t(fvgpu_handle* h) {
  const DevGrid& g = h->g;
  IntArgs& a = h->ia;
  const long mk = g.m1 * g.KB, nk = g.n1 * g.KB;
  LAUNCH("int_clock", k_int_clock, 1, 64, g, h->clk, a.qdis);
#define ADJ1_CALL(KM) TLAUNCH("adjust1_brough", k_adjust1_brough<KM>, g.N + 1, g, a)
  FV_KMAX_DISPATCH(g.KBM1, ADJ1_CALL);
  LAUNCH("adv3d_blk", (k_uv_blk<0, FT>), g.ub_tile.nblk, 256, g, a);
  TLAUNCH("baropg", k_baropg, g.N + 1, g, a, (int)h->md->p.barotropic);
  if (!h->md->p.barotropic) TLAUNCH("rho1_roundtrip", k_rho1_roundtrip, (long)g.KBM1 * g.m1, g, a);
  TLAUNCH("advave_edge", k_advave_edge<FT>, g.NE + 1, g, a.ua, a.va, a.d, a.el, a.ef2, a.elij2);
  TLAUNCH("advave_int", k_advave_int<FT>, g.N + 1, g, a);
  TLAUNCH("zero", k_zero, g.m1, a.egf, g.m1);
  if (g.IOBCN > 0) TLAUNCH("zero", k_zero, g.IOBCN, a.uard_obcn, (long)g.IOBCN);
  for (int iext = 1; iext <= g.isplit; ++iext) enqueue_external_step_t<FT>(h, iext, iext == 1);
#define ADJ2_CALL(KM) TLAUNCH("adjust2", k_adjust2<KM>, g.N + 1, g, a)
  FV_KMAX_DISPATCH(g.KBM1, ADJ2_CALL);
#define VERTVL_CALL(SM) LAUNCH("vertvl", k_vertvl_tile<SM>, g.nt_nblk, 256, g, a)
  FV_SMAX_DISPATCH(g.cv_max, VERTVL_CALL);
  N2E3D(a.wts, a.w);
#define VISCOFH_CALL(RM) TLAUNCH("viscofh", k_viscofh_r<RM>, g.M + 1, g, a)
  FV_RMAX_DISPATCH(g.ve_max, VISCOFH_CALL);
  LAUNCH("advuv_blk", (k_uv_blk<1, FT>), g.ub_run.nblk, 256, g, a);
#define VDIF_UV_CALL(KM) LAUNCH("vdif_uv", k_vdif_uv_r<KM>, dim3(nb(g.N + 1), 2), 256, g, a)
  FV_KMAX_DISPATCH(g.KBM1, VDIF_UV_CALL);
  TLAUNCH("uv_finish", k_uv_finish, (long)g.N * g.KBM1, g, a);
  LAUNCH("wreal", k_wreal, nb(g.N), 256, g, a);
  const int fct = h->md->p.scalar_positivity ? 1 : 0;
  if (h->md->p.vertical_closure && h->closure_mode == 1) {
    // learned closure: features, MFMA inference, realizability, then the
    // MY2.5 column solve (without horizontal advection of Q2/Q2L) for the
    // rejected columns only
    const long ns = (long)g.M * (g.KBM1 - 1);
#define CLF_CALL(RM) TLAUNCH("closure_features", k_closure_features<RM>, g.M, g, a, h->cl_x, (float*)nullptr)
    FV_RMAX_DISPATCH(g.ve_max, CLF_CALL);
    LAUNCH("closure_infer", k_mlp_infer_mfma, (unsigned)std::min<long>((ns + 31) / 32, 8192), 64, h->cl_w, h->cl_x, ns,
           h->cl_yp);
    LAUNCH("zero", k_zero_int, 1, 64, h->cl_nfall, 1);
    TLAUNCH("closure_apply", k_closure_apply, g.M, g, a, h->cl_yp, a.gx, h->cl_flag, h->cl_nfall);
    TLAUNCH("fallback_q2f", k_fallback_q2f, g.M + 1, g, a, (const int*)h->cl_flag);
#define VDIF_QM_CALL(KM) TLAUNCH("vdif_q", (k_vdif_q_r<KM, 10, FV_VDIFQ_VHREG>), g.M + 1, g, a, a.vhpair, h->cl_flag)
#define VDIF_QM_CALL16(KM) TLAUNCH("vdif_q", (k_vdif_q_r<KM, 16, FV_VDIFQ_VHREG>), g.M + 1, g, a, a.vhpair, h->cl_flag)
    // (vdif_q also stores Q2 = Q2F, Q2L = Q2LF for its columns)
    if (g.ve_max <= 10) FV_KMAX_DISPATCH(g.KBM1, VDIF_QM_CALL);
    else FV_KMAX_DISPATCH(g.KBM1, VDIF_QM_CALL16);
  } else if (h->md->p.vertical_closure) {
    // ADV_Q of Q2 on the main stream, of Q2L on stream2 (own scratch)
    fork_streams(h, FV_CONC_3D);
    on_stream2(h, FV_CONC_3D);
    enqueue_adv_q<FT>(h, h->ib, a.q2l, a.q2lf, fct);
    on_stream1(h);
    enqueue_adv_q<FT>(h, a, a.q2, a.q2f, fct);
    join_streams(h, FV_CONC_3D);
#define VDIF_Q_CALL(KM) TLAUNCH("vdif_q", (k_vdif_q_r<KM, 10, FV_VDIFQ_VHREG>), g.M + 1, g, a, a.vhpair, (const int*)nullptr)
#define VDIF_Q_CALL16(KM) TLAUNCH("vdif_q", (k_vdif_q_r<KM, 16, FV_VDIFQ_VHREG>), g.M + 1, g, a, a.vhpair, (const int*)nullptr)
    // VDIF_Q, then Q2 = Q2F and Q2L = Q2LF (internal_step.F:1029-1035) in
    // the same kernel, the elimination coefficients in `vhpair`.
    if (g.ve_max <= 10) FV_KMAX_DISPATCH(g.KBM1, VDIF_Q_CALL);
    else FV_KMAX_DISPATCH(g.KBM1, VDIF_Q_CALL16);
  } else {
    // VERTICAL_MIXING_TYPE 'constant'
    const double umol = h->md->p.umol;
    TLAUNCH("const_mixing", k_const_mixing, mk, a.km, a.kh, umol, umol * h->md->p.vprnu, mk);
  }
  N2E3D(a.km, a.km1);
  // ADV_T .. T1 = TF1 on the main stream, the same for S on stream2 (own
  // scratch); the two passes share no written array
  const bool t_on = h->md->p.temperature_active, s_on = h->md->p.salinity_active;
  if (t_on && s_on) {
    fork_streams(h, FV_CONC_3D);
    on_stream2(h, FV_CONC_3D);
    enqueue_ts_pass<FT>(h, h->ib, 2, fct);
    on_stream1(h);
    enqueue_ts_pass<FT>(h, a, 1, fct);
    join_streams(h, FV_CONC_3D);
  } else if (t_on) {
    enqueue_ts_pass<FT>(h, a, 1, fct);
  } else if (s_on) {
    enqueue_ts_pass<FT>(h, a, 2, fct);
  }
  if (!g.river_ts_specified && g.NUMQBC > 0) LAUNCH("adjust_ts", k_adjust_ts, 1, 256, g, a);
  if (!h->md->p.barotropic) {
    TLAUNCH("dens2", k_dens2, mk, g, a);
    N2E3D(a.rho1, a.rho);
  }
  TLAUNCH("copy", k_copy, nk, a.u, a.uf, nk);
  TLAUNCH("copy", k_copy, nk, a.v, a.vf, nk);
  TLAUNCH("end_2d", k_end_2d, std::max(g.M, g.N) + 1, g, a);
}

void enqueue_external_step(fvgpu_handle* h, int iext) {
  if (h->precision == FV_MIXED) enqueue_external_step_t<float>(h, iext, true);
  else enqueue_external_step_t<double>(h, iext, true);
}

void enqueue_internal_step(fvgpu_handle* h) {
  if (h->precision == FV_MIXED) enqueue_internal_step_t<float>(h);
  else enqueue_internal_step_t<double>(h);
}

// Wind stress of element wind speeds (PSIMPLE_DRAG, mod_bulk.F:58-90),
// elements 1..n1-1.
void wind_stress(const double* uu, const double* vv, size_t n1, double* wx, double* wy) {
  for (size_t i = 1; i < n1; ++i) psimple_drag(uu[i], vv[i], wx[i], wy[i]);
}

// ------------------------------------------------------ derived structures
int build_lists(fvgpu_handle* h) {
  const fv_model* md = h->md;
  const int M = md->m, N = md->n, NE = md->ne, NCV = md->ncv;
  DevGrid& g = h->g;
  // node -> CV sub-edges sorted by sub-edge index
  std::vector<int> cnt(M + 2, 0);
  for (int c = 1; c <= NCV; ++c) { cnt[md->niec[c]]++; cnt[md->niec[(size_t)(NCV + 1) + c]]++; }
  std::vector<int> off(M + 2, 0);
  for (int i = 0; i <= M; ++i) off[i + 1] = off[i] + cnt[i];
  std::vector<int> ent(off[M + 1]), fill(M + 1, 0);
  for (int c = 1; c <= NCV; ++c) {
    const int ia = md->niec[c], ib = md->niec[(size_t)(NCV + 1) + c];
    ent[off[ia] + fill[ia]++] = (c << 1) | 0;
    ent[off[ib] + fill[ib]++] = (c << 1) | 1;
  }
  // element -> 3 sub-edges, and each sub-edge's element-major flux position
  std::vector<int> pos(NCV + 1, 0);
  {
    const size_t n1 = (size_t)N + 1;
    std::vector<int> ec(3 * n1, 0), esw(3 * n1, 0), ecnt(n1, 0);
    for (int c = 1; c <= NCV; ++c) {
      const int e = md->ntrg[c];
      if (e < 1 || e > N) return fail("a sub-edge's element NTRG is out of range");
      if (ecnt[e]++ >= 3) return fail("an element has more than 3 control-volume sub-edges");
      const int ia = md->niec[c], ib = md->niec[(size_t)(NCV + 1) + c];
      int la = -1, lb = -1;
      for (int q = 0; q < 3; ++q) {
        const int nq = md->nv[(size_t)q * n1 + e];
        if (nq == ia) la = q;
        if (nq == ib) lb = q;
      }
      if (la < 0 || lb < 0 || la == lb) return fail("a sub-edge's nodes are not two nodes of its element");
      // slot = the element edge: 0 joins NV(e,1)-NV(e,2), 1 NV(e,2)-NV(e,3),
      // 2 NV(e,3)-NV(e,1); swap when IA is the second node of that pair
      const int slot = (la + lb == 1) ? 0 : (la + lb == 3) ? 1 : 2;
      const int swap = la != slot;
      if (ec[(size_t)slot * n1 + e] != 0) return fail("two sub-edges of an element on one edge");
      ec[(size_t)slot * n1 + e] = c;
      esw[(size_t)slot * n1 + e] = swap;
      pos[c] = 3 * (e - 1) + slot + 1;
    }
    for (int e = 1; e <= N; ++e)
      if (ecnt[e] != 3) return fail("an element has fewer than 3 control-volume sub-edges");
    std::vector<double> edx(3 * n1, 0.0), edy(3 * n1, 0.0);
    for (size_t q = 0; q < 3 * n1; ++q)
      if (ec[q] > 0) { edx[q] = md->dltxe[ec[q]]; edy[q] = md->dltye[ec[q]]; }
    if (dupload(h, &g.ecv_dx, edx.data(), edx.size())) return 1;
    if (dupload(h, &g.ecv_dy, edy.data(), edy.size())) return 1;
    if (dupload(h, &g.ecv_c, ec.data(), ec.size())) return 1;
    if (dupload(h, &g.ecv_swap, esw.data(), esw.size())) return 1;
  }
  g.ring_max = 1;
  g.ve_max = 1;
  for (int i = 1; i <= M; ++i) {
    if (md->ntsn[i] > FV_MAX_RING || md->ntsn[i] < 1) return fail("node ring NTSN outside [1, FV_MAX_RING]");
    if (md->ntve[i] > FV_MAX_RING || md->ntve[i] < 1) return fail("node ring NTVE outside [1, FV_MAX_RING]");
    g.ring_max = std::max(g.ring_max, md->ntsn[i]);
    g.ve_max = std::max(g.ve_max, md->ntve[i]);
  }
  if (dupload(h, &g.cv_off, off.data(), off.size())) return 1;
  if (dupload(h, &g.cv_ent, ent.data(), ent.size())) return 1;
  {
    int smax = 0;
    for (int i = 0; i <= M; ++i) smax = std::max(smax, cnt[i]);
    if (smax > FV_MAX_CV) return fail("a node has more control-volume sub-edges than FV_MAX_CV");
    g.cv_max = smax;
    const size_t m1 = (size_t)M + 1, ns = (size_t)std::max(smax, 1) * m1;
    std::vector<int> le(ns, 0);
    for (int i = 0; i <= M; ++i)
      for (int j = 0; j < cnt[i]; ++j) {
        const int e = ent[off[i] + j], c = e >> 1;
        le[(size_t)j * m1 + i] = (pos[c] << 1) | (e & 1);
      }
    if (dupload(h, &g.cv_cnt, cnt.data(), m1)) return 1;
    if (dupload(h, &g.cvl_ent, le.data(), ns)) return 1;
  }
  // element -> its three edges sorted by edge index
  std::vector<int> ed(3 * (size_t)(N + 1), 0), efill(N + 1, 0);
  for (int e = 1; e <= NE; ++e) {
    const int ia = md->iec[e], ib = md->iec[(size_t)(NE + 1) + e];
    if (ia > 0) { if (efill[ia] >= 3) return fail("element with >3 edges"); ed[3 * ia + efill[ia]++] = (e << 1) | 0; }
    if (ib > 0) { if (efill[ib] >= 3) return fail("element with >3 edges"); ed[3 * ib + efill[ib]++] = (e << 1) | 1; }
  }
  for (int i = 1; i <= N; ++i) if (efill[i] != 3) return fail("element without 3 edges");
  if (dupload(h, &g.ed_ent, ed.data(), ed.size())) return 1;
  // FLUXOBN contributions per element in OBC order
  const int IOBCN = md->iobcn;
  std::vector<std::vector<int>> fo(N + 1);
  for (int i = 0; i < IOBCN; ++i) {
    if (md->nadjc_obc[i] > 0) fo[md->adjc_obc[i]].push_back((i << 1) | 0);
    if (md->nadjc_obc[i] > 1) fo[md->adjc_obc[(size_t)IOBCN + i]].push_back((i << 1) | 1);
  }
  std::vector<int> fo_off(N + 2, 0), fo_ent;
  for (int i = 0; i <= N; ++i) { fo_off[i + 1] = fo_off[i] + (int)fo[i].size(); for (int x : fo[i]) fo_ent.push_back(x); }
  if (dupload(h, &g.fo_off, fo_off.data(), fo_off.size())) return 1;
  if (dupload(h, &g.fo_obc, fo_ent.data(), fo_ent.size())) return 1;
  std::vector<int> nobc(M + 1, -1), nriv(M + 1, -1), skip(M + 1, 0);
  for (int i = 0; i < IOBCN; ++i) { nobc[md->i_obc_n[i]] = i; skip[md->i_obc_n[i]] = 1; }
  // node -> the first river whose discharge enters it (its node river, or
  // an edge river of which it bounds the solid edge); FCT skips them all
  const int NQ = md->numqbc;
  for (int j = NQ - 1; j >= 0; --j) {
    const int nd[2] = {md->icellq[j] > 0 ? md->n_icellq[j] : md->inodeq[j],
                       md->icellq[j] > 0 ? md->n_icellq[NQ + j] : 0};
    for (int q = 0; q < 2; ++q)
      if (nd[q] > 0) { nriv[nd[q]] = j; skip[nd[q]] = 1; }
  }
  if (dupload(h, &g.node_obc, nobc.data(), nobc.size())) return 1;
  if (dupload(h, &g.node_riv, nriv.data(), nriv.size())) return 1;
  if (dupload(h, &g.fct_skip, skip.data(), skip.size())) return 1;
  std::vector<std::vector<int>> tile_own;   // the element tiles below, for the momentum blocks too
  // element blocks of the fused record + face-flux kernels: recursive
  // coordinate bisection of the element centroids (each range halved at the
  // median of its longer extent) down to blocks of at most FV_CVB elements
  // that use at most FV_CVB distinct nodes -- compact tiles, so that few of
  // a block's nodes are formed again by a neighbouring block.  Blocks in the
  // bisection's depth-first order (neighbouring tiles in neighbouring
  // blocks), each block's elements ascending.
  {
    // connected components of the mesh (the instances of a union of
    // meshes share coordinates but no node): the bisection starts from one
    // range per component, so that no tile mixes two of them
    std::vector<int> ecomp(N + 1, -1);
    int ncomp = 0;
    {
      std::vector<int> queue;
      for (int s0 = 1; s0 <= N; ++s0) {
        if (ecomp[s0] >= 0) continue;
        ecomp[s0] = ncomp;
        queue.assign(1, s0);
        for (size_t q = 0; q < queue.size(); ++q) {
          const int e = queue[q];
          for (int j = 0; j < 3; ++j) {
            const int x = md->nbe[(size_t)j * (N + 1) + e];
            if (x > 0 && ecomp[x] < 0) { ecomp[x] = ncomp; queue.push_back(x); }
          }
        }
        ++ncomp;
      }
    }
    std::vector<int> ncompv(M + 1, 0);
    for (int e = 1; e <= N; ++e)
      for (int j = 0; j < 3; ++j) ncompv[md->nv[(size_t)j * (N + 1) + e]] = ecomp[e];
    // ranges [lo, hi) of ids sorted by component, one per component
    auto component_ranges = [&](std::vector<int>& idv, const std::vector<int>& comp) {
      std::stable_sort(idv.begin(), idv.end(), [&](int a, int b) { return comp[a] < comp[b]; });
      std::vector<std::pair<int, int>> r;
      int lo = 0;
      for (int q = 1; q <= (int)idv.size(); ++q)
        if (q == (int)idv.size() || comp[idv[q]] != comp[idv[lo]]) { r.push_back({lo, q}); lo = q; }
      std::reverse(r.begin(), r.end());   // popped from the back: first component first
      return r;
    };
    std::vector<int> ids(N);
    for (int e = 1; e <= N; ++e) ids[e - 1] = e;
    std::vector<int> mark(M + 1, -1);
    int stamp = 0;
    auto distinct_nodes = [&](int lo, int hi) {
      ++stamp;
      int n = 0;
      for (int q = lo; q < hi; ++q)
        for (int j = 0; j < 3; ++j) {
          const int nd = md->nv[(size_t)j * (N + 1) + ids[q]];
          if (mark[nd] != stamp) { mark[nd] = stamp; ++n; }
        }
      return n;
    };
    std::vector<std::pair<int, int>> stack = component_ranges(ids, ecomp), leaves;
    while (!stack.empty()) {
      const std::pair<int, int> r = stack.back();
      stack.pop_back();
      const int lo = r.first, hi = r.second, n = hi - lo;
      if (n <= FV_CVB && distinct_nodes(lo, hi) <= FV_CVB) { leaves.push_back(r); continue; }
      double x0 = md->xc[ids[lo]], x1 = x0, y0 = md->yc[ids[lo]], y1 = y0;
      for (int q = lo + 1; q < hi; ++q) {
        const double x = md->xc[ids[q]], y = md->yc[ids[q]];
        x0 = std::min(x0, x); x1 = std::max(x1, x); y0 = std::min(y0, y); y1 = std::max(y1, y);
      }
      const double* key = (x1 - x0) >= (y1 - y0) ? md->xc : md->yc;
      const int mid = lo + (n + 1) / 2;
      std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi,
                       [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
      stack.push_back({mid, hi});
      stack.push_back({lo, mid});
    }
    const int nblk = (int)leaves.size();
    std::vector<int> bnn(nblk, 0), bne(nblk, 0), bnodes((size_t)nblk * FV_CVB, 1), belems((size_t)nblk * FV_CVB, 1);
    std::vector<int> loc((size_t)N + 1, 0), nloc(M + 1, 0);
    for (int b = 0; b < nblk; ++b) {
      const int lo = leaves[b].first, hi = leaves[b].second;
      std::sort(ids.begin() + lo, ids.begin() + hi);
      ++stamp;
      int nn = 0;
      for (int q = lo; q < hi; ++q) {
        const int e = ids[q];
        belems[(size_t)b * FV_CVB + (q - lo)] = e;
        int packed = 0;
        for (int j = 0; j < 3; ++j) {
          const int nd = md->nv[(size_t)j * (N + 1) + e];
          if (mark[nd] != stamp) {
            mark[nd] = stamp;
            nloc[nd] = nn;
            bnodes[(size_t)b * FV_CVB + nn++] = nd;
          }
          packed |= nloc[nd] << (8 * j);
        }
        loc[e] = packed;
      }
      bnn[b] = nn;
      bne[b] = hi - lo;
    }
    g.cvb_nblk = nblk;
    tile_own.assign(nblk, {});
    for (int b = 0; b < nblk; ++b)
      tile_own[b].assign(belems.begin() + (size_t)b * FV_CVB, belems.begin() + (size_t)b * FV_CVB + bne[b]);
    // node tiles (VERTVL): recursive coordinate bisection of the nodes down
    // to tiles of at most FV_NTN nodes whose elements (every element that
    // has one of them as a vertex) number at most FV_NTE; each owned node's
    // sub-edges as positions 3*(local element) + slot in the tile's element
    // list, sub-edge order kept, side in bit 0
    {
      std::vector<int> nids(M);
      for (int i = 1; i <= M; ++i) nids[i - 1] = i;
      std::vector<int> emark(N + 1, -1);
      int estamp = 0;
      auto elems_around = [&](int lo, int hi) {
        ++estamp;
        int n = 0;
        for (int q = lo; q < hi; ++q) {
          const int i = nids[q];
          for (int r = md->nbve_off[i]; r < md->nbve_off[i] + md->ntve[i]; ++r)
            if (emark[md->nbve[r]] != estamp) { emark[md->nbve[r]] = estamp; ++n; }
        }
        return n;
      };
      std::vector<std::pair<int, int>> st = component_ranges(nids, ncompv), tl;
      while (!st.empty()) {
        const std::pair<int, int> r = st.back();
        st.pop_back();
        const int lo = r.first, hi = r.second, n = hi - lo;
        if (n <= FV_NTN && elems_around(lo, hi) <= FV_NTE) { tl.push_back(r); continue; }
        if (n == 1) return fail("a node has more elements around it than a node tile holds");
        double x0 = md->vx[nids[lo]], x1 = x0, y0 = md->vy[nids[lo]], y1 = y0;
        for (int q = lo + 1; q < hi; ++q) {
          const double x = md->vx[nids[q]], y = md->vy[nids[q]];
          x0 = std::min(x0, x); x1 = std::max(x1, x); y0 = std::min(y0, y); y1 = std::max(y1, y);
        }
        const double* key = (x1 - x0) >= (y1 - y0) ? md->vx : md->vy;
        const int mid = lo + (n + 1) / 2;
        std::nth_element(nids.begin() + lo, nids.begin() + mid, nids.begin() + hi,
                         [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
        st.push_back({mid, hi});
        st.push_back({lo, mid});
      }
      const int ntb = (int)tl.size();
      const size_t m1 = (size_t)M + 1;
      std::vector<int> tnown(ntb, 0), tnel(ntb, 0), tnodes((size_t)ntb * FV_NTN, 1), telems((size_t)ntb * FV_NTE, 1);
      std::vector<int> tlpos((size_t)std::max(g.cv_max, 1) * m1, 0), eloc(N + 1, 0);
      for (int b = 0; b < ntb; ++b) {
        const int lo = tl[b].first, hi = tl[b].second;
        std::sort(nids.begin() + lo, nids.begin() + hi);
        std::vector<int> els;
        for (int q = lo; q < hi; ++q) {
          const int i = nids[q];
          tnodes[(size_t)b * FV_NTN + (q - lo)] = i;
          for (int r = md->nbve_off[i]; r < md->nbve_off[i] + md->ntve[i]; ++r) els.push_back(md->nbve[r]);
        }
        std::sort(els.begin(), els.end());
        els.erase(std::unique(els.begin(), els.end()), els.end());
        for (size_t q = 0; q < els.size(); ++q) {
          telems[(size_t)b * FV_NTE + q] = els[q];
          eloc[els[q]] = (int)q;
        }
        tnown[b] = hi - lo;
        tnel[b] = (int)els.size();
        for (int q = lo; q < hi; ++q) {
          const int i = nids[q];
          for (int j = 0; j < cnt[i]; ++j) {
            const int en = ent[off[i] + j], c = en >> 1;
            const int e = md->ntrg[c], slot = (pos[c] - 1) % 3;
            tlpos[(size_t)j * m1 + i] = ((3 * eloc[e] + slot) << 1) | (en & 1);
          }
        }
      }
      g.nt_nblk = ntb;
      if (dupload(h, &g.nt_nown, tnown.data(), tnown.size())) return 1;
      if (dupload(h, &g.nt_nodes, tnodes.data(), tnodes.size())) return 1;
      if (dupload(h, &g.nt_nel, tnel.data(), tnel.size())) return 1;
      if (dupload(h, &g.nt_elems, telems.data(), telems.size())) return 1;
      if (dupload(h, &g.nt_lpos, tlpos.data(), tlpos.size())) return 1;
    }
    std::vector<int> keep(M + 1, 0);
    for (int i = 0; i < md->iobcn; ++i) {
      const int j = md->i_obc_n[i];
      keep[j] = 1;
      for (int q = off[j]; q < off[j + 1]; ++q) {
        const int c = ent[q] >> 1;
        keep[md->niec[c]] = 1;
        keep[md->niec[(size_t)(NCV + 1) + c]] = 1;
      }
    }
    if (dupload(h, &g.cvb_nn, bnn.data(), bnn.size())) return 1;
    if (dupload(h, &g.cvb_nodes, bnodes.data(), bnodes.size())) return 1;
    if (dupload(h, &g.cvb_ne, bne.data(), bne.size())) return 1;
    if (dupload(h, &g.cvb_elems, belems.data(), belems.size())) return 1;
    if (dupload(h, &g.cvb_loc, loc.data(), loc.size())) return 1;
    if (dupload(h, &g.rec_keep, keep.data(), keep.size())) return 1;
  }
  // element blocks of the blocked ADVECTION / ADV_UV kernels.  ADVECTION
  // takes the element tiles above (compact, so few neighbours form their
  // records twice) if every tile needs at most 256 records (its own
  // elements and the other elements of their edges) and at most 256 edges;
  // ADV_UV, whose element part reads and writes many element-layer
  // arrays, takes runs of consecutive elements (contiguous addresses), the
  // largest run length (a multiple of 8, at most 128) that keeps every
  // block within both, and so does ADVECTION where a tile would not fit
  {
    std::vector<int> bns, bnown, belems, bne, bedges, beloc, epos;
    std::vector<int> rslot(N + 1, -1), rblk(N + 1, -1), eslot(NE + 1, -1), eblk(NE + 1, -1);
    auto build_blocks = [&](const std::vector<std::vector<int>>& own) {
      const int nblk = (int)own.size();
      bns.assign(nblk, 0); bnown.assign(nblk, 0); bne.assign(nblk, 0);
      belems.assign((size_t)nblk * 256, 0); bedges.assign((size_t)nblk * 256, 1); beloc.assign((size_t)nblk * 256, 0);
      epos.assign((size_t)N + 1, 0);
      std::fill(rblk.begin(), rblk.end(), -1);
      std::fill(eblk.begin(), eblk.end(), -1);
      int worst = 0;
      for (int b = 0; b < nblk; ++b) {
        int ns = 0, ne = 0;
        auto slot_of = [&](int x) {
          if (rblk[x] != b) {
            if (ns < 256) belems[(size_t)b * 256 + ns] = x;
            rblk[x] = b;
            rslot[x] = ns++;
          }
          return rslot[x];
        };
        for (int e : own[b]) slot_of(e);   // own elements first, in order
        for (int e : own[b]) {
          int packed = 0;
          for (int q = 0; q < 3; ++q) {
            const int c = ed[3 * (size_t)e + q] >> 1;
            if (eblk[c] != b) {
              const int ia = md->iec[c], ib = md->iec[(size_t)(NE + 1) + c];
              const int sa = slot_of(ia), sb = slot_of(ib);
              if (ne < 256) {
                bedges[(size_t)b * 256 + ne] = c;
                beloc[(size_t)b * 256 + ne] = (sa & 255) | ((sb & 255) << 8);
              }
              eblk[c] = b;
              eslot[c] = ne++;
            }
            packed |= (eslot[c] & 255) << (8 * q);
          }
          epos[e] = packed;
        }
        bns[b] = ns;
        bnown[b] = (int)own[b].size();
        bne[b] = ne;
        worst = std::max(worst, std::max(ns, ne));
      }
      return worst;
    };
    auto upload_blocks = [&](UvBlocks& u) {
      u.nblk = (int)bns.size();
      return dupload(h, &u.nown, bnown.data(), bnown.size()) || dupload(h, &u.ns, bns.data(), bns.size()) ||
             dupload(h, &u.elems, belems.data(), belems.size()) || dupload(h, &u.ne, bne.data(), bne.size()) ||
             dupload(h, &u.edges, bedges.data(), bedges.size()) ||
             dupload(h, &u.eloc, beloc.data(), beloc.size()) || dupload(h, &u.epos, epos.data(), epos.size());
    };
    int eb = 128;
    for (;;) {
      if (eb < 8) return fail("no element block size keeps the blocked momentum kernels within 256 records");
      std::vector<std::vector<int>> runs((N + eb - 1) / eb);
      for (int e = 1; e <= N; ++e) runs[(e - 1) / eb].push_back(e);
      const int worst = build_blocks(runs);
      if (worst <= 256) break;
      // shrink in proportion, to a multiple of 8, by at least 8
      eb = std::min(eb - 8, (int)((long)eb * 256 / worst) / 8 * 8);
    }
    if (upload_blocks(g.ub_run)) return 1;
    bool tiles_fit = true;
    for (const auto& t : tile_own) tiles_fit = tiles_fit && t.size() <= 128;
    if (tiles_fit && build_blocks(tile_own) <= 256) {
      if (upload_blocks(g.ub_tile)) return 1;
    } else {
      g.ub_tile = g.ub_run;
    }
  }
  // element -> rivers in river order: NBVE(J,1) then NBVE(J,NTVE(J)) of a
  // node river's node, an edge river's own cell
  std::vector<std::vector<int>> re(N + 1);
  for (int k = 0; k < md->numqbc; ++k) {
    if (md->icellq[k] > 0) { re[md->icellq[k]].push_back(k); continue; }
    const int j = md->inodeq[k];
    const int i1 = md->nbve[md->nbve_off[j]], i2 = md->nbve[md->nbve_off[j] + md->ntve[j] - 1];
    re[i1].push_back(k);
    re[i2].push_back(k);
  }
  std::vector<int> re_off(N + 2, 0), re_ent;
  for (int i = 0; i <= N; ++i) { re_off[i + 1] = re_off[i] + (int)re[i].size(); for (int x : re[i]) re_ent.push_back(x); }
  if (dupload(h, &g.re_off, re_off.data(), re_off.size())) return 1;
  if (dupload(h, &g.re_riv, re_ent.data(), re_ent.size())) return 1;
  return 0;
}

bool columns_uniform(const double* a, long l1, int kb, int first) {
  for (int k = 0; k < kb; ++k) {
    const double v = a[(long)k * l1 + first];
    for (long i = first; i < l1; ++i)
      if (std::memcmp(&a[(long)k * l1 + i], &v, sizeof(double)) != 0) return false;
  }
  return true;
}

int upload_grid(fvgpu_handle* h) {
  const fv_model* md = h->md;
  DevGrid& g = h->g;
  const int M = md->m, N = md->n, KB = md->kb;
  const size_t m1 = M + 1, n1 = N + 1;
  g.M = M; g.N = N; g.NE = md->ne; g.NCV = md->ncv; g.KB = KB; g.KBM1 = md->kbm1; g.KBM2 = md->kbm2;
  g.IOBCN = md->iobcn; g.NUMQBC = md->numqbc; g.NTIDE = md->ntide;
  g.m1 = (long)m1; g.n1 = (long)n1;
  if (KB > FV_MAXKB) return fail("KB exceeds FV_MAXKB");
  const fv_params& p = md->p;
  g.dte = p.dte; g.dti = p.dte * p.isplit; g.umol = p.umol; g.vprnu = p.vprnu; g.hprnu = p.hprnu;
  g.cbcmin = p.cbcmin; g.min_depth = p.min_depth;
  g.fact = p.horizontal_closure ? 1.0 : 0.0; g.fm1 = 1.0 - g.fact;
  g.const1 = std::pow(16.6, .6666667) * 1.00;
  g.cvhp1 = std::pow(15.8 * 100.0, 0.6666667);
  g.isplit = p.isplit; g.iramp = p.iramp; g.noflux_bot = p.noflux_bot;
  g.river_ts_specified = p.river_ts_specified; g.type_tsobc = p.type_tsobc; g.obc_on = p.obc_on;
  if (p.type_tsobc < 1 || p.type_tsobc > 4) return fail("TYPE_TSOBC must be 1..4");
  g.obc_elev = p.obc_elevation_on; g.wind_on = p.wind_on;
#define UP(field, n) if (dupload(h, &g.field, md->field, (n))) return 1
  UP(vx, m1); UP(vy, m1); UP(h, m1); UP(art1, m1); UP(art2, m1); UP(grav_n, m1); UP(nn_hvc, m1); UP(phpn, m1);
  UP(isonb, m1); UP(ntsn, m1); UP(ntve, m1); UP(nbsn_off, m1 + 1); UP(nbve_off, m1 + 1);
  UP(nbsn, md->nring_sn); UP(dltxtrie, md->nring_sn); UP(dltytrie, md->nring_sn);
  UP(nbve, md->nring_ve); UP(dltxecec, md->nring_ve); UP(dltyecec, md->nring_ve);
  UP(dltxnec, md->nring_ve); UP(dltynec, md->nring_ve);
  UP(xc, n1); UP(yc, n1); UP(art, n1); UP(h1, n1); UP(cor, n1); UP(cc_hvc, n1); UP(cc_sponge, n1);
  UP(cc_z0b, n1); UP(grav_e, n1); UP(epor, n1); UP(isbce, n1); UP(iucp, n1);
  UP(nv, 3 * n1); UP(nbe, 3 * n1); UP(a1u, 4 * n1); UP(a2u, 4 * n1); UP(awx, 3 * n1); UP(awy, 3 * n1);
  {
    std::vector<double> a12(8 * n1);
    for (size_t e = 0; e < n1; ++e)
      for (int j = 0; j < 4; ++j) {
        a12[8 * e + j] = md->a1u[(size_t)j * n1 + e];
        a12[8 * e + 4 + j] = md->a2u[(size_t)j * n1 + e];
      }
    if (dupload(h, &g.a12, a12.data(), a12.size())) return 1;
  }
  const size_t ne1 = md->ne + 1, cv1 = md->ncv + 1;
  UP(iec, 2 * ne1); UP(ienode, 2 * ne1); UP(isbc, ne1); UP(xijc, ne1); UP(yijc, ne1); UP(dltxc, ne1); UP(dltyc, ne1);
  UP(niec, 2 * cv1); UP(ntrg, cv1); UP(dltxe, cv1); UP(dltye, cv1); UP(dltxncve, 2 * cv1); UP(dltyncve, 2 * cv1);
  const size_t ob = md->iobcn, nt = md->ntide, nr = md->numqbc;
  for (size_t k = 0; k < ob; ++k)
    if (md->type_obc[k] < 1 || md->type_obc[k] > 10) return fail("TYPE_OBC must be 1..10");
  if (p.type_tsobc == 4)
    for (size_t k = 0; k < ob; ++k)
      if (md->next_obc2[k] <= 0) return fail("TYPE_TSOBC 4 needs NEXT_OBC2 for every OBC node");
  g.obc_radiate = 0;
  g.obc_ore = 0;
  for (size_t k = 0; k < ob; ++k) {
    if (md->type_obc[k] >= 5) g.obc_radiate = 1;
    if (md->type_obc[k] >= 9) g.obc_ore = 1;
  }
  UP(i_obc_n, ob); UP(next_obc, ob); UP(nadjc_obc, ob); UP(adjc_obc, 2 * ob); UP(fluxf_obc, 2 * ob);
  UP(type_obc, ob); UP(next_obc2, ob);
  UP(dltn_obc, ob); UP(emean, ob); UP(apt, ob * nt); UP(phai, ob * nt); UP(period, nt);
  for (size_t k = 0; k < nr; ++k)
    if ((md->inodeq[k] > 0) == (md->icellq[k] > 0))
      return fail("each river must be either a node river (INODEQ) or an edge river (ICELLQ)");
  UP(inodeq, nr); UP(qarea, nr); UP(qdis0, nr); UP(tdis, nr); UP(sdis, nr); UP(vqdist, nr * md->kbm1);
  UP(icellq, nr); UP(n_icellq, 2 * nr); UP(rdisq, 2 * nr);
#undef UP
  // host-evaluated trigonometry of setup angles (the same libm the setup uses)
  std::vector<double> cs(m1), sn(m1);
  for (size_t i = 0; i < m1; ++i) { cs[i] = std::cos(md->sita_gd[i]); sn[i] = std::sin(md->sita_gd[i]); }
  if (dupload(h, &g.cos_sita, cs.data(), m1)) return 1;
  if (dupload(h, &g.sin_sita, sn.data(), m1)) return 1;
  std::vector<double> ca(n1), sa(n1);
  for (size_t i = 0; i < n1; ++i) { ca[i] = std::cos(md->alpha[i]); sa[i] = std::sin(md->alpha[i]); }
  if (dupload(h, &g.cos_alpha, ca.data(), n1)) return 1;
  if (dupload(h, &g.sin_alpha, sa.data(), n1)) return 1;
  std::vector<double> cq(nr), sq_(nr);
  for (size_t i = 0; i < nr; ++i) { cq[i] = std::cos(md->angleq[i]); sq_[i] = std::sin(md->angleq[i]); }
  if (dupload(h, &g.cos_angleq, cq.data(), nr)) return 1;
  if (dupload(h, &g.sin_angleq, sq_.data(), nr)) return 1;
  std::vector<double> wx(n1, 0.0), wy(n1, 0.0);
  if (p.wind_on) wind_stress(md->uuwind, md->vvwind, n1, wx.data(), wy.data());
  if (dupload(h, &g.wsx, wx.data(), n1)) return 1;
  if (dupload(h, &g.wsy, wy.data(), n1)) return 1;
  if (p.wind_on) {
    const double* w0s[4] = {md->uuwind0, md->vvwind0, md->uuwind, md->vvwind};
    double** dst[4] = {&h->wu0, &h->wv0, &h->wu1, &h->wv1};
    for (int q = 0; q < 4; ++q) {
      const double* src = nullptr;
      if (dupload(h, &src, w0s[q], n1)) return 1;
      *dst[q] = (double*)src;
    }
    if (dalloc(h, &h->ws2x, n1) || dalloc(h, &h->ws2y, n1)) return 1;
    h->wind_varies = std::memcmp(md->uuwind0, md->uuwind, n1 * sizeof(double)) != 0 ||
                     std::memcmp(md->vvwind0, md->vvwind, n1 * sizeof(double)) != 0;
  }
  // sigma: one column when every column is identical
  const bool un = columns_uniform(md->z, m1, KB, 0) && columns_uniform(md->dz, m1, KB, 0) &&
                  columns_uniform(md->dzz, m1, KB, 0) && columns_uniform(md->zz, m1, KB, 0);
  const bool ue = columns_uniform(md->z1, n1, KB, 1) && columns_uniform(md->dz1, n1, KB, 1) &&
                  columns_uniform(md->dzz1, n1, KB, 1) && columns_uniform(md->zz1, n1, KB, 1);
  g.sig_uniform = (un && ue) ? 1 : 0;
  auto col = [&](const double* a, size_t l1, size_t first) {
    std::vector<double> c(KB);
    for (int k = 0; k < KB; ++k) c[k] = a[(size_t)k * l1 + first];
    return c;
  };
  if (g.sig_uniform) {
    std::vector<double> c;
    c = col(md->z, m1, 0); if (dupload(h, &g.z, c.data(), KB)) return 1;
    c = col(md->zz, m1, 0); if (dupload(h, &g.zz, c.data(), KB)) return 1;
    c = col(md->dz, m1, 0); if (dupload(h, &g.dz, c.data(), KB)) return 1;
    c = col(md->dzz, m1, 0); if (dupload(h, &g.dzz, c.data(), KB)) return 1;
    c = col(md->z1, n1, 1); if (dupload(h, &g.z1, c.data(), KB)) return 1;
    c = col(md->zz1, n1, 1); if (dupload(h, &g.zz1, c.data(), KB)) return 1;
    c = col(md->dz1, n1, 1); if (dupload(h, &g.dz1, c.data(), KB)) return 1;
    c = col(md->dzz1, n1, 1); if (dupload(h, &g.dzz1, c.data(), KB)) return 1;
  } else {
    if (dupload(h, &g.z, md->z, m1 * KB)) return 1;
    if (dupload(h, &g.zz, md->zz, m1 * KB)) return 1;
    if (dupload(h, &g.dz, md->dz, m1 * KB)) return 1;
    if (dupload(h, &g.dzz, md->dzz, m1 * KB)) return 1;
    if (dupload(h, &g.z1, md->z1, n1 * KB)) return 1;
    if (dupload(h, &g.zz1, md->zz1, n1 * KB)) return 1;
    if (dupload(h, &g.dz1, md->dz1, n1 * KB)) return 1;
    if (dupload(h, &g.dzz1, md->dzz1, n1 * KB)) return 1;
  }
  return build_lists(h);
}

int upload_state(fvgpu_handle* h) {
  const fv_model* md = h->md;
  const DevGrid& g = h->g;
  const size_t m1 = g.m1, n1 = g.n1, mk = m1 * g.KB, nk = n1 * g.KB;
  IntArgs& a = h->ia;
  for (int b = 0; b < 3; ++b) {
    Ext2DState& s = h->ext[b];
    if (dupload_rw(h, &s.el, b == 0 ? md->el : nullptr, m1)) return 1;
    if (dupload_rw(h, &s.d, b == 0 ? md->d : nullptr, m1)) return 1;
    if (dupload_rw(h, &s.ua, b == 0 ? md->ua : nullptr, n1)) return 1;
    if (dupload_rw(h, &s.va, b == 0 ? md->va : nullptr, n1)) return 1;
    if (dupload_rw(h, &s.el1, b == 0 ? md->el1 : nullptr, n1)) return 1;
    if (dupload_rw(h, &s.d1, b == 0 ? md->d1 : nullptr, n1)) return 1;
  }
  a.el = h->ext[0].el; a.d = h->ext[0].d; a.ua = h->ext[0].ua; a.va = h->ext[0].va;
  a.el1 = h->ext[0].el1; a.d1 = h->ext[0].d1;
  a.elf = a.el;   // after the external loop ELF = EL
#define RW(field, src, n) if (dupload_rw(h, &a.field, src, (n))) return 1
  RW(et, md->et, m1); RW(dt, md->dt, m1); RW(et1, md->et1, n1); RW(dt1, md->dt1, n1);
  RW(egf, nullptr, m1); RW(dtfa, md->d, m1); RW(elm1, md->elm1, m1); RW(elm2, md->elm2, m1);
  RW(u, md->u, nk); RW(v, md->v, nk); RW(w, md->w, nk); RW(ww, md->ww, nk); RW(uf, nullptr, nk); RW(vf, nullptr, nk);
  RW(drhox, nullptr, nk); RW(drhoy, nullptr, nk); RW(rho, md->rho, nk); RW(rmean, md->rmean, nk);
  RW(t, md->t, nk); RW(s, md->s, nk); RW(km1, md->km1, nk); RW(viscofm, nullptr, nk);
  RW(t1, md->t1, mk); RW(s1, md->s1, mk); RW(tf1, nullptr, mk); RW(sf1, nullptr, mk); RW(rho1, md->rho1, mk);
  RW(rmean1, md->rmean1, mk); RW(tmean1, md->tmean1, mk); RW(smean1, md->smean1, mk);
  RW(q2, md->q2, mk); RW(q2l, md->q2l, mk); RW(q2f, nullptr, mk); RW(q2lf, nullptr, mk);
  RW(l, md->l, mk); RW(km, md->km, mk); RW(kh, md->kh, mk); RW(kq, md->kq, mk); RW(wts, md->wts, mk);
  RW(viscofh, nullptr, mk);
  RW(gx, nullptr, mk);
  // flux buffers sized for FP64 (the mixed form uses the first half)
  {
    EdgeFlux4T<double>* p2 = nullptr; CvRecT<double>* pr = nullptr;
    double* pc = nullptr;
    if (dalloc(h, &p2, (size_t)g.NE + 1)) return 1;
    if (dalloc(h, &pr, mk)) return 1;
    if (dalloc(h, &pc, ((size_t)g.NCV + 1) * g.KB)) return 1;
    a.ef2 = p2; a.rec = pr; a.cvflux = pc;
  }
  RW(elij2, nullptr, (size_t)g.NE + 1);
  if (dalloc(h, &a.vhpair, mk)) return 1;
  if (dalloc(h, &a.obcsum, 5 * (size_t)g.IOBCN)) return 1;
  RW(adx2d, nullptr, n1); RW(ady2d, nullptr, n1); RW(drx2d, nullptr, n1); RW(dry2d, nullptr, n1);
  RW(cbc, nullptr, n1); RW(wubot, nullptr, n1); RW(wvbot, nullptr, n1); RW(tps, nullptr, n1);
  RW(uard, nullptr, n1); RW(vard, nullptr, n1); RW(fluxobn, nullptr, n1);
  RW(pfpxb, nullptr, m1); RW(pfpyb, nullptr, m1); RW(ah_bottom, nullptr, m1);
  RW(qdis, nullptr, (size_t)g.NUMQBC); RW(xflux_obc, nullptr, (size_t)g.IOBCN * g.KB);
  RW(uard_obcn, nullptr, (size_t)g.IOBCN);
#undef RW
  if (dalloc(h, &h->xflux_obcn, (size_t)g.IOBCN)) return 1;
  if (dalloc(h, &h->phi, (size_t)g.NCV + 1)) return 1;
  if (dalloc(h, &h->fobn, (size_t)g.IOBCN)) return 1;
  if (dalloc(h, &h->clk, 1)) return 1;
  {
    double* dp = nullptr;
    if (dalloc(h, &dp, 256 * 8)) return 1;  // 256 x DiagAcc (8 words each)
    h->diag_part = dp;
  }
  DevClock c{};
  c.iint = md->p.iint0; c.iext = 0; c.ramp_int = 0; c.ramp_ext = 0;
  c.ext_time = md->p.time0; c.time0 = md->p.time0; c.dti = g.dti; c.dte = g.dte;
  c.isplit = g.isplit; c.iramp = g.iramp; c.depth_fail = 0;
  HIPCHK(hipMemcpy(h->clk, &c, sizeof(c), hipMemcpyHostToDevice));
  a.clk = h->clk;
  // the stream2 chains: the same arrays, their own scratch (node-layer
  // records, face fluxes, bottom gradients, OBC fluxes)
  IntArgs& b = h->ib;
  b = a;
  {
    CvRecT<double>* pr = nullptr;
    double* pc = nullptr;
    if (dalloc(h, &pr, mk)) return 1;
    if (dalloc(h, &pc, ((size_t)g.NCV + 1) * g.KB)) return 1;
    b.rec = pr; b.cvflux = pc;
  }
  if (dalloc(h, &b.pfpxb, m1) || dalloc(h, &b.pfpyb, m1) || dalloc(h, &b.ah_bottom, m1)) return 1;
  if (dalloc(h, &b.xflux_obc, (size_t)g.IOBCN * g.KB)) return 1;
  if (dalloc(h, &b.obcsum, 5 * (size_t)g.IOBCN)) return 1;
  return 0;
}

double* lookup(fvgpu_handle* h, const std::string& n, size_t* len) {
  IntArgs& a = h->ia;
  const size_t m1 = h->g.m1, n1 = h->g.n1, mk = m1 * h->g.KB, nk = n1 * h->g.KB;
  struct E { const char* n; double* p; size_t l; };
  const E t[] = {
      {"el", a.el, m1}, {"d", a.d, m1}, {"et", a.et, m1}, {"dt", a.dt, m1}, {"elf", a.el, m1},
      {"egf", a.egf, m1}, {"dtfa", a.dtfa, m1}, {"elm1", a.elm1, m1}, {"elm2", a.elm2, m1},
      {"ua", a.ua, n1}, {"va", a.va, n1}, {"el1", a.el1, n1}, {"d1", a.d1, n1}, {"et1", a.et1, n1},
      {"dt1", a.dt1, n1}, {"adx2d", a.adx2d, n1}, {"ady2d", a.ady2d, n1}, {"drx2d", a.drx2d, n1},
      {"dry2d", a.dry2d, n1}, {"cbc", a.cbc, n1}, {"wubot", a.wubot, n1}, {"wvbot", a.wvbot, n1},
      {"uard", a.uard, n1}, {"vard", a.vard, n1},
      {"u", a.u, nk}, {"v", a.v, nk}, {"w", a.w, nk}, {"ww", a.ww, nk}, {"uf", a.uf, nk}, {"vf", a.vf, nk},
      {"drhox", a.drhox, nk}, {"drhoy", a.drhoy, nk}, {"rho", a.rho, nk}, {"rmean", a.rmean, nk},
      {"t", a.t, nk}, {"s", a.s, nk}, {"km1", a.km1, nk}, {"viscofm", a.viscofm, nk},
      {"t1", a.t1, mk}, {"s1", a.s1, mk}, {"tf1", a.tf1, mk}, {"sf1", a.sf1, mk}, {"rho1", a.rho1, mk},
      {"rmean1", a.rmean1, mk}, {"tmean1", a.tmean1, mk}, {"smean1", a.smean1, mk}, {"q2", a.q2, mk},
      {"q2l", a.q2l, mk}, {"q2f", a.q2f, mk}, {"q2lf", a.q2lf, mk}, {"l", a.l, mk}, {"km", a.km, mk},
      {"kh", a.kh, mk}, {"kq", a.kq, mk}, {"wts", a.wts, mk}, {"viscofh", a.viscofh, mk},
  };
  for (const E& e : t)
    if (n == e.n) { *len = e.l; return e.p; }
  return nullptr;
}

}  // namespace

// ---------------------------------------------------------------- C-ABI
extern "C" {

const char* fvgpu_last_error(void) { return g_err.c_str(); }

int fvgpu_device_count(int32_t* n) {
  int c = 0;
  HIPCHK(hipGetDeviceCount(&c));
  *n = c;
  return 0;
}

int fvgpu_create(const fv_model* md, int32_t precision, fvgpu_handle** out) {
  if (!md || !out) return fail("null argument");
  if (precision != FV_FP64 && precision != FV_MIXED) return fail("unknown precision mode");
  if (md->kb > FV_MAXKB) return fail("KB too large");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail("no HIP device");
  fvgpu_handle* h = new fvgpu_handle();
  h->md = md;
  h->precision = precision;
  HIPCHK(hipGetDevice(&h->device));
  HIPCHK(hipDeviceGetAttribute(&h->cu_count, hipDeviceAttributeMultiprocessorCount, h->device));
  if (!scratch_of(h->device)) { delete h; return 1; }
  // FP64: CV records + face fluxes in one launch; the FP32 flux form keeps
  // the two launches (measured faster for it: its records are half the
  // size, and the two kernels' higher occupancy outweighs their traffic)
  // the faster form per precision: in FP32 the fused tracer pass (980
  // against 505 + 620 us) but the two-phase Q2 pass (386 + 491 against
  // 1013 us), in FP64 both fused
  h->two_phase_ts = 0;
  h->two_phase_q = precision == FV_MIXED ? 1 : 0;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) { delete h; return fail("stream"); }
  h->cur = h->stream;
  if (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess) {
    fvgpu_destroy(h);
    return fail("stream2 / events");
  }
  if (upload_grid(h) || upload_state(h)) { fvgpu_destroy(h); return 1; }
  HIPCHK(hipStreamSynchronize(h->stream));
  *out = h;
  return 0;
}

int fvgpu_internal_step(fvgpu_handle* h) {
  enqueue_internal_step(h);
  HIPCHK(hipGetLastError());
  if (h->profiling) return flush_profile(h);
  return 0;
}

int fvgpu_external_step(fvgpu_handle* h, int32_t iext) {
  // the clock as INTERNAL_STEP would set it for IINT+1, then one EXTERNAL_STEP
  LAUNCH("int_clock", k_int_clock, 1, 64, h->g, h->clk, h->ia.qdis);
  enqueue_external_step(h, iext);
  // restore IINT (this call does not complete an internal step)
  DevClock c;
  HIPCHK(hipMemcpyAsync(&c, h->clk, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  c.iint -= 1;
  HIPCHK(hipMemcpyAsync(h->clk, &c, sizeof(c), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fvgpu_run(fvgpu_handle* h, int64_t nsteps, int32_t use_graph) {
  if (!use_graph || h->profiling) {
    for (int64_t s = 0; s < nsteps; ++s) {
      enqueue_internal_step(h);
      // bound the number of outstanding timing events without a per-step sync
      if (h->profiling && h->pending.size() > 20000 && flush_profile(h)) return 1;
    }
    HIPCHK(hipGetLastError());
    if (h->profiling) return flush_profile(h);
    return 0;
  }
  if (!h->gexec) {
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_internal_step(h);
    HIPCHK(hipStreamEndCapture(h->stream, &h->graph));
    HIPCHK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  }
  for (int64_t s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(h->gexec, h->stream));
  HIPCHK(hipGetLastError());
  return 0;
}

int fvgpu_get_state(fvgpu_handle* h, const char* name, double* host, int64_t len) {
  size_t l = 0;
  double* p = lookup(h, name, &l);
  if (!p) return fail(std::string("unknown state array ") + name);
  if ((size_t)len != l) return fail(std::string("length mismatch for ") + name);
  HIPCHK(hipMemcpyAsync(host, p, l * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fvgpu_put_state(fvgpu_handle* h, const char* name, const double* host, int64_t len) {
  size_t l = 0;
  double* p = lookup(h, name, &l);
  if (!p) return fail(std::string("unknown state array ") + name);
  if ((size_t)len != l) return fail(std::string("length mismatch for ") + name);
  HIPCHK(hipMemcpyAsync(p, host, l * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

namespace {
// fvgpu_ensemble_update with its staging (X5, then the member offsets) in
// `stage`, or, when `stage` is null, in a buffer the handle allocates on
// first use.
int ens_update(fvgpu_handle* h, const char* name, int32_t nrens, const int64_t* node_off, const int64_t* elem_off,
               int64_t m, int64_t n, const double* x5, double* stage) {
  const DevGrid& g = h->g;
  if (nrens < 2 || nrens > FV_ENS_MAX) return fail("ensemble size must be 2..64");
  size_t l = 0;
  double* arr = lookup(h, name ? name : "", &l);
  if (!arr) return fail(std::string("unknown state array ") + (name ? name : ""));
  const size_t m1 = g.m1, n1 = g.n1, kb = (size_t)g.KB;
  bool node;
  int nl;
  if (l == m1) { node = true; nl = 1; }
  else if (l == n1) { node = false; nl = 1; }
  else if (l == m1 * kb) { node = true; nl = (int)kb; }
  else if (l == n1 * kb) { node = false; nl = (int)kb; }
  else return fail("ensemble update: array extent not recognised");
  const int64_t len = node ? m : n;
  const int64_t* off = node ? node_off : elem_off;
  const size_t ls = node ? m1 : n1;
  if (len < 1) return fail("ensemble update: empty member slice");
  for (int i = 0; i < nrens; ++i)
    if (off[i] < 0 || (size_t)(off[i] + len + 1) > ls) return fail("ensemble update: member slice outside the array");
  if (!stage) {
    if (!h->ens_buf) {
      void* q = nullptr;
      HIPCHK(hipMalloc(&q, (FV_ENS_MAX * FV_ENS_MAX + FV_ENS_MAX) * sizeof(double)));
      h->allocs.push_back(q);
      h->ens_buf = (double*)q;
    }
    stage = h->ens_buf;
  }
  double* dx5 = stage;
  long* doff = (long*)(stage + FV_ENS_MAX * FV_ENS_MAX);
  std::vector<long> offl(off, off + nrens);
  HIPCHK(hipMemcpyAsync(dx5, x5, (size_t)nrens * nrens * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(doff, offl.data(), (size_t)nrens * sizeof(long), hipMemcpyHostToDevice, h->stream));
  const long total = (long)len * nl;
  // persistent blocks: two per CU (the 65 KB of LDS each allows two)
  const long ntiles = (total + FV_ENS_TILE - 1) / FV_ENS_TILE;
  const unsigned blocks = (unsigned)std::max(1L, std::min(ntiles, 2L * std::max(h->cu_count, 1)));
  hipLaunchKernelGGL(k_ens_update, dim3(blocks), dim3(256), 0, h->stream, arr, (const long*)doff, (long)len,
                     (long)ls, nl, (int)nrens, (const double*)dx5);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}
}  // namespace

int fvgpu_ensemble_update(fvgpu_handle* h, const char* name, int32_t nrens, const int64_t* node_off,
                          const int64_t* elem_off, int64_t m, int64_t n, const double* x5) {
  return ens_update(h, name, nrens, node_off, elem_off, m, n, x5, nullptr);
}

int fvgpu_ensemble_update_noalloc(fvgpu_handle* h, const char* name, int32_t nrens, const int64_t* node_off,
                                  const int64_t* elem_off, int64_t m, int64_t n, const double* x5) {
  if (nrens > FV_ENS_MAX) return fail("ensemble size beyond the fixed staging of 64 members");
  double* stage = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (h->device >= 0 && h->device < FV_SCRATCH_DEVICES) stage = g_scratch_ptr[h->device];
  }
  if (!stage) return fail("no staging address for this device (the handle was not created on it)");
  return ens_update(h, name, nrens, node_off, elem_off, m, n, x5, stage);
}


int fvgpu_set_forcing(fvgpu_handle* h, const char* name, const double* host, int64_t len) {
  const DevGrid& g = h->g;
  const std::string n = name ? name : "";
  const size_t n1 = g.n1, nr = (size_t)g.NUMQBC;
  if (n == "wind") {
    if (!h->md->p.wind_on) return fail("the model was created without wind");
    if ((size_t)len != 2 * n1) return fail("wind: expected UUWIND then VVWIND, 2*(N+1) values");
    std::vector<double> wx(n1, 0.0), wy(n1, 0.0);
    wind_stress(host, host + n1, n1, wx.data(), wy.data());
    HIPCHK(hipMemcpyAsync((void*)g.wsx, wx.data(), n1 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync((void*)g.wsy, wy.data(), n1 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    // the previous end speeds become the start of the coming step
    std::vector<double> old_u(n1), old_v(n1);
    HIPCHK(hipMemcpyAsync(old_u.data(), h->wu1, n1 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(old_v.data(), h->wv1, n1 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpyAsync(h->wu0, old_u.data(), n1 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->wv0, old_v.data(), n1 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->wu1, host, n1 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->wv1, host + n1, n1 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    const bool varies = std::memcmp(old_u.data(), host, n1 * sizeof(double)) != 0 ||
                        std::memcmp(old_v.data(), host + n1, n1 * sizeof(double)) != 0;
    if (varies && !h->wind_varies) {
      // the captured step gains the per-external-step stress kernel
      h->wind_varies = 1;
      drop_graphs(h);
    }
  } else if (n == "qdis" || n == "tdis" || n == "sdis") {
    if ((size_t)len != nr) return fail(n + ": expected one value per river (NUMQBC)");
    const double* dst = n == "qdis" ? g.qdis0 : n == "tdis" ? g.tdis : g.sdis;
    if (nr) HIPCHK(hipMemcpyAsync((void*)dst, host, nr * sizeof(double), hipMemcpyHostToDevice, h->stream));
  } else {
    return fail("unknown forcing " + n + " (wind, qdis, tdis, sdis)");
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fvgpu_diagnostics(fvgpu_handle* h, fv_diag* out);  // fvgpu_diag.hip

int fvgpu_kernel_time(fvgpu_handle* h, const char* name, double* avg_us, int64_t* count) {
  if (flush_profile(h)) return 1;
  auto it = h->stats.find(name);
  if (it == h->stats.end()) return fail(std::string("no timing for ") + name);
  *count = it->second.count;
  *avg_us = it->second.count ? 1000.0 * it->second.total_ms / it->second.count : 0.0;
  return 0;
}

int fvgpu_kernel_stat(fvgpu_handle* h, int32_t idx, const char** name, double* avg_us, int64_t* count) {
  if (flush_profile(h)) return 1;
  if (idx < 0 || idx >= (int32_t)h->stats.size()) return fail("kernel statistics index out of range");
  auto it = h->stats.begin();
  std::advance(it, idx);
  *name = it->first.c_str();
  *count = it->second.count;
  *avg_us = it->second.count ? 1000.0 * it->second.total_ms / it->second.count : 0.0;
  return 0;
}

int fvgpu_set_profiling(fvgpu_handle* h, int32_t on) {
  if (h->profiling && flush_profile(h)) return 1;
  h->profiling = on != 0;
  if (h->profiling) { h->stats.clear(); h->profile_seen = 0; }   // statistics stay readable after switching off
  return 0;
}

int fvgpu_profile_every(fvgpu_handle* h, int32_t every) {
  if (every < 1) return fail("fvgpu_profile_every: every must be >= 1");
  h->profile_every = every;
  return 0;
}

int fvgpu_profile_only(fvgpu_handle* h, const char* name) {
  const std::string n = name ? name : "";
  if (n != h->profile_only) drop_graphs(h);
  h->profile_only = n;
  return 0;
}

int fvgpu_stream_copy(fvgpu_handle* h, int64_t n, int32_t reps, double* avg_us) {
  double *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc((void**)&a, n * sizeof(double)));
  HIPCHK(hipMalloc((void**)&b, n * sizeof(double)));
  hipLaunchKernelGGL(k_zero, dim3(nb(n)), dim3(256), 0, h->stream, a, (long)n);
  hipLaunchKernelGGL(k_copy, dim3(nb(n)), dim3(256), 0, h->stream, b, a, (long)n);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy, dim3(nb(n)), dim3(256), 0, h->stream, b, a, (long)n);
  HIPCHK(hipEventRecord(e1, h->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *avg_us = 1000.0 * ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipFree(a));
  HIPCHK(hipFree(b));
  return 0;
}

int fvgpu_synchronize(fvgpu_handle* h) {
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

void fvgpu_destroy(fvgpu_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  drop_graphs(h);
  for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"

// --------------------------------------------------------- diagnostics
// Deterministic two-level reduction with Neumaier compensation for the FP64
// budgets; flags for non-finite values and the minimum depth.
namespace {
struct DiagAcc { double vol, vc, salt, sc, heat, hc, dmin; int bad; };

__device__ void neumaier(double& s, double& c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}

__global__ void k_diag_partial(DevGrid g, IntArgs a, DiagAcc* part) {
  __shared__ DiagAcc sh[256];
  const int tid = threadIdx.x;
  DiagAcc acc{0, 0, 0, 0, 0, 0, 1e300, 0};
  for (long i = (long)blockIdx.x * blockDim.x + tid + 1; i <= g.M; i += (long)gridDim.x * blockDim.x) {
    const double d = a.d[i];
    neumaier(acc.vol, acc.vc, g.art1[i] * d);
    acc.dmin = fmin(acc.dmin, d);
    if (!isfinite(a.el[i])) acc.bad = 1;
    for (int k = 1; k <= g.KBM1; ++k) {
      const double w = g.art1[i] * d * DZ((int)i, k);
      const double s = a.s1[NI(g, (int)i, k)], t = a.t1[NI(g, (int)i, k)];
      neumaier(acc.salt, acc.sc, w * s);
      neumaier(acc.heat, acc.hc, w * t);
      if (!isfinite(s) || !isfinite(t)) acc.bad = 1;
    }
  }
  for (long i = (long)blockIdx.x * blockDim.x + tid + 1; i <= g.N; i += (long)gridDim.x * blockDim.x) {
    if (!isfinite(a.ua[i]) || !isfinite(a.va[i])) acc.bad = 1;
    for (int k = 1; k <= g.KBM1; ++k)
      if (!isfinite(a.u[EI(g, (int)i, k)]) || !isfinite(a.v[EI(g, (int)i, k)])) acc.bad = 1;
  }
  sh[tid] = acc;
  __syncthreads();
  if (tid == 0) {
    DiagAcc r = sh[0];
    for (int j = 1; j < blockDim.x; ++j) {
      neumaier(r.vol, r.vc, sh[j].vol); r.vc += sh[j].vc;
      neumaier(r.salt, r.sc, sh[j].salt); r.sc += sh[j].sc;
      neumaier(r.heat, r.hc, sh[j].heat); r.hc += sh[j].hc;
      r.dmin = fmin(r.dmin, sh[j].dmin);
      r.bad |= sh[j].bad;
    }
    part[blockIdx.x] = r;
  }
}
}  // namespace

extern "C" int fvgpu_diagnostics(fvgpu_handle* h, fv_diag* out) {
  const int nblk = 256;
  DiagAcc* part = (DiagAcc*)h->diag_part;
  hipLaunchKernelGGL(k_diag_partial, dim3(nblk), dim3(256), 0, h->stream, h->g, h->ia, part);
  std::vector<DiagAcc> hp(nblk);
  HIPCHK(hipMemcpyAsync(hp.data(), part, sizeof(DiagAcc) * nblk, hipMemcpyDeviceToHost, h->stream));
  DevClock c;
  HIPCHK(hipMemcpyAsync(&c, h->clk, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  double vol = 0, vc = 0, salt = 0, sc = 0, heat = 0, hc = 0, dmin = 1e300;
  int bad = 0;
  auto nm = [](double& s, double& cc, double x) {
    const double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) cc += (s - t) + x; else cc += (x - t) + s;
    s = t;
  };
  for (const DiagAcc& r : hp) {
    nm(vol, vc, r.vol); vc += r.vc;
    nm(salt, sc, r.salt); sc += r.sc;
    nm(heat, hc, r.heat); hc += r.hc;
    dmin = std::min(dmin, r.dmin);
    bad |= r.bad;
  }
  out->nonfinite = bad;
  out->depth_fail = c.depth_fail;
  out->min_depth = dmin;
  out->volume = vol + vc;
  out->salt = salt + sc;
  out->heat = heat + hc;
  out->iint = c.iint;
  return 0;
}

// ------------------------------------------------------------ learned closure
extern "C" {

int fvgpu_set_closure(fvgpu_handle* h, int32_t mode, const float* params) {
  if (mode != 0 && mode != 1) return fail("closure mode must be 0 (MY2.5) or 1 (learned)");
  if (mode == 1) {
    if (!params) return fail("learned closure needs parameters");
    if (!h->md->p.vertical_closure) return fail("the learned closure replaces VERTICAL_MIXING_TYPE='closure'");
    const DevGrid& g = h->g;
    const size_t ns = (size_t)g.M * (g.KBM1 - 1);
    if (!h->cl_x) {
      if (dalloc(h, &h->cl_x, ns * MLP_IN) || dalloc(h, &h->cl_yp, ns * MLP_OUT) ||
          dalloc(h, &h->cl_w, 1) || dalloc(h, &h->cl_flag, (size_t)g.m1) || dalloc(h, &h->cl_nfall, 1))
        return 1;
    }
    float* dp = nullptr;
    HIPCHK(hipMalloc((void**)&dp, MLP_NPARAM * sizeof(float)));
    HIPCHK(hipMemcpy(dp, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, h->stream, dp, h->cl_w);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(dp));
  }
  h->closure_mode = mode;
  drop_graphs(h);   // the step changed
  return 0;
}

int fvgpu_closure_fallbacks(fvgpu_handle* h, int64_t* fallback_columns, int64_t* columns) {
  int n = 0;
  if (h->cl_nfall) {
    HIPCHK(hipMemcpyAsync(&n, h->cl_nfall, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  *fallback_columns = n;
  *columns = h->g.M;
  return 0;
}

int fvgpu_closure_samples(fvgpu_handle* h, float* x, float* y, int64_t max_samples, int64_t* n) {
  const DevGrid& g = h->g;
  const size_t ns = (size_t)g.M * (g.KBM1 - 1);
  *n = (int64_t)ns;
  if (!x || !y) return 0;   // size query
  if (max_samples < (int64_t)ns) return fail("sample buffer too small");
  float *dx = nullptr, *dy = nullptr;
  HIPCHK(hipMalloc((void**)&dx, ns * MLP_IN * sizeof(float)));
  HIPCHK(hipMalloc((void**)&dy, ns * MLP_OUT * sizeof(float)));
  IntArgs a = h->ia;
#define CLD_CALL(RM) hipLaunchKernelGGL(k_closure_features<RM>, dim3(nb(g.M)), dim3(256), 0, h->stream, g, a, dx, dy)
  FV_RMAX_DISPATCH(g.ve_max, CLD_CALL);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(x, dx, ns * MLP_IN * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(y, dy, ns * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipFree(dx));
  HIPCHK(hipFree(dy));
  return 0;
}

}  // extern "C"

// The trainer: device-resident samples, FP32 master weights and AdamW state.
struct fvmlp {
  int64_t n = 0;
  int bmax = 0;
  int step = 0;
  hipStream_t stream = nullptr;
  float *X = nullptr, *Y = nullptr, *P = nullptr, *G = nullptr, *Mo = nullptr, *Ve = nullptr;
  float *H1 = nullptr, *H2 = nullptr, *D1 = nullptr, *D2 = nullptr, *DO = nullptr, *loss = nullptr, *lsum = nullptr;
  float* Yp = nullptr;
  int* idx = nullptr;
  MlpPacked* packed = nullptr;
  std::vector<void*> allocs;
};

namespace {
template <typename T>
int malloc_mlp(fvmlp* t, T** p, size_t n) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
  HIPCHK(hipMemset(q, 0, (n ? n : 1) * sizeof(T)));
  t->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

// Minibatch of one training step: fwd/bwd, weight gradients, loss sum.
int mlp_batch(fvmlp* t, const int* didx, int B, float* loss_out) {
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3((B + 255) / 256), dim3(256), 0, t->stream, t->P, t->X, t->Y, didx, B, t->H1,
                     t->H2, t->D1, t->D2, t->DO, t->loss);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, t->stream, t->X, didx, B, t->H1, t->H2,
                     t->D1, t->D2, t->DO, t->G);
  hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(256), 0, t->stream, t->loss, (long)B, t->lsum);
  HIPCHK(hipGetLastError());
  if (loss_out) {
    float l = 0.0f;
    HIPCHK(hipMemcpyAsync(&l, t->lsum, sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    *loss_out = l / (float)B;
  }
  return 0;
}
}  // namespace

extern "C" {

int fvmlp_create(int64_t n, const float* x, const float* y, int32_t bmax, fvmlp** out) {
  if (n < 1 || bmax < 1 || !x || !y || !out) return fail("fvmlp_create: bad arguments");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail("no HIP device");
  fvmlp* t = new fvmlp();
  t->n = n;
  t->bmax = bmax;
  const size_t bb = (size_t)bmax;
  if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess ||
      malloc_mlp(t, &t->X, (size_t)n * MLP_IN) || malloc_mlp(t, &t->Y, (size_t)n * MLP_OUT) ||
      malloc_mlp(t, &t->P, MLP_NPARAM) || malloc_mlp(t, &t->G, MLP_NPARAM) || malloc_mlp(t, &t->Mo, MLP_NPARAM) ||
      malloc_mlp(t, &t->Ve, MLP_NPARAM) || malloc_mlp(t, &t->H1, bb * MLP_H) || malloc_mlp(t, &t->H2, bb * MLP_H) ||
      malloc_mlp(t, &t->D1, bb * MLP_H) || malloc_mlp(t, &t->D2, bb * MLP_H) || malloc_mlp(t, &t->DO, bb * MLP_OUT) ||
      malloc_mlp(t, &t->loss, bb) || malloc_mlp(t, &t->lsum, 1) || malloc_mlp(t, &t->idx, bb) ||
      malloc_mlp(t, &t->Yp, (size_t)n * MLP_OUT) || malloc_mlp(t, &t->packed, 1)) {
    fvmlp_destroy(t);
    return 1;
  }
  HIPCHK(hipMemcpy(t->X, x, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(t->Y, y, (size_t)n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice));
  *out = t;
  return 0;
}

int fvmlp_set_params(fvmlp* t, const float* p) {
  HIPCHK(hipMemcpy(t->P, p, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(t->Mo, 0, MLP_NPARAM * sizeof(float)));
  HIPCHK(hipMemset(t->Ve, 0, MLP_NPARAM * sizeof(float)));
  t->step = 0;
  return 0;
}

int fvmlp_get_params(fvmlp* t, float* p) {
  HIPCHK(hipStreamSynchronize(t->stream));
  HIPCHK(hipMemcpy(p, t->P, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int32_t fvmlp_nparams(void) { return MLP_NPARAM; }

int fvmlp_grad(fvmlp* t, const int32_t* idx, int32_t B, float* grad, float* loss) {
  if (B < 1 || B > t->bmax) return fail("fvmlp_grad: batch size outside [1, bmax]");
  const int* didx = nullptr;
  if (idx) {
    for (int b = 0; b < B; ++b)
      if (idx[b] < 0 || idx[b] >= t->n) return fail("fvmlp_grad: sample index out of range");
    HIPCHK(hipMemcpy(t->idx, idx, (size_t)B * sizeof(int), hipMemcpyHostToDevice));
    didx = t->idx;
  } else if (B > t->n) {
    return fail("fvmlp_grad: batch larger than the sample set");
  }
  if (mlp_batch(t, didx, B, loss)) return 1;
  HIPCHK(hipMemcpy(grad, t->G, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int fvmlp_train(fvmlp* t, int32_t steps, int32_t batch, int64_t ntrain, float lr, float wd, uint64_t seed,
                float* last_loss) {
  if (batch < 1 || batch > t->bmax) return fail("fvmlp_train: batch size outside [1, bmax]");
  if (ntrain < 1 || ntrain > t->n) return fail("fvmlp_train: ntrain outside [1, n]");
  std::vector<int> hidx((size_t)batch);
  uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
  float l = 0.0f;
  for (int it = 0; it < steps; ++it) {
    for (int b = 0; b < batch; ++b) {   // seeded minibatch (PCG-style LCG, high bits)
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      hidx[b] = (int)((st >> 33) % (uint64_t)ntrain);
    }
    HIPCHK(hipMemcpyAsync(t->idx, hidx.data(), (size_t)batch * sizeof(int), hipMemcpyHostToDevice, t->stream));
    if (mlp_batch(t, t->idx, batch, it + 1 == steps ? &l : nullptr)) return 1;
    t->step += 1;
    hipLaunchKernelGGL(k_adamw, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, t->stream, t->P, t->G, t->Mo, t->Ve,
                       t->step, lr, wd, 0.9f, 0.999f, 1e-8f);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->stream));   // the host index buffer is reused
  }
  if (last_loss) *last_loss = l;
  return 0;
}

int fvmlp_infer(fvmlp* t, int64_t lo, int64_t n, int32_t use_mfma, float* y) {
  if (lo < 0 || n < 1 || lo + n > t->n) return fail("fvmlp_infer: range outside the sample set");
  if (use_mfma) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, t->stream, t->P, t->packed);
    const long groups = (n + 31) / 32;
    hipLaunchKernelGGL(k_mlp_infer_mfma, dim3((unsigned)std::min<long>(groups, 8192)), dim3(64), 0, t->stream,
                       t->packed, t->X + lo * MLP_IN, (long)n, t->Yp);
  } else {
    hipLaunchKernelGGL(k_mlp_infer_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, t->stream, t->P,
                       t->X + lo * MLP_IN, (long)n, t->Yp);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, t->Yp, (size_t)n * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  return 0;
}

void fvmlp_destroy(fvmlp* t) {
  if (!t) return;
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  for (void* p : t->allocs) (void)hipFree(p);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
}

}  // extern "C"

// ------------------------------------------------------------ autotuner
extern "C" {

int fvgpu_set_block(fvgpu_handle* h, const char* family, int32_t block) {
  if (block != 64 && block != 128 && block != 256) return fail("workgroup size must be 64, 128 or 256");
  h->tune[family] = block;
  drop_graphs(h);
  return 0;
}

int fvgpu_set_concurrency(fvgpu_handle* h, int32_t mode) {
  if (mode < 0 || mode > 3) return fail("concurrency must be 0..3");
  h->concurrent = mode;
  drop_graphs(h);
  return 0;
}

int fvgpu_autotune(fvgpu_handle* h, int32_t steps, char* report, int64_t report_len) {
  static const char* fams[] = {"ext_elevation", "advave_edge", "ext_momentum", "cvrec_ts", "cvrec_q", "cvflux_ts",
                               "cvflux_q", "adv_ts", "adv_q",
                               "viscofh", "vdif_q", "vdif_ts", "adjust1_brough", "adjust2", "baropg",
                               "uv_finish", "copy", "closure_features", "closure_apply"};
  static const int cands[] = {64, 128, 256};
  if (steps < 1) return fail("autotune needs at least one step per candidate");
  // save the state; every candidate runs real steps
  size_t tot = 0;
  for (auto& b : h->rw) tot += (b.second + 255) & ~(size_t)255;
  char* arena = nullptr;
  HIPCHK(hipMalloc((void**)&arena, tot + sizeof(DevClock) + 256));
  size_t off = 0;
  HIPCHK(hipStreamSynchronize(h->stream));
  for (auto& b : h->rw) {
    HIPCHK(hipMemcpyAsync(arena + off, b.first, b.second, hipMemcpyDeviceToDevice, h->stream));
    off += (b.second + 255) & ~(size_t)255;
  }
  HIPCHK(hipMemcpyAsync(arena + off, h->clk, sizeof(DevClock), hipMemcpyDeviceToDevice, h->stream));
  const bool prof = h->profiling;
  const std::string only = h->profile_only;
  const int conc = h->concurrent, every = h->profile_every;
  h->concurrent = 0;   // each family timed alone, at every launch
  h->profile_every = 1;
  std::string rep;
  int rc = 0;
  for (const char* f : fams) {
    double best = 1e30;
    int bestc = 256;
    bool seen = false;
    for (int c : cands) {
      h->tune[f] = c;
      h->profiling = true;
      h->profile_only = f;
      h->stats.clear();
      for (int st = 0; st < steps && rc == 0; ++st) enqueue_internal_step(h);
      if (rc == 0) rc = flush_profile(h);
      auto it = h->stats.find(f);
      if (it == h->stats.end() || it->second.count == 0) break;   // family not in this step
      seen = true;
      const double t = it->second.total_ms / it->second.count;
      if (t < best) { best = t; bestc = c; }
    }
    if (seen) {
      h->tune[f] = bestc;
      char buf[96];
      std::snprintf(buf, sizeof(buf), "%s%s=%d:%.1fus", rep.empty() ? "" : ";", f, bestc, 1000.0 * best);
      rep += buf;
    } else {
      h->tune.erase(f);
    }
  }
  h->profiling = prof;
  h->profile_only = only;
  h->concurrent = conc;
  h->profile_every = every;
  h->stats.clear();
  off = 0;
  for (auto& b : h->rw) {
    HIPCHK(hipMemcpyAsync(b.first, arena + off, b.second, hipMemcpyDeviceToDevice, h->stream));
    off += (b.second + 255) & ~(size_t)255;
  }
  HIPCHK(hipMemcpyAsync(h->clk, arena + off, sizeof(DevClock), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipFree(arena));
  drop_graphs(h);
  if (report && report_len > 0) {
    std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return rc;
}

}  // extern "C"
