// fvgrid.cpp -- host side of the mode-split time step: seeded synthetic
// cases, grid topology and metrics, and the cold-start state.
//
// The metric definitions follow the reference's GRID_METRICS sequence
// (FVCOM_source/grid_metrics.F:55-134):
//   TRIANGLE_GRID_EDGE  FVCOM_source/tge.F:29-867
//   CELL_AREA           FVCOM_source/cell_area.F:40-352
//   SHAPE_COEF_GCN      FVCOM_source/shape_coef_gcn.F:29-368
//   EDGE_LEN            FVCOM_source/edge_len.F:29-241
//   DEPTH_GRADIENT      FVCOM_source/depth_grad.F
//   SETUP_OBC           FVCOM_source/mod_obcs.F:868-1173
// and the setup of sigma levels (mod_setup.F:772-1091), water depth
// (mod_startup.F:1408-1495), rivers (mod_force.F:3470-3511) and the cold
// start (mod_startup.F:1062-1181, 1332-1408).  The O(M*NT) searches of the
// reference are replaced by node->cell lists that visit cells in the same
// increasing order, so every array comes out identical.
#include "../../include/fvgpu.h"

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <array>
#include <type_traits>

namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return 1; }

// SplitMix64 / xoshiro256** -- seeded, platform independent.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed ? seed : 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < 4; ++i) {
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
  double normal() {                                                          // Box-Muller
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

// Gaussian random field with covariance exp(-r^2/(2 L^2)) by random Fourier
// modes: g(x) = sqrt(2/Nm) sum cos(k_m.x + phi_m), k_m ~ N(0, 1/L^2).
struct RandomField {
  std::vector<double> kx, ky, ph;
  RandomField(Rng& r, double corr, int nm = 192) {
    kx.resize(nm); ky.resize(nm); ph.resize(nm);
    for (int i = 0; i < nm; ++i) {
      kx[i] = r.normal() / corr; ky[i] = r.normal() / corr; ph[i] = 2.0 * M_PI * r.uniform();
    }
  }
  double operator()(double x, double y) const {
    double s = 0.0;
    for (size_t i = 0; i < kx.size(); ++i) s += std::cos(kx[i] * x + ky[i] * y + ph[i]);
    return s * std::sqrt(2.0 / (double)kx.size());
  }
};

inline int nxt(int j) { return j % 3 + 1; }        // JJ+1-INT((JJ+1)/4)*3
inline int nxt2(int j) { return (j + 1) % 3 + 1; } // JJ+2-INT((JJ+2)/4)*3

struct Builder {
  fv_model* md;
  int M, N;
  // 1-based accessors into component-major arrays
  int32_t& NV(int i, int j) { return md->nv[(size_t)(j - 1) * (N + 1) + i]; }
  int32_t& NBE(int i, int j) { return md->nbe[(size_t)(j - 1) * (N + 1) + i]; }
  double& A1U(int i, int j) { return md->a1u[(size_t)(j - 1) * (N + 1) + i]; }
  double& A2U(int i, int j) { return md->a2u[(size_t)(j - 1) * (N + 1) + i]; }
  double& AWX(int i, int j) { return md->awx[(size_t)(j - 1) * (N + 1) + i]; }
  double& AWY(int i, int j) { return md->awy[(size_t)(j - 1) * (N + 1) + i]; }
};

void* xalloc(size_t bytes) {
  void* p = std::calloc(bytes ? bytes : 8, 1);
  return p;
}

}  // namespace

extern "C" const char* fv_last_error(void) { return g_err.c_str(); }

extern "C" int64_t fv_extent_len(const fv_model* md, int32_t ext) {
  const int64_t m1 = md->m + 1, n1 = md->n + 1;
  switch (ext) {
    case FVX_M1: return m1;
    case FVX_M2: return md->m + 2;
    case FVX_N1: return n1;
    case FVX_N1x3: return 3 * n1;
    case FVX_N1x4: return 4 * n1;
    case FVX_NE1: return md->ne + 1;
    case FVX_NE1x2: return 2 * (int64_t)(md->ne + 1);
    case FVX_CV1: return md->ncv + 1;
    case FVX_CV1x2: return 2 * (int64_t)(md->ncv + 1);
    case FVX_M1K: return m1 * md->kb;
    case FVX_N1K: return n1 * md->kb;
    case FVX_RSN: return md->nring_sn;
    case FVX_RVE: return md->nring_ve;
    case FVX_OBC: return md->iobcn;
    case FVX_OBCx2: return 2 * (int64_t)md->iobcn;
    case FVX_OBCxT: return (int64_t)md->iobcn * md->ntide;
    case FVX_TIDE: return md->ntide;
    case FVX_RIV: return md->numqbc;
    case FVX_RIVxK: return (int64_t)md->numqbc * md->kbm1;
    case FVX_RIVx2: return 2 * (int64_t)md->numqbc;
  }
  return -1;
}

namespace {
struct FieldInfo { const char* name; int ext; int dtype; size_t off; };
#define FV_FIELD_INFO(T, NAME, EXT) \
  {#NAME, EXT, (sizeof(T) == 8 ? 0 : 1), offsetof(fv_model, NAME)},
const FieldInfo kFields[] = { FV_ALL_FIELDS(FV_FIELD_INFO) };
#undef FV_FIELD_INFO
const int kNFields = (int)(sizeof(kFields) / sizeof(kFields[0]));

void alloc_all(fv_model* md) {
  for (int i = 0; i < kNFields; ++i) {
    int64_t len = fv_extent_len(md, kFields[i].ext);
    size_t esz = kFields[i].dtype == 0 ? 8 : 4;
    void** slot = (void**)((char*)md + kFields[i].off);
    std::free(*slot);
    *slot = xalloc((size_t)len * esz);
  }
}
}  // namespace

extern "C" int32_t fv_model_nfields(void) { return kNFields; }
extern "C" const char* fv_model_field_name(int32_t i) {
  return (i >= 0 && i < kNFields) ? kFields[i].name : nullptr;
}

extern "C" int fv_model_field(const fv_model* md, const char* name, void** ptr,
                              int64_t* len, int32_t* dtype) {
  for (int i = 0; i < kNFields; ++i) {
    if (std::strcmp(kFields[i].name, name) == 0) {
      *ptr = *(void* const*)((const char*)md + kFields[i].off);
      *len = fv_extent_len(md, kFields[i].ext);
      *dtype = kFields[i].dtype;
      return 0;
    }
  }
  return fail(std::string("unknown field ") + name);
}

extern "C" void fv_model_free(fv_model* md) {
  if (!md) return;
  for (int i = 0; i < kNFields; ++i) std::free(*(void**)((char*)md + kFields[i].off));
  std::free(md);
}

// --------------------------------------------------------------------------
// Grid metrics from (VX, VY, NV, H, OBC list).
// --------------------------------------------------------------------------
namespace {

int build_metrics(fv_model* md, const double* x, const double* y, const int32_t* nvrows,
                  const double* hin, int32_t iobcn, const int32_t* obc) {
  const int M = md->m, N = md->n, KB = md->kb;
  Builder B{md, M, N};
  // Provisional sizes so that the node/element/sigma arrays exist; the edge,
  // CV and ring arrays are re-allocated once their counts are known.
  md->ne = 0; md->ncv = 0; md->nring_sn = 0; md->nring_ve = 0; md->iobcn = iobcn;
  alloc_all(md);

  // SETUP_CENTER_COORDS (mod_setup.F:62-110): shift to the lower-left corner.
  double vxmin = 1e300, vymin = 1e300;
  for (int i = 1; i <= M; ++i) { vxmin = std::min(vxmin, x[i - 1]); vymin = std::min(vymin, y[i - 1]); }
  for (int i = 1; i <= M; ++i) { md->vx[i] = x[i - 1] - vxmin; md->vy[i] = y[i - 1] - vymin; md->h[i] = hin[i - 1]; }
  for (int i = 1; i <= N; ++i)
    for (int j = 1; j <= 3; ++j) {
      int v = nvrows[(size_t)(i - 1) * 3 + (j - 1)];
      if (v < 1 || v > M) return fail("NV out of range");
      B.NV(i, j) = v;
    }
  const double one_third = 1.0 / 3.0;
  for (int i = 1; i <= N; ++i) {
    md->xc[i] = one_third * (md->vx[B.NV(i, 1)] + md->vx[B.NV(i, 2)] + md->vx[B.NV(i, 3)]);
    md->yc[i] = one_third * (md->vy[B.NV(i, 1)] + md->vy[B.NV(i, 2)] + md->vy[B.NV(i, 3)]);
  }

  // ---------------- TRIANGLE_GRID_EDGE (tge.F) ----------------
  std::vector<int> cellcnt(M + 1, 0), celloff(M + 2, 0), cells;
  for (int i = 1; i <= N; ++i) for (int j = 1; j <= 3; ++j) cellcnt[B.NV(i, j)]++;
  for (int i = 1; i <= M; ++i) celloff[i + 1] = celloff[i] + cellcnt[i];
  cells.resize(celloff[M + 1]);
  {
    std::vector<int> fill(M + 1, 0);
    for (int i = 1; i <= N; ++i)
      for (int j = 1; j <= 3; ++j) { int n = B.NV(i, j); cells[celloff[n] + fill[n]++] = i; }
  }
  auto shares = [&](int a, int b, int self) {  // last common cell of nodes a,b other than self
    int r = 0;
    for (int p = celloff[a]; p < celloff[a + 1]; ++p)
      for (int q = celloff[b]; q < celloff[b + 1]; ++q)
        if (cells[p] == cells[q] && cells[p] != self) r = cells[p];
    return r;
  };
  for (int i = 1; i <= N; ++i) {
    int n1 = B.NV(i, 1), n2 = B.NV(i, 2), n3 = B.NV(i, 3);
    B.NBE(i, 3) = shares(n1, n2, i);
    B.NBE(i, 1) = shares(n2, n3, i);
    B.NBE(i, 2) = shares(n1, n3, i);
    if (B.NBE(i, 1) + B.NBE(i, 2) + B.NBE(i, 3) == 0) return fail("isolated element");
  }
  for (int i = 1; i <= N; ++i) {
    if (std::min(B.NBE(i, 1), std::min(B.NBE(i, 2), B.NBE(i, 3))) == 0) {
      md->isbce[i] = 1;
      if (B.NBE(i, 1) == 0) { md->isonb[B.NV(i, 2)] = 1; md->isonb[B.NV(i, 3)] = 1; }
      if (B.NBE(i, 2) == 0) { md->isonb[B.NV(i, 1)] = 1; md->isonb[B.NV(i, 3)] = 1; }
      if (B.NBE(i, 3) == 0) { md->isonb[B.NV(i, 1)] = 1; md->isonb[B.NV(i, 2)] = 1; }
    }
  }
  // NBVE/NBVT: cells of each node in increasing element order, then ordered
  // as a ring walk; NBSN the surrounding node ring.
  std::vector<std::vector<int>> nbve(M + 1), nbvt(M + 1), nbsn(M + 1);
  for (int i = 1; i <= M; ++i) {
    int ntve = cellcnt[i];
    std::vector<int> ve(ntve), vt(ntve);
    for (int k = 0; k < ntve; ++k) {
      int c = cells[celloff[i] + k];  // increasing element index
      ve[k] = c;
      for (int j = 1; j <= 3; ++j) if (B.NV(c, j) == i) vt[k] = j;
    }
    md->ntve[i] = ntve;
    auto locidx = [&](int c) { int r = 0; for (int j = 1; j <= 3; ++j) if (B.NV(c, j) == i) r = j; return r; };
    if (md->isonb[i] == 0) {
      std::vector<int> e(ntve + 1), t(ntve + 1);
      e[0] = ve[0]; t[0] = vt[0];
      for (int j = 1; j <= ntve; ++j) {
        int ii = e[j - 1], jj = t[j - 1];
        e[j] = B.NBE(ii, nxt(jj));
        if (e[j] == 0) return fail("broken interior ring");
        t[j] = locidx(e[j]);
      }
      if (e[0] != e[ntve] || t[0] != t[ntve]) return fail("NBVE ring not closed");
      nbve[i].assign(e.begin(), e.begin() + ntve);
      nbvt[i].assign(t.begin(), t.begin() + ntve);
      std::vector<int> sn(ntve + 1);
      for (int j = 0; j < ntve; ++j) sn[j] = B.NV(nbve[i][j], nxt(nbvt[i][j]));
      sn[ntve] = sn[0];
      nbsn[i] = sn;
    } else {
      int jjb = 0, se = 0, st = 0;
      for (int k = 0; k < ntve; ++k) {
        int jj = vt[k];
        if (B.NBE(ve[k], nxt2(jj)) == 0) { jjb++; se = ve[k]; st = jj; }
      }
      if (jjb != 1) return fail("ERROR IN TGE DETERMINING ISONB");
      std::vector<int> e(ntve), t(ntve);
      e[0] = se; t[0] = st;
      for (int j = 1; j < ntve; ++j) {
        int ii = e[j - 1], jj = t[j - 1];
        e[j] = B.NBE(ii, nxt(jj));
        if (e[j] == 0) return fail("broken boundary ring");
        t[j] = locidx(e[j]);
      }
      nbve[i] = e; nbvt[i] = t;
      std::vector<int> sn;
      sn.push_back(i);
      for (int j = 0; j < ntve; ++j) sn.push_back(B.NV(e[j], nxt(t[j])));
      sn.push_back(B.NV(e[ntve - 1], nxt2(t[ntve - 1])));
      sn.push_back(i);
      nbsn[i] = sn;
    }
    md->ntsn[i] = (int)nbsn[i].size();
  }
  // Edges in the reference's (element, local side) discovery order.
  std::vector<int> iec1, iec2, ien1, ien2;
  {
    std::vector<char> iset((size_t)(N + 1) * 3, 0);
    for (int i = 1; i <= N; ++i)
      for (int j = 1; j <= 3; ++j) {
        if (iset[(size_t)i * 3 + j - 1]) continue;
        int iney = B.NBE(i, j);
        iset[(size_t)i * 3 + j - 1] = 1;
        if (iney) for (int jn = 1; jn <= 3; ++jn) if (B.NBE(iney, jn) == i) iset[(size_t)iney * 3 + jn - 1] = 1;
        iec1.push_back(i); iec2.push_back(iney);
        ien1.push_back(B.NV(i, nxt(j))); ien2.push_back(B.NV(i, nxt2(j)));
      }
  }
  const int NE = (int)iec1.size();
  // Open boundary nodes: ISONB = 2; open-boundary element rule.
  for (int k = 0; k < iobcn; ++k) {
    if (obc[k] < 1 || obc[k] > M) return fail("OBC node out of range");
    md->isonb[obc[k]] = 2;
  }
  for (int i = 1; i <= N; ++i) {
    int s = md->isonb[B.NV(i, 1)] + md->isonb[B.NV(i, 2)] + md->isonb[B.NV(i, 3)];
    if (s == 4) md->isbce[i] = 2;
    else if (s > 4) return fail("boundary cell with two open sides or open+solid (tge.F check)");
  }
  for (int i = 1; i <= N; ++i) {
    int a = B.NBE(i, 1), b = B.NBE(i, 2), c = B.NBE(i, 3);
    if (md->isbce[i] != 2 && (a + b + c == 0 || a + b == 0 || b + c == 0 || a + c == 0)) md->isbce[i] = 3;
  }

  // Sizes are known: allocate edge/CV/ring arrays (keeping the node and
  // element arrays filled so far).
  int nsn = 0, nve = 0;
  for (int i = 1; i <= M; ++i) { nsn += (int)nbsn[i].size(); nve += (int)nbve[i].size(); }
  {
    fv_model tmp = *md;
    md->ne = NE; md->ncv = 3 * N; md->ncv_i = 3 * N; md->nring_sn = nsn; md->nring_ve = nve;
    // re-allocate only the arrays whose extent depends on NE/NCV/rings
    for (int f = 0; f < kNFields; ++f) {
      int ext = kFields[f].ext;
      if (ext == FVX_NE1 || ext == FVX_NE1x2 || ext == FVX_CV1 || ext == FVX_CV1x2 ||
          ext == FVX_RSN || ext == FVX_RVE) {
        void** slot = (void**)((char*)md + kFields[f].off);
        std::free(*slot);
        *slot = xalloc((size_t)fv_extent_len(md, ext) * (kFields[f].dtype == 0 ? 8 : 4));
      }
    }
    (void)tmp;
  }
  int32_t* IEC = md->iec; int32_t* IEN = md->ienode;
  const size_t ne1 = NE + 1;
  for (int e = 1; e <= NE; ++e) {
    IEC[e] = iec1[e - 1]; IEC[ne1 + e] = iec2[e - 1];
    IEN[e] = ien1[e - 1]; IEN[ne1 + e] = ien2[e - 1];
    md->isbc[e] = (iec1[e - 1] == 0 || iec2[e - 1] == 0) ? 1 : 0;
    double x1 = md->vx[IEN[e]], x2 = md->vx[IEN[ne1 + e]], y1 = md->vy[IEN[e]], y2 = md->vy[IEN[ne1 + e]];
    md->dltxc[e] = x2 - x1;
    md->dltyc[e] = y2 - y1;
    md->xijc[e] = (x1 + x2) / 2.0;
    md->yijc[e] = (y1 + y2) / 2.0;
  }
  // Control-volume sub-edges.
  const size_t cv1 = md->ncv + 1;
  std::vector<double> xije1(cv1), yije1(cv1), xije2(cv1), yije2(cv1);
  int npt = 0;
  for (int e = 1; e <= NE; ++e) {
    int ia = IEC[e], ib = IEC[ne1 + e];
    if (ia == 0) return fail("IEC(I,1)==0");
    auto add = [&](int tri, int na, int nb) {
      ++npt;
      xije1[npt] = md->xc[tri]; yije1[npt] = md->yc[tri];
      xije2[npt] = md->xijc[e]; yije2[npt] = md->yijc[e];
      md->niec[npt] = na; md->niec[cv1 + npt] = nb; md->ntrg[npt] = tri;
      md->dltxe[npt] = xije2[npt] - xije1[npt];
      md->dltye[npt] = yije2[npt] - yije1[npt];
    };
    if (md->isbc[e] == 0) { add(ia, IEN[e], IEN[ne1 + e]); add(ib, IEN[ne1 + e], IEN[e]); }
    else add(ia, IEN[e], IEN[ne1 + e]);
  }
  if (npt != 3 * N) return fail("NCV IS NOT CORRECT");
  for (int i = 1; i <= N; ++i) md->epor[i] = (md->isbce[i] == 2) ? 0.0 : 1.0;
  md->epor[0] = 1.0;
  // rings to CSR
  md->nbsn_off[1] = 0; md->nbve_off[1] = 0;
  md->nbsn_off[0] = 0; md->nbve_off[0] = 0;
  for (int i = 1; i <= M; ++i) {
    md->nbsn_off[i + 1] = md->nbsn_off[i] + (int)nbsn[i].size();
    md->nbve_off[i + 1] = md->nbve_off[i] + (int)nbve[i].size();
    for (size_t k = 0; k < nbsn[i].size(); ++k) md->nbsn[md->nbsn_off[i] + k] = nbsn[i][k];
    for (size_t k = 0; k < nbve[i].size(); ++k) {
      md->nbve[md->nbve_off[i] + k] = nbve[i][k];
      md->nbvt[md->nbve_off[i] + k] = nbvt[i][k];
    }
  }

  // ---------------- CELL_AREA (cell_area.F) ----------------
  for (int i = 1; i <= N; ++i) {
    double a = (md->vx[B.NV(i, 2)] - md->vx[B.NV(i, 1)]) * (md->vy[B.NV(i, 3)] - md->vy[B.NV(i, 1)]) -
               (md->vx[B.NV(i, 3)] - md->vx[B.NV(i, 1)]) * (md->vy[B.NV(i, 2)] - md->vy[B.NV(i, 1)]);
    md->art[i] = std::fabs(0.5 * a);
  }
  {
    std::vector<double> XX, YY;
    for (int i = 1; i <= M; ++i) {
      const int nt = md->ntve[i];
      const int* ve = &md->nbve[md->nbve_off[i]];
      const int* vt = &md->nbvt[md->nbve_off[i]];
      XX.assign(2 * nt + 4, 0.0); YY.assign(2 * nt + 4, 0.0);
      for (int j = 1; j <= nt; ++j) {
        int ii = ve[j - 1], j1 = vt[j - 1], j2 = nxt(j1);
        XX[2 * j - 1] = (md->vx[B.NV(ii, j1)] + md->vx[B.NV(ii, j2)]) * 0.5 - md->vx[i];
        YY[2 * j - 1] = (md->vy[B.NV(ii, j1)] + md->vy[B.NV(ii, j2)]) * 0.5 - md->vy[i];
        XX[2 * j] = md->xc[ii] - md->vx[i];
        YY[2 * j] = md->yc[ii] - md->vy[i];
      }
      double a1 = 0.0;
      if (md->isonb[i] == 0) {
        XX[2 * nt + 1] = XX[1]; YY[2 * nt + 1] = YY[1];
        for (int j = 1; j <= 2 * nt; ++j) a1 = a1 + 0.5 * (XX[j + 1] * YY[j] - XX[j] * YY[j + 1]);
      } else {
        int j = nt + 1;
        int ii = ve[j - 2], j1 = vt[nt - 1], j2 = nxt2(j1);
        XX[2 * j - 1] = (md->vx[B.NV(ii, j1)] + md->vx[B.NV(ii, j2)]) * 0.5 - md->vx[i];
        YY[2 * j - 1] = (md->vy[B.NV(ii, j1)] + md->vy[B.NV(ii, j2)]) * 0.5 - md->vy[i];
        XX[2 * j] = md->vx[i] - md->vx[i];
        YY[2 * j] = md->vy[i] - md->vy[i];
        XX[2 * j + 1] = XX[1]; YY[2 * j + 1] = YY[1];
        for (int jj = 1; jj <= 2 * nt + 2; ++jj) a1 = a1 + 0.5 * (XX[jj + 1] * YY[jj] - XX[jj] * YY[jj + 1]);
      }
      md->art1[i] = std::fabs(a1);
    }
  }
  for (int i = 1; i <= M; ++i) {
    double s = 0.0;
    for (int k = md->nbve_off[i]; k < md->nbve_off[i + 1]; ++k) s += md->art[md->nbve[k]];
    md->art2[i] = s;
  }
  md->art[0] = md->art[1];
  md->art1[0] = md->art1[1];

  // ---------------- SHAPE_COEF_GCN (shape_coef_gcn.F) ----------------
  for (int i = 1; i <= N; ++i) {
    if (md->isbce[i] == 0) {
      double y1 = md->yc[B.NBE(i, 1)] - md->yc[i], y2 = md->yc[B.NBE(i, 2)] - md->yc[i], y3 = md->yc[B.NBE(i, 3)] - md->yc[i];
      double x1 = md->xc[B.NBE(i, 1)] - md->xc[i], x2 = md->xc[B.NBE(i, 2)] - md->xc[i], x3 = md->xc[B.NBE(i, 3)] - md->xc[i];
      x1 /= 1000.0; x2 /= 1000.0; x3 /= 1000.0; y1 /= 1000.0; y2 /= 1000.0; y3 /= 1000.0;
      double delt = (x1 * y2 - x2 * y1) * (x1 * y2 - x2 * y1) + (x1 * y3 - x3 * y1) * (x1 * y3 - x3 * y1) +
                    (x2 * y3 - x3 * y2) * (x2 * y3 - x3 * y2);
      delt = delt * 1000.0;
      double sxy = x1 * y1 + x2 * y2 + x3 * y3;
      double syy = y1 * y1 + y2 * y2 + y3 * y3;
      double sxx = x1 * x1 + x2 * x2 + x3 * x3;
      B.A1U(i, 1) = ((y1 + y2 + y3) * sxy - (x1 + x2 + x3) * syy) / delt;
      B.A1U(i, 2) = (syy * x1 - sxy * y1) / delt;
      B.A1U(i, 3) = (syy * x2 - sxy * y2) / delt;
      B.A1U(i, 4) = (syy * x3 - sxy * y3) / delt;
      B.A2U(i, 1) = ((x1 + x2 + x3) * sxy - (y1 + y2 + y3) * sxx) / delt;
      B.A2U(i, 2) = (sxx * y1 - sxy * x1) / delt;
      B.A2U(i, 3) = (sxx * y2 - sxy * x2) / delt;
      B.A2U(i, 4) = (sxx * y3 - sxy * x3) / delt;
    }
    double x1 = md->vx[B.NV(i, 1)] - md->xc[i], x2 = md->vx[B.NV(i, 2)] - md->xc[i], x3 = md->vx[B.NV(i, 3)] - md->xc[i];
    double y1 = md->vy[B.NV(i, 1)] - md->yc[i], y2 = md->vy[B.NV(i, 2)] - md->yc[i], y3 = md->vy[B.NV(i, 3)] - md->yc[i];
    double ai1 = y2 - y3, ai2 = y3 - y1, ai3 = y1 - y2;
    double bi1 = x3 - x2, bi2 = x1 - x3, bi3 = x2 - x1;
    B.AWX(i, 1) = -ai1 / 2. / md->art[i]; B.AWX(i, 2) = -ai2 / 2. / md->art[i]; B.AWX(i, 3) = -ai3 / 2. / md->art[i];
    B.AWY(i, 1) = -bi1 / 2. / md->art[i]; B.AWY(i, 2) = -bi2 / 2. / md->art[i]; B.AWY(i, 3) = -bi3 / 2. / md->art[i];
  }
  for (int i = 1; i <= N; ++i) {
    if (md->isbce[i] > 1) {
      for (int j = 1; j <= 4; ++j) { B.A1U(i, j) = 0.0; B.A2U(i, j) = 0.0; }
    } else if (md->isbce[i] == 1) {
      int jj = 0;
      for (int j = 1; j <= 3; ++j) if (B.NBE(i, j) == 0) jj = j;
      int j1 = nxt(jj), j2 = nxt2(jj);
      double deltx = md->vx[B.NV(i, j1)] - md->vx[B.NV(i, j2)];
      double delty = md->vy[B.NV(i, j1)] - md->vy[B.NV(i, j2)];
      md->alpha[i] = std::atan2(delty, deltx) - 3.1415926 / 2.0;
      for (int j = 1; j <= 4; ++j) { B.A1U(i, j) = 0.0; B.A2U(i, j) = 0.0; }
    }
  }

  // ---------------- EDGE_LEN (edge_len.F) ----------------
  for (int c = 1; c <= md->ncv_i; ++c) {
    int ia = md->niec[c], ib = md->niec[cv1 + c];
    double xi = 0.5 * (xije1[c] + xije2[c]), yi = 0.5 * (yije1[c] + yije2[c]);
    md->dltxncve[c] = xi - md->vx[ia]; md->dltyncve[c] = yi - md->vy[ia];
    md->dltxncve[cv1 + c] = xi - md->vx[ib]; md->dltyncve[cv1 + c] = yi - md->vy[ib];
  }
  for (int i = 1; i <= M; ++i) {
    if (md->ntsn[i] > 13) return fail("MORE THAN 12 NODES AROUND ONE NODE");
    if (md->ntve[i] > 13) return fail("MORE THAN 12 CELLS AROUND ONE NODE");
    const int o = md->nbsn_off[i];
    for (int j = 0; j < md->ntsn[i] - 1; ++j) {
      int i1 = md->nbsn[o + j], i2 = md->nbsn[o + j + 1];
      md->dltytrie[o + j] = md->vy[i1] - md->vy[i2];
      md->dltxtrie[o + j] = md->vx[i2] - md->vx[i1];
    }
    const int ov = md->nbve_off[i];
    for (int j = 0; j < md->ntve[i]; ++j) {
      int i1 = md->nbve[ov + j], jt = md->nbvt[ov + j];
      int j1 = nxt(jt), j2 = nxt2(jt);
      double x11 = 0.5 * (md->vx[i] + md->vx[B.NV(i1, j1)]), y11 = 0.5 * (md->vy[i] + md->vy[B.NV(i1, j1)]);
      double x33 = 0.5 * (md->vx[i] + md->vx[B.NV(i1, j2)]), y33 = 0.5 * (md->vy[i] + md->vy[B.NV(i1, j2)]);
      md->dltyecec[ov + j] = y11 - y33;
      md->dltxecec[ov + j] = x33 - x11;
      md->dltynec[ov + j] = md->vy[i] - y11;
      md->dltxnec[ov + j] = x11 - md->vx[i];
    }
  }

  // ---------------- DEPTH_GRADIENT (depth_grad.F) ----------------
  for (int i = 1; i <= M; ++i) {
    double phpx = 0.0, phpy = 0.0;
    const int o = md->nbsn_off[i];
    for (int j = 0; j < md->ntsn[i] - 1; ++j) {
      int i1 = md->nbsn[o + j], i2 = md->nbsn[o + j + 1];
      double f1 = 0.50 * (md->h[i1] + md->h[i2]);
      phpx = phpx + f1 * (md->vy[i1] - md->vy[i2]);
      phpy = phpy + f1 * (md->vx[i2] - md->vx[i1]);
    }
    phpx = phpx / md->art2[i];
    phpy = phpy / md->art2[i];
    if (phpx == 0.0 && phpy == 0.0) { md->sita_gd[i] = 0.0; md->phpn[i] = 0.0; }
    else { md->sita_gd[i] = std::atan2(phpy, phpx); md->phpn[i] = std::sqrt(phpy * phpy + phpx * phpx); }
  }

  // ---------------- SETUP_OBC (mod_obcs.F:868-1173) ----------------
  for (int i = 0; i <= N; ++i) md->iucp[i] = 1;
  for (int k = 0; k < iobcn; ++k) {
    const int inode = obc[k];
    md->i_obc_n[k] = inode;
    md->type_obc[k] = 1;
    int nadjn = 0, adjn[2] = {0, 0};
    const int o = md->nbsn_off[inode];
    for (int j = 0; j < md->ntsn[inode] - 1; ++j) {
      int jn = md->nbsn[o + j];
      if (md->isonb[jn] == 2 && inode != jn) { if (nadjn < 2) adjn[nadjn] = jn; nadjn++; }
    }
    if (nadjn == 0) return fail("NO ADJACENT NODE FOUND FOR BOUNDARY NODE");
    double nxo = 0, nyo = 0, nxt_ = 0, nyt_ = 0;
    int nadjc = 0;
    for (int pass = 0; pass < (nadjn > 1 ? 2 : 1); ++pass) {
      int i2 = adjn[pass];
      for (int kk = md->nbve_off[inode]; kk < md->nbve_off[inode + 1]; ++kk) {
        int ic = md->nbve[kk];
        if (B.NV(ic, 1) == i2 || B.NV(ic, 2) == i2 || B.NV(ic, 3) == i2) {
          double dxn = md->vx[i2] - md->vx[inode], dyn = md->vy[i2] - md->vy[inode];
          double dxc = md->xc[ic] - md->vx[inode], dyc = md->yc[ic] - md->vy[inode];
          double cross = (dxc * dyn - dyc * dxn) >= 0 ? 1.0 : -1.0;
          double len = std::sqrt(dxn * dxn + dyn * dyn);
          if (pass == 0) { nxt_ = cross * dyn / len; nyt_ = -cross * dxn / len; nxo = nxt_; nyo = nyt_; }
          else {
            nxt_ = nxt_ + cross * dyn / len; nyt_ = nyt_ - cross * dxn / len;
            double nn = std::sqrt(nxt_ * nxt_ + nyt_ * nyt_);
            nxo = nxt_ / nn; nyo = nyt_ / nn;
          }
          if (nadjc < 2) md->adjc_obc[(size_t)nadjc * iobcn + k] = ic;
          nadjc++;
          md->iucp[ic] = 0;  // TYPE_OBC = 1 (odd)
        }
      }
    }
    md->nadjc_obc[k] = std::min(nadjc, 2);
    if (nadjn == 1) { md->fluxf_obc[k] = 1.0; md->fluxf_obc[iobcn + k] = 0.0; }
    else {
      double e1 = std::sqrt((md->vx[inode] - md->vx[adjn[0]]) * (md->vx[inode] - md->vx[adjn[0]]) +
                            (md->vy[inode] - md->vy[adjn[0]]) * (md->vy[inode] - md->vy[adjn[0]]));
      double e2 = std::sqrt((md->vx[inode] - md->vx[adjn[1]]) * (md->vx[inode] - md->vx[adjn[1]]) +
                            (md->vy[inode] - md->vy[adjn[1]]) * (md->vy[inode] - md->vy[adjn[1]]));
      md->fluxf_obc[k] = e1 / (e1 + e2);
      md->fluxf_obc[iobcn + k] = e2 / (e1 + e2);
    }
    double dotmax = -1.0;
    int nextn = 0;
    for (int j = 0; j < md->ntsn[inode] - 1; ++j) {
      int jn = md->nbsn[o + j];
      if (md->isonb[jn] != 2 && inode != jn) {
        double dx = md->vx[jn] - md->vx[inode], dy = md->vy[jn] - md->vy[inode];
        double ln = std::sqrt(dx * dx + dy * dy);
        double dot = dx / ln * nxo + dy / ln * nyo;
        if (dot > dotmax) { dotmax = dot; nextn = jn; }
      }
    }
    if (nextn == 0) return fail("no interior neighbour for OBC node");
    md->next_obc[k] = nextn;
    // NEXT_OBC2: NEXT_OBC's ring neighbour off the boundary most normal to it
    // (mod_obcs.F:1072-1101, the direction normalised with a 1e-9 guard)
    {
      double dotmax2 = -1.0;
      int next2 = 0;
      const int o2 = md->nbsn_off[nextn];
      for (int j = 0; j < md->ntsn[nextn] - 1; ++j) {
        const int jn = md->nbsn[o2 + j];
        if (md->isonb[jn] == 2) continue;
        const double dxt = md->vx[jn] - md->vx[nextn], dyt = md->vy[jn] - md->vy[nextn];
        const double dxn = dxt / (std::sqrt(dxt * dxt + dyt * dyt) + 1.0e-9);
        const double dyn = dyt / (std::sqrt(dxt * dxt + dyt * dyt) + 1.0e-9);
        const double dot = dxn * nxo + dyn * nyo;
        if (dot > dotmax2) { dotmax2 = dot; next2 = jn; }
      }
      md->next_obc2[k] = next2;
    }
    md->dltn_obc[k] = std::fabs((md->vx[nextn] - md->vx[inode]) * nxo + (md->vy[nextn] - md->vy[inode]) * nyo);
  }

  // ---------------- sigma levels, uniform (mod_setup.F:848-877, 1009-1091) ----------------
  const int KBM1 = KB - 1, KBM2 = KB - 2;
  std::vector<double> ztmp(KB + 1);
  for (int k = 1; k <= KB; ++k) ztmp[k] = -(double)((float)(k - 1) / (float)(KB - 1));  // FLOAT() is default real
  const size_t m1 = M + 1, n1 = N + 1;
  for (int k = 1; k <= KB; ++k) for (int i = 0; i <= M; ++i) md->z[(k - 1) * m1 + i] = ztmp[k];
  auto sig = [&](double* zz, double* dz, double* dzz, const double* zb, size_t L) {
    for (int k = 1; k <= KB - 1; ++k)
      for (size_t i = 0; i < L; ++i) {
        dz[(k - 1) * L + i] = zb[(k - 1) * L + i] - zb[k * L + i];
        zz[(k - 1) * L + i] = .5 * (zb[(k - 1) * L + i] + zb[k * L + i]);
      }
    for (size_t i = 0; i < L; ++i) zz[(KB - 1) * L + i] = 2.0 * zz[(KB - 2) * L + i] - zz[(KB - 3) * L + i];
    for (int k = 1; k <= KBM2; ++k)
      for (size_t i = 0; i < L; ++i) dzz[(k - 1) * L + i] = zz[(k - 1) * L + i] - zz[k * L + i];
    for (size_t i = 0; i < L; ++i) { dzz[(KBM1 - 1) * L + i] = 0.0; dz[(KB - 1) * L + i] = 0.0; }
  };
  sig(md->zz, md->dz, md->dzz, md->z, m1);
  for (int i = 1; i <= N; ++i) {
    md->h1[i] = (md->h[B.NV(i, 1)] + md->h[B.NV(i, 2)] + md->h[B.NV(i, 3)]) / 3.0;
    for (int k = 1; k <= KB; ++k)
      md->z1[(k - 1) * n1 + i] = (md->z[(k - 1) * m1 + B.NV(i, 1)] + md->z[(k - 1) * m1 + B.NV(i, 2)] +
                                  md->z[(k - 1) * m1 + B.NV(i, 3)]) / 3.0;
  }
  sig(md->zz1, md->dz1, md->dzz1, md->z1, n1);
  // Constant physics coefficients (mod_setup.F:314-453).
  for (int i = 0; i <= M; ++i) md->grav_n[i] = 9.81;
  for (int i = 0; i <= N; ++i) md->grav_e[i] = 9.81;
  return 0;
}

}  // namespace

extern "C" int fv_model_from_mesh(int32_t m, int32_t n, const double* x, const double* y,
                                  const int32_t* nv_rows, const double* h, int32_t kb,
                                  int32_t iobcn, const int32_t* obc_nodes, const fv_params* p,
                                  fv_model** out) {
  if (m < 3 || n < 1 || kb < 4) return fail("mesh too small (need M>=3, N>=1, KB>=4)");
  fv_model* md = (fv_model*)xalloc(sizeof(fv_model));
  md->m = m; md->n = n; md->kb = kb; md->kbm1 = kb - 1; md->kbm2 = kb - 2;
  md->p = *p;
  int rc = build_metrics(md, x, y, nv_rows, h, iobcn, obc_nodes);
  if (rc) { fv_model_free(md); return rc; }
  // constant physics coefficients (mod_setup.F:314-371, 402-427)
  for (int i = 0; i <= n; ++i) {
    md->cc_hvc[i] = i ? 0.2 : 0.0;
    md->cc_z0b[i] = i ? 0.001 : 0.0;
    md->cor[i] = i ? 1.0e-4 : 0.0;
    md->cc_sponge[i] = 0.0;
  }
  for (int i = 0; i <= m; ++i) md->nn_hvc[i] = i ? 0.2 : 0.0;
  rc = fv_model_cold_start(md);
  if (rc) { fv_model_free(md); return rc; }
  *out = md;
  return 0;
}

namespace {

// Breadth-first search of the node graph from `root` over the nodes whose
// level is unset (-1): returns them in visiting order with their levels
// set; the caller clears the levels again.
std::vector<int> bfs_levels(int root, const std::vector<int>& off, const std::vector<int>& adj,
                            std::vector<int>& level) {
  std::vector<int> order;
  order.push_back(root);
  level[root] = 0;
  for (size_t h = 0; h < order.size(); ++h) {
    const int u = order[h];
    for (int p = off[u]; p < off[u + 1]; ++p) {
      const int v = adj[p];
      if (level[v] < 0) {
        level[v] = level[u] + 1;
        order.push_back(v);
      }
    }
  }
  return order;
}

}  // namespace

// Reverse Cuthill-McKee numbering of a triangle mesh's nodes (the node
// graph of the triangles' edges), each connected component started from a
// pseudo-peripheral node (George and Liu: repeated breadth-first searches
// from a node of least degree in the last level until the depth stops
// growing), neighbours visited by increasing degree, then the order
// reversed.  Elements are numbered by their lowest, middle and highest new
// node index.  Outputs are 1-based "new index of old index" maps:
// node_new[old-1] and elem_new[old-1].
extern "C" int fv_rcm_order(int32_t m, int32_t n, const int32_t* nv_rows, int32_t* node_new, int32_t* elem_new) {
  if (m < 1 || n < 1 || !nv_rows || !node_new || !elem_new) return fail("fv_rcm_order: bad arguments");
  for (int64_t q = 0; q < 3LL * n; ++q)
    if (nv_rows[q] < 1 || nv_rows[q] > m) return fail("fv_rcm_order: node index out of range");
  // node adjacency (CSR, 0-based), duplicates removed
  std::vector<std::vector<int>> nb(m);
  for (int e = 0; e < n; ++e)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (a != b) nb[nv_rows[3 * e + a] - 1].push_back(nv_rows[3 * e + b] - 1);
  std::vector<int> off(m + 1, 0), adj, deg(m);
  for (int u = 0; u < m; ++u) {
    std::sort(nb[u].begin(), nb[u].end());
    nb[u].erase(std::unique(nb[u].begin(), nb[u].end()), nb[u].end());
    deg[u] = (int)nb[u].size();
    off[u + 1] = off[u] + deg[u];
  }
  adj.reserve(off[m]);
  for (int u = 0; u < m; ++u) {
    std::sort(nb[u].begin(), nb[u].end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
    adj.insert(adj.end(), nb[u].begin(), nb[u].end());
  }
  std::vector<int> level(m, -1), order;
  std::vector<char> done(m, 0);
  order.reserve(m);
  for (int seed = 0; seed < m; ++seed) {
    if (done[seed]) continue;
    // least-degree node of this component
    std::vector<int> comp = bfs_levels(seed, off, adj, level);
    int root = comp[0];
    for (int u : comp) {
      level[u] = -1;
      if (deg[u] < deg[root] || (deg[u] == deg[root] && u < root)) root = u;
    }
    // pseudo-peripheral node: from the least-degree node of the last level
    // while the number of levels grows
    for (int it = 0; it < 16; ++it) {
      std::vector<int> o = bfs_levels(root, off, adj, level);
      const int d0 = level[o.back()];
      int cand = -1;
      for (int u : o)
        if (level[u] == d0 && (cand < 0 || deg[u] < deg[cand] || (deg[u] == deg[cand] && u < cand))) cand = u;
      for (int u : o) level[u] = -1;
      std::vector<int> o2 = bfs_levels(cand, off, adj, level);
      const int d1 = level[o2.back()];
      for (int u : o2) level[u] = -1;
      if (d1 <= d0) break;
      root = cand;
    }
    // Cuthill-McKee from root (neighbours already sorted by degree)
    const size_t start = order.size();
    order.push_back(root);
    done[root] = 1;
    for (size_t h = start; h < order.size(); ++h) {
      const int u = order[h];
      for (int p = off[u]; p < off[u + 1]; ++p)
        if (!done[adj[p]]) { done[adj[p]] = 1; order.push_back(adj[p]); }
    }
  }
  if ((int)order.size() != m) return fail("fv_rcm_order: ordering incomplete");
  std::reverse(order.begin(), order.end());
  for (int k = 0; k < m; ++k) node_new[order[k]] = k + 1;
  // elements by (lowest, middle, highest) new node index
  std::vector<int> eo(n);
  std::vector<std::array<int, 3>> key(n);
  for (int e = 0; e < n; ++e) {
    std::array<int, 3> t = {node_new[nv_rows[3 * e] - 1], node_new[nv_rows[3 * e + 1] - 1],
                            node_new[nv_rows[3 * e + 2] - 1]};
    std::sort(t.begin(), t.end());
    key[e] = t;
    eo[e] = e;
  }
  std::stable_sort(eo.begin(), eo.end(), [&](int a, int b) { return key[a] < key[b]; });
  for (int k = 0; k < n; ++k) elem_new[eo[k]] = k + 1;
  return 0;
}

// Room for one more river: every per-river array grows by one entry, the
// existing rivers kept in their order; returns the new river's index.
static int add_river_slot(fv_model* md) {
  const int KBM1 = md->kbm1, k = md->numqbc;
  auto grow = [&](auto*& a, int comps) {   // (k, comps) -> (k + 1, comps), component-major
    using T = std::remove_reference_t<decltype(*a)>;
    T* b = (T*)xalloc(sizeof(T) * (size_t)(k + 1) * comps);
    for (int c = 0; c < comps; ++c)
      for (int j = 0; j < k; ++j) b[(size_t)c * (k + 1) + j] = a[(size_t)c * k + j];
    std::free(a);
    a = b;
  };
  grow(md->inodeq, 1); grow(md->qarea, 1); grow(md->angleq, 1);
  grow(md->qdis0, 1); grow(md->tdis, 1); grow(md->sdis, 1);
  grow(md->icellq, 1); grow(md->n_icellq, 2); grow(md->rdisq, 2);
  grow(md->vqdist, KBM1);
  md->numqbc = k + 1;
  md->inodeq[k] = 0; md->icellq[k] = 0;
  md->n_icellq[k] = 0; md->n_icellq[k + 1 + k] = 0;
  md->rdisq[k] = 0.0; md->rdisq[k + 1 + k] = 0.0;
  return k;
}

extern "C" int fv_model_set_river(fv_model* md, int32_t node, double q, double t, double s) {
  const int KBM1 = md->kbm1;
  const size_t m1 = md->m + 1;
  if (node < 1 || node > md->m) return fail("river node out of range");
  if (md->isonb[node] != 1) return fail("river node is not on the solid boundary");
  const int k = add_river_slot(md), nq = md->numqbc;
  md->inodeq[k] = node; md->qdis0[k] = q; md->tdis[k] = t; md->sdis[k] = s;
  // QAREA / ANGLEQ of a node river (mod_force.F:3470-3511)
  const int o = md->nbsn_off[node];
  const int i1 = md->nbsn[o + 1], i2 = md->nbsn[o + 0], i3 = md->nbsn[o + md->ntsn[node] - 2];
  const double dy12 = md->vy[i1] - md->vy[i2], dy32 = md->vy[i3] - md->vy[i2];
  const double dx12 = md->vx[i1] - md->vx[i2], dx32 = md->vx[i3] - md->vx[i2];
  double a1 = std::atan2(dy12, dx12);
  const double a2 = std::atan2(dy32, dx32);
  if (a1 < a2) a1 = a1 + 2.0 * 3.1415927;
  const double dxy = std::sqrt(dx12 * dx12 + dy12 * dy12) + std::sqrt(dx32 * dx32 + dy32 * dy32);
  md->qarea[k] = 0.5 * dxy * md->h[node];
  md->angleq[k] = (a1 - a2) / 2. + a2;
  for (int l = 0; l < KBM1; ++l) md->vqdist[(size_t)l * nq + k] = md->dz[(size_t)l * m1 + node];
  return 0;
}

extern "C" int fv_model_set_river_edge(fv_model* md, int32_t cell, double q, double t, double s) {
  const int KBM1 = md->kbm1;
  const size_t n1 = md->n + 1;
  if (cell < 1 || cell > md->n) return fail("river cell out of range");
  if (md->isbce[cell] != 1) return fail("river cell is not on the solid boundary (ISBCE /= 1)");
  int itmp = 0, jtmp = 0;
  for (int j = 1; j <= 3; ++j)
    if (md->nbe[(size_t)(j - 1) * n1 + cell] == 0) { jtmp = j; ++itmp; }
  if (itmp != 1) return fail("river cell must have exactly one solid boundary edge");
  // the solid edge joins the two nodes other than node JTMP (mod_force.F:3540-3547)
  const int j1 = jtmp + 1 - ((jtmp + 1) / 4) * 3, j2 = jtmp + 2 - ((jtmp + 2) / 4) * 3;
  const int i1 = md->nv[(size_t)(j1 - 1) * n1 + cell], i2 = md->nv[(size_t)(j2 - 1) * n1 + cell];
  const int k = add_river_slot(md), nq = md->numqbc;
  md->icellq[k] = cell; md->qdis0[k] = q; md->tdis[k] = t; md->sdis[k] = s;
  md->n_icellq[k] = i1; md->n_icellq[nq + k] = i2;
  // QAREA, ANGLEQ (its pi/2 a default-real constant) and RDISQ (mod_force.F:3548-3567)
  const double htmp = 0.5 * (md->h[i1] + md->h[i2]);
  const double dy12 = md->vy[i1] - md->vy[i2], dx12 = md->vx[i1] - md->vx[i2];
  const double atmp1 = std::atan2(dy12, dx12);
  md->qarea[k] = std::sqrt(dx12 * dx12 + dy12 * dy12) * htmp;
  md->angleq[k] = atmp1 + (double)(3.1415927f / 2.0f);
  const double areatmp = md->art1[i1] + md->art1[i2];
  md->rdisq[k] = md->art1[i1] / areatmp;
  md->rdisq[nq + k] = md->art1[i2] / areatmp;
  // vertical distribution: the cell's layer thicknesses
  for (int l = 0; l < KBM1; ++l) md->vqdist[(size_t)l * nq + k] = md->dz1[(size_t)l * n1 + cell];
  return 0;
}

extern "C" int fv_case_defaults(int32_t config, fv_case* c) {
  std::memset(c, 0, sizeof(*c));
  fv_params& p = c->p;
  p.umol = 1.0e-4; p.vprnu = 1.0; p.hprnu = 1.0; p.cbcmin = 0.0025; p.min_depth = 0.05;
  p.horizontal_closure = 1; p.vertical_closure = 1; p.scalar_positivity = 1; p.noflux_bot = 1;
  p.barotropic = 0; p.temperature_active = 1; p.salinity_active = 1;
  p.river_ts_specified = 1; p.type_tsobc = 3; p.obc_on = 1; p.obc_elevation_on = 1;
  p.iint0 = 0; p.time0 = 0.0;
  c->coriolis = 1.0e-4;
  c->tide_amp[0] = 1.0;
  // a sponge along the open boundary, as the reference's Estuary example
  // carries one (Examples/Estuary/tstinp/tst_spg.dat)
  c->sponge_radius = 4000.0;
  c->sponge_coef = 0.01;
  c->seed = 20240601;
  switch (config) {
    case 0:
      c->nx = 81; c->ny = 21; c->dx = 500.0; c->jitter = 0.0; c->kb = 11; c->depth_kind = 0;
      c->ntide = 1; c->river = 1; c->river_q = 100.0; c->river_t = 15.0; c->river_s = 0.0;
      c->wind = 0; c->ts_kind = 0;
      p.dte = 5.0; p.isplit = 10; p.iramp = 72; p.wind_on = 0;
      return 0;
    case 1:
    case 3:
      c->nx = config == 1 ? 1001 : 501; c->ny = config == 1 ? 501 : 251; c->dx = 200.0;
      c->jitter = 0.2; c->kb = config == 1 ? 41 : 21; c->depth_kind = 1; c->n_seamounts = 20;
      c->ntide = 3; c->tide_amp[1] = 0.4; c->tide_amp[2] = 0.3;
      c->river = 0; c->wind = 1; c->wind_mean = 8.0; c->wind_std = 2.0; c->wind_corr = 50000.0;
      c->ts_kind = 1;
      if (config == 3) { c->el_pert_std = 0.05; c->el_pert_corr = 20000.0; }
      p.dte = 2.0; p.isplit = 10; p.iramp = 180; p.wind_on = 1;
      return 0;
  }
  return fail("unknown configuration");
}

extern "C" int fv_model_build(const fv_case* c, fv_model** out) {
  const int nx = c->nx, ny = c->ny;
  if (nx < 3 || ny < 3) return fail("nx, ny must be >= 3");
  const int M = nx * ny, N = 2 * (nx - 1) * (ny - 1);
  Rng rng(c->seed);
  std::vector<double> x(M), y(M), h(M);
  const double Lx = (nx - 1) * c->dx;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int k = j * nx + i;
      double jx = c->jitter > 0 ? (2.0 * rng.uniform() - 1.0) * c->jitter * c->dx : 0.0;
      double jy = c->jitter > 0 ? (2.0 * rng.uniform() - 1.0) * c->jitter * c->dx : 0.0;
      x[k] = i * c->dx + jx; y[k] = j * c->dx + jy;
    }
  // depth
  std::vector<double> smx, smy, sma, sms;
  for (int s = 0; s < c->n_seamounts; ++s) {
    smx.push_back(rng.uniform() * Lx); smy.push_back(rng.uniform() * (ny - 1) * c->dx);
    sma.push_back(30.0 + 50.0 * rng.uniform()); sms.push_back(2000.0 + 6000.0 * rng.uniform());
  }
  for (int k = 0; k < M; ++k) {
    double xf = std::min(std::max(x[k] / Lx, 0.0), 1.0);
    if (c->depth_kind == 0) h[k] = 5.0 + 15.0 * xf;
    else {
      double d = 5.0 + 195.0 * std::pow(xf, 1.5);
      for (size_t s = 0; s < smx.size(); ++s) {
        double r2 = (x[k] - smx[s]) * (x[k] - smx[s]) + (y[k] - smy[s]) * (y[k] - smy[s]);
        d += sma[s] * std::exp(-r2 / (2.0 * sms[s] * sms[s]));
      }
      h[k] = std::max(d, 5.0);
    }
  }
  // triangles, clockwise (the reference swaps the file's counter-clockwise
  // order to clockwise on input, mod_input.F:3739-3742)
  std::vector<int32_t> nv((size_t)N * 3);
  int e = 0;
  for (int j = 0; j < ny - 1; ++j)
    for (int i = 0; i < nx - 1; ++i) {
      int a = j * nx + i + 1, b = a + 1, d = a + nx, cc = d + 1;  // a=(i,j) b=(i+1,j) d=(i,j+1) cc=(i+1,j+1)
      int tri[2][3];
      if (((i + j) & 1) == 0) { tri[0][0] = a; tri[0][1] = b; tri[0][2] = cc; tri[1][0] = a; tri[1][1] = cc; tri[1][2] = d; }
      else { tri[0][0] = a; tri[0][1] = b; tri[0][2] = d; tri[1][0] = b; tri[1][1] = cc; tri[1][2] = d; }
      for (int t = 0; t < 2; ++t) {  // store clockwise: (n1, n3, n2) of the counter-clockwise triple
        nv[(size_t)e * 3 + 0] = tri[t][0]; nv[(size_t)e * 3 + 1] = tri[t][2]; nv[(size_t)e * 3 + 2] = tri[t][1];
        ++e;
      }
    }
  // open boundary: east column without its two corners
  std::vector<int32_t> obc;
  for (int j = 1; j < ny - 1; ++j) obc.push_back(j * nx + (nx - 1) + 1);
  fv_model* md = nullptr;
  int rc = fv_model_from_mesh(M, N, x.data(), y.data(), nv.data(), h.data(), c->kb, (int)obc.size(),
                              obc.data(), &c->p, &md);
  if (rc) return rc;
  const int KB = md->kb, KBM1 = md->kbm1;
  const size_t m1 = M + 1, n1 = N + 1;
  // tides: M2 (44712 s), S2 (43200 s), K1 (86164 s); seeded phases
  md->ntide = c->ntide;
  {
    const double periods[3] = {44712.0, 43200.0, 86164.0};
    double ph[3] = {0.0, 0.0, 0.0};
    if (c->ntide > 1) for (int t = 0; t < 3; ++t) ph[t] = 360.0 * rng.uniform();
    std::free(md->apt); std::free(md->phai); std::free(md->period);
    md->apt = (double*)xalloc(sizeof(double) * md->iobcn * md->ntide);
    md->phai = (double*)xalloc(sizeof(double) * md->iobcn * md->ntide);
    md->period = (double*)xalloc(sizeof(double) * md->ntide);
    for (int t = 0; t < md->ntide; ++t) {
      md->period[t] = periods[t];
      for (int k = 0; k < md->iobcn; ++k) {
        md->apt[(size_t)t * md->iobcn + k] = c->tide_amp[t];
        md->phai[(size_t)t * md->iobcn + k] = ph[t];
      }
    }
  }
  // physics coefficients (mod_setup.F:314-371, 402-427)
  for (int i = 0; i <= N; ++i) {
    md->cc_hvc[i] = i ? 0.2 : 0.0;
    md->cc_z0b[i] = i ? 0.001 : 0.0;
    md->cor[i] = i ? c->coriolis : 0.0;
    md->cc_sponge[i] = 0.0;
  }
  for (int i = 0; i <= M; ++i) md->nn_hvc[i] = i ? 0.2 : 0.0;
  if (c->sponge_radius > 0.0 && c->sponge_coef > 0.0 && md->iobcn > 0) {
    std::vector<double> r((size_t)md->iobcn, c->sponge_radius), cf((size_t)md->iobcn, c->sponge_coef);
    if (fv_model_set_sponge(md, md->iobcn, md->i_obc_n, r.data(), cf.data())) { fv_model_free(md); return 1; }
  }
  // river at the west-end node (middle of the west wall)
  if (c->river) {
    const int jn = (ny / 2) * nx + 1;
    if (fv_model_set_river(md, jn, c->river_q, c->river_t, c->river_s)) { fv_model_free(md); return 1; }
  }
  // wind field at elements
  if (c->wind) {
    Rng wr(c->seed ^ 0xA5A5A5A5ull);
    RandomField fu(wr, c->wind_corr), fv(wr, c->wind_corr);
    double dir = 2.0 * M_PI * wr.uniform();
    for (int i = 1; i <= N; ++i) {
      md->uuwind[i] = c->wind_mean * std::cos(dir) + c->wind_std * fu(md->xc[i], md->yc[i]);
      md->vvwind[i] = c->wind_mean * std::sin(dir) + c->wind_std * fv(md->xc[i], md->yc[i]);
      md->uuwind0[i] = md->uuwind[i];   // held constant: the same at the start of every step
      md->vvwind0[i] = md->vvwind[i];
    }
  }
  // initial state
  for (int i = 1; i <= M; ++i) {
    for (int k = 1; k <= KBM1; ++k) {
      double t, s;
      if (c->ts_kind == 0) { t = 15.0; s = 30.0 * (md->vx[i] / Lx); }
      else {
        double zr = -md->zz[(k - 1) * m1 + i];  // depth fraction of the layer centre
        t = 20.0 - 10.0 * zr; s = 33.0 + 2.0 * zr;
      }
      md->t1[(k - 1) * m1 + i] = t; md->s1[(k - 1) * m1 + i] = s;
    }
  }
  if (c->el_pert_std > 0) {
    Rng er(c->seed ^ 0x5EED5EEDull);
    RandomField fe(er, c->el_pert_corr);
    for (int i = 1; i <= M; ++i) md->el[i] = c->el_pert_std * fe(md->vx[i], md->vy[i]);
  }
  (void)KB;
  rc = fv_model_cold_start(md);
  if (rc) { fv_model_free(md); return rc; }
  *out = md;
  return 0;
}

extern "C" int fv_model_set_obc_type(fv_model* md, int32_t n, const int32_t* types) {
  if (!md) return fail("fv_model_set_obc_type: null model");
  if (n != md->iobcn) return fail("fv_model_set_obc_type: n must equal IOBCN");
  for (int k = 0; k < n; ++k)
    if (types[k] < 1 || types[k] > 10) return fail("fv_model_set_obc_type: TYPE_OBC must be 1..10");
  for (int k = 0; k < n; ++k) md->type_obc[k] = types[k];
  // IUCP: SETUP_OBC marks the adjacent cells of linear (odd-type) nodes
  for (int i = 0; i <= md->n; ++i) md->iucp[i] = 1;
  for (int k = 0; k < n; ++k) {
    if (md->type_obc[k] % 2 != 1) continue;
    for (int j = 0; j < md->nadjc_obc[k]; ++j) md->iucp[md->adjc_obc[(size_t)j * n + k]] = 0;
  }
  return 0;
}

extern "C" int fv_model_set_sponge(fv_model* md, int32_t n, const int32_t* nodes, const double* radius,
                                   const double* coef) {
  if (n < 0 || (n > 0 && (!nodes || !radius || !coef))) return fail("fv_model_set_sponge: bad arguments");
  for (int j = 0; j < n; ++j) {
    if (nodes[j] < 1 || nodes[j] > md->m) return fail("fv_model_set_sponge: node out of range");
    if (!(radius[j] > 0.0)) return fail("fv_model_set_sponge: radius must be positive");
  }
  for (int i = 1; i <= md->n; ++i)
    for (int j = 0; j < n; ++j) {
      const double dx = md->xc[i] - md->vx[nodes[j]], dy = md->yc[i] - md->vy[nodes[j]];
      const double dtmp = std::sqrt(dx * dx + dy * dy) / radius[j];
      if (dtmp <= 1.0) md->cc_sponge[i] = std::max(coef[j] * (1.0 - dtmp), md->cc_sponge[i]);
    }
  return 0;
}

extern "C" int fv_model_set_tide(fv_model* md, int32_t ntide, const double* period,
                                 const double* amp, const double* phase_deg) {
  if (ntide < 0 || ntide > 16) return fail("ntide out of range");
  std::free(md->apt); std::free(md->phai); std::free(md->period);
  md->ntide = ntide;
  md->apt = (double*)xalloc(sizeof(double) * md->iobcn * ntide);
  md->phai = (double*)xalloc(sizeof(double) * md->iobcn * ntide);
  md->period = (double*)xalloc(sizeof(double) * ntide);
  for (int t = 0; t < ntide; ++t) {
    md->period[t] = period[t];
    for (int k = 0; k < md->iobcn; ++k) {
      md->apt[(size_t)t * md->iobcn + k] = amp[t];
      md->phai[(size_t)t * md->iobcn + k] = phase_deg[t];
    }
  }
  return 0;
}

// DENS2 polynomial (eqs_of_state.F:254-321).
static double dens2(double s, double t) {
  double r = s * s * s * 6.76786136E-6 - s * s * 4.8249614E-4 + s * 8.14876577E-1 - 0.22584586E0;
  r = r * (t * t * t * 1.667E-8 - t * t * 8.164E-7 + t * 1.803E-5);
  r = r + 1.0 - t * t * t * 1.0843E-6 + t * t * 9.8185E-5 - t * 4.786E-3;
  r = r * (s * s * s * 6.76786136E-6 - s * s * 4.8249614E-4 + s * 8.14876577E-1 + 3.895414E-2);
  r = r - (t - 3.98) * (t - 3.98) * (t + 283.0) / (503.57 * (t + 67.26));
  return r * 1.e-3;
}

extern "C" int fv_model_cold_start(fv_model* md) {
  const int M = md->m, N = md->n, KB = md->kb, KBM1 = md->kbm1;
  const size_t m1 = M + 1, n1 = N + 1;
  const double one_third = 1.0 / 3.0;
  const int32_t* nv = md->nv;
  auto n2e = [&](const double* nvar, double* evar) {
    for (int k = 0; k < KB; ++k)
      for (int i = 1; i <= N; ++i)
        evar[k * n1 + i] = one_third * (nvar[k * m1 + nv[i]] + nvar[k * m1 + nv[n1 + i]] + nvar[k * m1 + nv[2 * n1 + i]]);
  };
  // water depth (mod_startup.F:1408-1462); EL may carry an initial perturbation
  for (int i = 1; i <= M; ++i) { md->et[i] = md->el[i]; md->d[i] = md->h[i] + md->el[i]; md->dt[i] = md->h[i] + md->et[i]; }
  for (int i = 1; i <= N; ++i) {
    md->el1[i] = one_third * (md->el[nv[i]] + md->el[nv[n1 + i]] + md->el[nv[2 * n1 + i]]);
    md->et1[i] = md->el1[i];
    md->d1[i] = md->h1[i] + md->el1[i];
    md->dt1[i] = md->h1[i] + md->et1[i];
    md->ua[i] = 0.0; md->va[i] = 0.0;
  }
  // density (DENS2, eqs_of_state.F:381-403) and the mean fields (SET_*_TS)
  for (int k = 1; k <= KB; ++k)
    for (int i = 0; i <= M; ++i) {
      size_t q = (k - 1) * m1 + i;
      md->rho1[q] = (k == KB || i == 0) ? 0.0 : dens2(md->s1[q], md->t1[q]);
    }
  if (md->p.barotropic) for (size_t q = 0; q < m1 * KB; ++q) md->rho1[q] = 2.3E-2;
  n2e(md->rho1, md->rho);
  std::memcpy(md->tmean1, md->t1, sizeof(double) * m1 * KB);
  std::memcpy(md->smean1, md->s1, sizeof(double) * m1 * KB);
  std::memcpy(md->rmean1, md->rho1, sizeof(double) * m1 * KB);
  std::memcpy(md->rmean, md->rho, sizeof(double) * n1 * KB);
  n2e(md->t1, md->t);
  n2e(md->s1, md->s);
  // SET_DEFAULT_TURB (mod_startup.F:1332-1406); 1.E-8 is a default-real literal
  const double umol = md->p.umol, vprnu = md->p.vprnu;
  for (int i = 0; i <= M; ++i) {
    for (int k = 1; k <= KB; ++k) {
      size_t q = (k - 1) * m1 + i;
      md->km[q] = md->kh[q] = md->kq[q] = md->l[q] = md->q2[q] = md->q2l[q] = 0.0;
      md->wts[q] = 0.0;
    }
  }
  for (int k = 2; k <= KBM1; ++k)
    for (int i = 1; i <= M; ++i) {
      size_t q = (k - 1) * m1 + i;
      if (md->d[i] > 0.0) {
        md->q2[q] = (double)1.E-8f; md->q2l[q] = (double)1.E-8f; md->l[q] = 1.;
        md->km[q] = 2. * umol; md->kq[q] = 2. * umol; md->kh[q] = 2. * umol / vprnu;
      }
    }
  n2e(md->km, md->km1);
  for (size_t q = 0; q < n1 * KB; ++q) { md->u[q] = md->v[q] = md->w[q] = md->ww[q] = 0.0; }
  return 0;
}

// --------------------------------------------------------------------------
// Batched instances as one mesh: the disjoint union of independent models.
// Every kernel of the time step is local to a node's or element's own
// neighbourhood, and each gather list is sorted by sub-edge / edge index, so
// instance p of the union advances exactly as it would alone.  Entities are
// renumbered by offsets (instance p's node i becomes offM[p] + i); the dummy
// entry 0 is shared.
// --------------------------------------------------------------------------
namespace {

enum RefKind { REF_NONE = 0, REF_NODE, REF_ELEM };

RefKind field_ref(const char* name) {
  static const char* node_refs[] = {"nbsn", "inodeq", "i_obc_n", "next_obc", "next_obc2", "ienode", "niec", "nv",
                                    "n_icellq"};
  static const char* elem_refs[] = {"nbe", "nbve", "iec", "ntrg", "adjc_obc", "icellq"};
  for (const char* n : node_refs) if (std::strcmp(n, name) == 0) return REF_NODE;
  for (const char* n : elem_refs) if (std::strcmp(n, name) == 0) return REF_ELEM;
  return REF_NONE;
}

// entity count of an extent with a dummy 0 (or a plain list) and its components
struct ExtShape { int kind; int comps; };   // kind 0: 0..cnt with dummy, 1: list 0..cnt-1, 2: CSR offsets, 3: shared
ExtShape ext_shape(int ext, const fv_model* md) {
  switch (ext) {
    case FVX_M1: case FVX_N1: case FVX_NE1: case FVX_CV1: return {0, 1};
    case FVX_N1x3: return {0, 3};
    case FVX_N1x4: return {0, 4};
    case FVX_NE1x2: case FVX_CV1x2: return {0, 2};
    case FVX_M1K: case FVX_N1K: return {0, md->kb};
    case FVX_M2: return {2, 1};
    case FVX_RSN: case FVX_RVE: case FVX_OBC: case FVX_RIV: return {1, 1};
    case FVX_OBCx2: return {1, 2};
    case FVX_OBCxT: return {1, md->ntide};
    case FVX_RIVxK: return {1, md->kbm1};
    case FVX_RIVx2: return {1, 2};
    case FVX_TIDE: return {3, 1};
  }
  return {3, 1};
}

int64_t ext_count(int ext, const fv_model* md) {
  switch (ext) {
    case FVX_M1: case FVX_M1K: case FVX_M2: return md->m;
    case FVX_N1: case FVX_N1x3: case FVX_N1x4: case FVX_N1K: return md->n;
    case FVX_NE1: case FVX_NE1x2: return md->ne;
    case FVX_CV1: case FVX_CV1x2: return md->ncv;
    case FVX_RSN: return md->nring_sn;
    case FVX_RVE: return md->nring_ve;
    case FVX_OBC: case FVX_OBCx2: case FVX_OBCxT: return md->iobcn;
    case FVX_RIV: case FVX_RIVxK: case FVX_RIVx2: return md->numqbc;
    case FVX_TIDE: return md->ntide;
  }
  return 0;
}

}  // namespace

extern "C" int fv_model_union(const fv_model* const* parts, int32_t nparts, fv_model** out) {
  if (nparts < 1 || !parts) return fail("fv_model_union: no parts");
  const fv_model* p0 = parts[0];
  for (int p = 0; p < nparts; ++p) {
    const fv_model* q = parts[p];
    if (q->kb != p0->kb || q->ntide != p0->ntide)
      return fail("fv_model_union: parts differ in KB or NTIDE");
    if (std::memcmp(&q->p, &p0->p, sizeof(fv_params)) != 0)
      return fail("fv_model_union: parts differ in run parameters");
    for (int t = 0; t < p0->ntide; ++t)
      if (q->period[t] != p0->period[t]) return fail("fv_model_union: parts differ in tidal periods");
    if (q->nbsn_off[0] != 0 || q->nbsn_off[1] != 0 || q->nbve_off[0] != 0 || q->nbve_off[1] != 0)
      return fail("fv_model_union: node 0 must have empty rings");
  }
  fv_model* md = (fv_model*)xalloc(sizeof(fv_model));
  *md = fv_model{};
  md->kb = p0->kb; md->kbm1 = p0->kbm1; md->kbm2 = p0->kbm2; md->ntide = p0->ntide; md->p = p0->p;
  std::vector<int64_t> offM(nparts), offN(nparts);
  for (int p = 0; p < nparts; ++p) {
    const fv_model* q = parts[p];
    offM[p] = md->m; offN[p] = md->n;
    md->m += q->m; md->n += q->n; md->ne += q->ne; md->ncv += q->ncv; md->ncv_i += q->ncv_i;
    md->iobcn += q->iobcn; md->numqbc += q->numqbc; md->nring_sn += q->nring_sn; md->nring_ve += q->nring_ve;
  }
  alloc_all(md);
  for (int f = 0; f < kNFields; ++f) {
    const int ext = kFields[f].ext;
    const size_t esz = kFields[f].dtype == 0 ? 8 : 4;
    const RefKind ref = kFields[f].dtype == 1 ? field_ref(kFields[f].name) : REF_NONE;
    char* dst = *(char**)((char*)md + kFields[f].off);
    const ExtShape sh = ext_shape(ext, md);
    const int64_t tot = ext_count(ext, md);
    if (sh.kind == 3) {   // shared (tidal periods)
      std::memcpy(dst, *(char* const*)((const char*)p0 + kFields[f].off), (size_t)fv_extent_len(md, ext) * esz);
      continue;
    }
    if (sh.kind == 2) {   // CSR offsets over nodes: ring starts shifted by the preceding parts' entries
      int32_t* o = (int32_t*)dst;
      const bool sn = std::strcmp(kFields[f].name, "nbsn_off") == 0;
      int32_t shift = 0;
      o[0] = 0;
      for (int p = 0; p < nparts; ++p) {
        const int32_t* src = *(int32_t* const*)((const char*)parts[p] + kFields[f].off);
        for (int i = 1; i <= parts[p]->m; ++i) o[offM[p] + i] = src[i] + shift;
        shift += sn ? parts[p]->nring_sn : parts[p]->nring_ve;
      }
      o[md->m + 1] = shift;
      continue;
    }
    const int base = sh.kind == 0 ? 1 : 0;            // first copied entry of each component
    const int64_t stride = tot + (sh.kind == 0 ? 1 : 0);
    int64_t at = 0;                                   // entities written so far (per component)
    for (int p = 0; p < nparts; ++p) {
      const fv_model* q = parts[p];
      const char* src = *(char* const*)((const char*)q + kFields[f].off);
      const int64_t cnt = ext_count(ext, q);
      const int64_t sstride = cnt + (sh.kind == 0 ? 1 : 0);
      const int32_t add = ref == REF_NODE ? (int32_t)offM[p] : ref == REF_ELEM ? (int32_t)offN[p] : 0;
      for (int c = 0; c < sh.comps; ++c) {
        if (p == 0 && sh.kind == 0)   // the shared dummy entry 0 of every component
          std::memcpy(dst + (size_t)(c * stride) * esz, src + (size_t)(c * sstride) * esz, esz);
        for (int64_t i = 0; i < cnt; ++i) {
          const size_t d = (size_t)(c * stride + base + at + i), s = (size_t)(c * sstride + base + i);
          if (ref != REF_NONE) {
            const int32_t v = ((const int32_t*)src)[s];
            ((int32_t*)dst)[d] = v > 0 ? v + add : v;
          } else {
            std::memcpy(dst + d * esz, src + s * esz, esz);
          }
        }
      }
      at += cnt;
    }
  }
  *out = md;
  return 0;
}
