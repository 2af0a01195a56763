/*
 * fvcom_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial FP64 restatement of FVCOM's mode-split time step (INTERNAL_STEP,
 * FVCOM_source/internal_step.F:29-1607, and EXTERNAL_STEP,
 * FVCOM_source/external_step.F:30-443) in the reference's own loop and
 * scatter order, for the build variant GCN + LIMITED_NO + DOUBLE_PRECISION,
 * Cartesian, no wet/dry, no heating; open-boundary types 1-10 with the
 * spectral tide, TYPE_TSOBC 1-4, node and edge rivers.
 *
 * The HIP path is tested against it, and it is the benchmark's CPU baseline.
 * Only tests/, the smoke check and the benchmark's CPU leg may load it; the
 * product library never links or calls it.
 *
 * Every routine cites the reference lines it restates.  Grid and metric
 * arrays are read from the fv_model built by libfvgrid (the reference's
 * GRID_METRICS output); state arrays are copied into the oracle.
 *
 * Pinning (what checks each routine beyond restatement) is listed in
 * tests/test_oracle.py (whole steps: budgets, seiche, tide, densities),
 * tests/test_oracle_pins.py and tests/test_oracle_pins_limiters.py (single
 * routines through orc_routine against closed forms and published
 * results), tests/test_obc_types.py (open-boundary types) and
 * tests/test_river_edge.py (edge rivers).  Every routine has such a check;
 * these parts of routines have none ("parity unpinned": restatement only,
 * an error there would be shared by the HIP path):
 *   - which bound BCOND_TS's clip lands on (the vertical interpolants
 *     inside the range);
 *   - VISCOF_H's closure at boundary nodes (ISONB != 0);
 *   - VDIF_Q's stable-side STF factor;
 *   - the DZZ weights of FCT_Q2/FCT_Q2L's vertical bounds on unequal
 *     spacing.
 */
#include "../include/fvgpu.h"
#include "../include/fv_detmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef struct orc {
  const fv_model* g;
  fv_params p;
  int M, N, NE, NCV, KB, KBM1, KBM2, IOBCN, NUMQBC;
  size_t m1, n1;
  double dte, dti;
  long long iint;
  int iext;
  double ramp, int_time, ext_time, rk_time;
  /* state, 2-D nodes */
  double *el, *et, *d, *dt, *elf, *elrk, *dtfa, *egf, *elm1, *elm2;
  int has_ore;   /* some OBC node is of TYPE_OBC 9/10 */
  /* state, 2-D elements */
  double *ua, *va, *el1, *et1, *d1, *dt1, *elf1, *elrk1, *uaf, *vaf, *uark, *vark;
  double *uard, *vard, *adx2d, *ady2d, *drx2d, *dry2d, *advua, *advva, *pstx, *psty;
  double *fluxobn, *cbc, *wubot, *wvbot, *wusurf, *wvsurf, *wusurf2, *wvsurf2, *tps;
  double *uard_obcn, *xflux_obcn, *xflux_obc, *qdis;
  double *pfpxb, *pfpyb, *ah_bottom;
  /* 3-D */
  double *u, *v, *w, *ww, *uf, *vf, *advx, *advy, *drhox, *drhoy, *rho, *rmean, *t, *s, *km1, *viscofm;
  double *t1, *s1, *tf1, *sf1, *rho1, *rmean1, *tmean1, *smean1, *q2, *q2l, *q2f, *q2lf;
  double *l, *km, *kh, *kq, *wts, *viscofh;
  double *t1m1, *t1m2, *s1m1, *s1m2;   /* T1/S1 one and two RK stages back (TYPE_TSOBC 4) */
  /* scratch */
  double *sc_m1[8];
  double *sc_mk[4];
  double *sc_nk[4];
  int depth_fail;
} orc;

/* Column work arrays hold levels 0..KB; KB is limited to ORC_MAXKB, the
 * device library's own limit. */
#define ORC_MAXKB 64
#define ORC_COL (ORC_MAXKB + 2)

static const double ALPHA_RK[5] = {0.0, 0.2500, 1.0 / 3.0, 0.5000, 1.0};
static const double ONE_THIRD = 1.0 / 3.0;
static const double GRAV = 9.81;

#define G (o->g)
#define M1 (o->m1)
#define N1 (o->n1)
#define ND(a, i, k) (a)[(size_t)((k)-1) * M1 + (size_t)(i)]
#define EL3(a, i, k) (a)[(size_t)((k)-1) * N1 + (size_t)(i)]
#define NV(i, j) G->nv[(size_t)((j)-1) * N1 + (size_t)(i)]
#define NBE(i, j) G->nbe[(size_t)((j)-1) * N1 + (size_t)(i)]
#define A1U(i, j) G->a1u[(size_t)((j)-1) * N1 + (size_t)(i)]
#define A2U(i, j) G->a2u[(size_t)((j)-1) * N1 + (size_t)(i)]
#define AWX(i, j) G->awx[(size_t)((j)-1) * N1 + (size_t)(i)]
#define AWY(i, j) G->awy[(size_t)((j)-1) * N1 + (size_t)(i)]
#define IEC(i, j) G->iec[(size_t)((j)-1) * (size_t)(o->NE + 1) + (size_t)(i)]
#define IENODE(i, j) G->ienode[(size_t)((j)-1) * (size_t)(o->NE + 1) + (size_t)(i)]
#define NIEC(i, j) G->niec[(size_t)((j)-1) * (size_t)(o->NCV + 1) + (size_t)(i)]
#define DLTXNCVE(i, j) G->dltxncve[(size_t)((j)-1) * (size_t)(o->NCV + 1) + (size_t)(i)]
#define DLTYNCVE(i, j) G->dltyncve[(size_t)((j)-1) * (size_t)(o->NCV + 1) + (size_t)(i)]
#define NTSN(i) G->ntsn[i]
#define NTVE(i) G->ntve[i]
#define NBSN(i, j) G->nbsn[G->nbsn_off[i] + (j)-1]
#define NBVE(i, j) G->nbve[G->nbve_off[i] + (j)-1]
#define DLTXTRIE(i, j) G->dltxtrie[G->nbsn_off[i] + (j)-1]
#define DLTYTRIE(i, j) G->dltytrie[G->nbsn_off[i] + (j)-1]
#define DLTXECEC(i, j) G->dltxecec[G->nbve_off[i] + (j)-1]
#define DLTYECEC(i, j) G->dltyecec[G->nbve_off[i] + (j)-1]
#define DLTXNEC(i, j) G->dltxnec[G->nbve_off[i] + (j)-1]
#define DLTYNEC(i, j) G->dltynec[G->nbve_off[i] + (j)-1]
#define ADJC_OBC(i, j) G->adjc_obc[(size_t)((j)-1) * o->IOBCN + (i)]
#define FLUXF_OBC(i, j) G->fluxf_obc[(size_t)((j)-1) * o->IOBCN + (i)]
#define VQDIST(j, k) G->vqdist[(size_t)((k)-1) * o->NUMQBC + (j)]
/* edge rivers (RIVER_INFLOW_LOCATION = 'edge'): N_ICELLQ(J,1:2), RDISQ(J,1:2) */
#define N_ICELLQ(j, c) G->n_icellq[(size_t)((c)-1) * o->NUMQBC + (j)]
#define RDISQ(j, c) G->rdisq[(size_t)((c)-1) * o->NUMQBC + (j)]

static double fsign(double a, double b) { return (b >= 0.0) ? fabs(a) : -fabs(a); }
static double sq(double x) { return x * x; }

static double* zal(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

/* ----------------------------------------------------------------- setup */
orc* orc_create(const fv_model* g) {
  if (g->kb < 2 || g->kb > ORC_MAXKB) return NULL;
  orc* o = (orc*)calloc(1, sizeof(orc));
  o->g = g;
  o->p = g->p;
  o->M = g->m; o->N = g->n; o->NE = g->ne; o->NCV = g->ncv; o->KB = g->kb;
  o->KBM1 = g->kbm1; o->KBM2 = g->kbm2; o->IOBCN = g->iobcn; o->NUMQBC = g->numqbc;
  o->m1 = (size_t)g->m + 1; o->n1 = (size_t)g->n + 1;
  o->dte = g->p.dte; o->dti = g->p.dte * g->p.isplit;
  o->iint = g->p.iint0;
  o->int_time = g->p.time0;
  o->ext_time = g->p.time0;
  const size_t mk = o->m1 * o->KB, nk = o->n1 * o->KB;
#define AM(x) o->x = zal(o->m1)
#define AN(x) o->x = zal(o->n1)
#define AMK(x) o->x = zal(mk)
#define ANK(x) o->x = zal(nk)
  AM(el); AM(et); AM(d); AM(dt); AM(elf); AM(elrk); AM(dtfa); AM(egf); AM(elm1); AM(elm2);
  for (int i = 0; i < o->IOBCN; ++i) if (g->type_obc[i] >= 9) o->has_ore = 1;
  AN(ua); AN(va); AN(el1); AN(et1); AN(d1); AN(dt1); AN(elf1); AN(elrk1); AN(uaf); AN(vaf);
  AN(uark); AN(vark); AN(uard); AN(vard); AN(adx2d); AN(ady2d); AN(drx2d); AN(dry2d);
  AN(advua); AN(advva); AN(pstx); AN(psty); AN(fluxobn); AN(cbc); AN(wubot); AN(wvbot);
  AN(wusurf); AN(wvsurf); AN(wusurf2); AN(wvsurf2); AN(tps);
  AM(pfpxb); AM(pfpyb); AM(ah_bottom);
  ANK(u); ANK(v); ANK(w); ANK(ww); ANK(uf); ANK(vf); ANK(advx); ANK(advy); ANK(drhox); ANK(drhoy);
  ANK(rho); ANK(rmean); ANK(t); ANK(s); ANK(km1); ANK(viscofm);
  AMK(t1); AMK(s1); AMK(tf1); AMK(sf1); AMK(rho1); AMK(rmean1); AMK(tmean1); AMK(smean1);
  AMK(q2); AMK(q2l); AMK(q2f); AMK(q2lf); AMK(l); AMK(km); AMK(kh); AMK(kq); AMK(wts); AMK(viscofh);
  AMK(t1m1); AMK(t1m2); AMK(s1m1); AMK(s1m2);
  for (int i = 0; i < 8; ++i) o->sc_m1[i] = zal(o->m1 > o->n1 ? o->m1 : o->n1);
  for (int i = 0; i < 4; ++i) o->sc_mk[i] = zal(mk);
  for (int i = 0; i < 4; ++i) o->sc_nk[i] = zal(nk);
  o->uard_obcn = zal(o->IOBCN); o->xflux_obcn = zal(o->IOBCN);
  o->xflux_obc = zal((size_t)o->IOBCN * o->KB);
  o->qdis = zal(o->NUMQBC);
  /* state from the model */
#define CPY(x, n) memcpy(o->x, g->x, sizeof(double) * (n))
  CPY(el, o->m1); CPY(et, o->m1); CPY(d, o->m1); CPY(dt, o->m1); CPY(elm1, o->m1); CPY(elm2, o->m1);
  CPY(ua, o->n1); CPY(va, o->n1); CPY(el1, o->n1); CPY(et1, o->n1); CPY(d1, o->n1); CPY(dt1, o->n1);
  CPY(u, nk); CPY(v, nk); CPY(w, nk); CPY(ww, nk); CPY(t, nk); CPY(s, nk); CPY(rho, nk);
  CPY(rmean, nk); CPY(km1, nk);
  CPY(t1, mk); CPY(s1, mk); CPY(rho1, mk); CPY(rmean1, mk); CPY(tmean1, mk); CPY(smean1, mk);
  CPY(q2, mk); CPY(q2l, mk); CPY(l, mk); CPY(km, mk); CPY(kh, mk); CPY(kq, mk); CPY(wts, mk);
  memcpy(o->dtfa, o->d, sizeof(double) * o->m1);
  return o;
}

void orc_destroy(orc* o) {
  if (!o) return;
  double** f[] = {&o->el, &o->elm1, &o->elm2, &o->et, &o->d, &o->dt, &o->elf, &o->elrk, &o->dtfa, &o->egf, &o->ua, &o->va,
                  &o->el1, &o->et1, &o->d1, &o->dt1, &o->elf1, &o->elrk1, &o->uaf, &o->vaf, &o->uark,
                  &o->vark, &o->uard, &o->vard, &o->adx2d, &o->ady2d, &o->drx2d, &o->dry2d, &o->advua,
                  &o->advva, &o->pstx, &o->psty, &o->fluxobn, &o->cbc, &o->wubot, &o->wvbot, &o->wusurf,
                  &o->wvsurf, &o->wusurf2, &o->wvsurf2, &o->tps, &o->uard_obcn, &o->xflux_obcn,
                  &o->xflux_obc, &o->qdis, &o->pfpxb, &o->pfpyb, &o->ah_bottom, &o->u, &o->v, &o->w,
                  &o->ww, &o->uf, &o->vf, &o->advx, &o->advy, &o->drhox, &o->drhoy, &o->rho, &o->rmean,
                  &o->t, &o->s, &o->km1, &o->viscofm, &o->t1, &o->s1, &o->tf1, &o->sf1, &o->rho1,
                  &o->rmean1, &o->tmean1, &o->smean1, &o->q2, &o->q2l, &o->q2f, &o->q2lf, &o->l, &o->km,
                  &o->kh, &o->kq, &o->wts, &o->viscofh, &o->t1m1, &o->t1m2, &o->s1m1, &o->s1m2};
  for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i) free(*f[i]);
  for (int i = 0; i < 8; ++i) free(o->sc_m1[i]);
  for (int i = 0; i < 4; ++i) { free(o->sc_mk[i]); free(o->sc_nk[i]); }
  free(o);
}

/* name -> (pointer, length) for get/put */
static double* orc_lookup(orc* o, const char* name, size_t* len) {
  const size_t mk = o->m1 * o->KB, nk = o->n1 * o->KB;
  struct { const char* n; double* p; size_t l; } t[] = {
      {"el", o->el, o->m1}, {"et", o->et, o->m1}, {"d", o->d, o->m1}, {"dt", o->dt, o->m1},
      {"elf", o->elf, o->m1}, {"dtfa", o->dtfa, o->m1}, {"egf", o->egf, o->m1},
      {"elm1", o->elm1, o->m1}, {"elm2", o->elm2, o->m1},
      {"ua", o->ua, o->n1}, {"va", o->va, o->n1}, {"el1", o->el1, o->n1}, {"et1", o->et1, o->n1},
      {"d1", o->d1, o->n1}, {"dt1", o->dt1, o->n1}, {"uard", o->uard, o->n1}, {"vard", o->vard, o->n1},
      {"adx2d", o->adx2d, o->n1}, {"ady2d", o->ady2d, o->n1}, {"drx2d", o->drx2d, o->n1},
      {"dry2d", o->dry2d, o->n1}, {"advua", o->advua, o->n1}, {"advva", o->advva, o->n1},
      {"pstx", o->pstx, o->n1}, {"psty", o->psty, o->n1}, {"cbc", o->cbc, o->n1},
      {"wubot", o->wubot, o->n1}, {"wvbot", o->wvbot, o->n1}, {"wusurf", o->wusurf, o->n1},
      {"wvsurf", o->wvsurf, o->n1}, {"wusurf2", o->wusurf2, o->n1}, {"wvsurf2", o->wvsurf2, o->n1},
      {"uaf", o->uaf, o->n1}, {"vaf", o->vaf, o->n1}, {"uark", o->uark, o->n1}, {"vark", o->vark, o->n1},
      {"elrk1", o->elrk1, o->n1}, {"elf1", o->elf1, o->n1}, {"elrk", o->elrk, o->m1},
      {"u", o->u, nk}, {"v", o->v, nk}, {"w", o->w, nk}, {"ww", o->ww, nk}, {"uf", o->uf, nk},
      {"vf", o->vf, nk}, {"advx", o->advx, nk}, {"advy", o->advy, nk}, {"drhox", o->drhox, nk},
      {"drhoy", o->drhoy, nk}, {"rho", o->rho, nk}, {"rmean", o->rmean, nk}, {"t", o->t, nk},
      {"s", o->s, nk}, {"km1", o->km1, nk}, {"viscofm", o->viscofm, nk},
      {"t1", o->t1, mk}, {"s1", o->s1, mk}, {"tf1", o->tf1, mk}, {"sf1", o->sf1, mk},
      {"rho1", o->rho1, mk}, {"rmean1", o->rmean1, mk}, {"tmean1", o->tmean1, mk},
      {"smean1", o->smean1, mk}, {"q2", o->q2, mk}, {"q2l", o->q2l, mk}, {"q2f", o->q2f, mk},
      {"q2lf", o->q2lf, mk}, {"l", o->l, mk}, {"km", o->km, mk}, {"kh", o->kh, mk},
      {"kq", o->kq, mk}, {"wts", o->wts, mk}, {"viscofh", o->viscofh, mk},
      {"t1m1", o->t1m1, mk}, {"t1m2", o->t1m2, mk}, {"s1m1", o->s1m1, mk}, {"s1m2", o->s1m2, mk},
      {"ramp", &o->ramp, 1},
      {"uard_obcn", o->uard_obcn, (size_t)o->IOBCN},
      {"xflux_obc", o->xflux_obc, (size_t)o->IOBCN * o->KB},
  };
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i)
    if (strcmp(t[i].n, name) == 0) { *len = t[i].l; return t[i].p; }
  return NULL;
}

int orc_get(orc* o, const char* name, double* out, long long len) {
  size_t l;
  double* p = orc_lookup(o, name, &l);
  if (!p || (size_t)len != l) return 1;
  memcpy(out, p, l * sizeof(double));
  return 0;
}

int orc_put(orc* o, const char* name, const double* in, long long len) {
  size_t l;
  double* p = orc_lookup(o, name, &l);
  if (!p || (size_t)len != l) return 1;
  memcpy(p, in, l * sizeof(double));
  return 0;
}

long long orc_iint(orc* o) { return o->iint; }
int orc_depth_fail(orc* o) { return o->depth_fail; }

/* ------------------------------------------------------ small helpers */
/* N2E2D / N2E3D, FVCOM_source/mod_main.F:1633-1664 */
static void n2e2d(orc* o, const double* nv_, double* ev) {
  for (int i = 1; i <= o->N; ++i) ev[i] = ONE_THIRD * (nv_[NV(i, 1)] + nv_[NV(i, 2)] + nv_[NV(i, 3)]);
}
static void n2e3d(orc* o, const double* nv_, double* ev) {
  for (int k = 1; k <= o->KB; ++k)
    for (int i = 1; i <= o->N; ++i)
      EL3(ev, i, k) = ONE_THIRD * (ND(nv_, NV(i, 1), k) + ND(nv_, NV(i, 2), k) + ND(nv_, NV(i, 3), k));
}

/* PSIMPLE_DRAG, FVCOM_source/mod_bulk.F:58-90 (CD=1.2E-3 is a default-real literal) */
static void simple_drag(double tx, double ty, double* sx, double* sy) {
  double wds = sqrt(tx * tx + ty * ty);
  double cd = (double)1.2E-3f;
  if (wds >= 11.0) cd = (0.49 + 0.065 * wds) * 1.E-3;
  if (wds >= 25.0) cd = (0.49 + 0.065 * 25.0) * 1.E-3;
  *sx = 1.2 * cd * tx * wds;
  *sy = 1.2 * cd * ty * wds;
}

/* ----------------------------------------------------- BCOND_GCN cases */
/* Wall projection of one velocity pair for ISBCE==1 elements, with the
 * river-mouth exception; FVCOM_source/bcond_gcn.F:325-420 (case 2),
 * :424-530 (case 3), :537-616 (case 5) share this body. */
static void wall_project(orc* o, int i, double* pu, double* pv) {
  if (G->isbce[i] == 3) { *pu = 0.0; *pv = 0.0; }
  if (G->isbce[i] != 1) return;
  double alpha1 = G->alpha[i];
  for (int j = 0; j < o->NUMQBC; ++j) {
    int mouth;
    if (G->icellq[j] > 0) {          /* edge river: its cell */
      mouth = i == G->icellq[j];
    } else {                         /* node river: the first and last cells of its node */
      int i1 = G->inodeq[j];
      int j1 = NBVE(i1, 1), j2 = NBVE(i1, NTVE(i1));
      mouth = i == j1 || i == j2;
    }
    if (mouth) {
      double aq = G->angleq[j];
      double untmp = *pu * cos(aq) + *pv * sin(aq);
      double vntmp = -*pu * sin(aq) + *pv * cos(aq);
      untmp = fmax(untmp, 0.0);
      *pu = untmp * cos(aq) - vntmp * sin(aq);
      *pv = untmp * sin(aq) + vntmp * cos(aq);
      return;
    }
  }
  double ui = *pu * sq(sin(alpha1)) - *pv * sin(alpha1) * cos(alpha1);
  double vi = -*pu * sin(alpha1) * cos(alpha1) + *pv * sq(cos(alpha1));
  *pu = ui;
  *pv = vi;
}

/* BCOND_GCN(8,0): internal-mode surface forcing, bcond_gcn.F:624-700 */
static void bcond_forcing_int(orc* o) {
  for (int j = 0; j < o->NUMQBC; ++j) o->qdis[j] = G->qdis0[j] * o->ramp;
  if (o->p.wind_on) {
    for (int i = 1; i <= o->N; ++i) {
      double sx, sy;
      simple_drag(G->uuwind[i], G->vvwind[i], &sx, &sy);
      o->wusurf[i] = -sx * o->ramp * 0.001;
      o->wvsurf[i] = -sy * o->ramp * 0.001;
    }
  }
}

/* BCOND_GCN(9,0): external-mode surface forcing, bcond_gcn.F:851-905.  The
 * wind speed at ExtTime is interpolated linearly in time between its values
 * at the start (UUWIND0) and the end (UUWIND) of the internal step, as
 * UPDATE_WIND interpolates its records (bcond_gcn.F:878-885). */
static void bcond_forcing_ext(orc* o) {
  if (o->p.wind_on) {
    const double frac = (double)o->iext / (double)o->p.isplit;
    for (int i = 1; i <= o->N; ++i) {
      double sx, sy;
      const double uu = G->uuwind0[i] + frac * (G->uuwind[i] - G->uuwind0[i]);
      const double vv = G->vvwind0[i] + frac * (G->vvwind[i] - G->vvwind0[i]);
      simple_drag(uu, vv, &sx, &sy);
      o->wusurf2[i] = sx * o->ramp * 0.001;
      o->wvsurf2[i] = sy * o->ramp * 0.001;
    }
  }
}

/* ------------------------------------------------------- external mode */
/* EXTEL_EDGE(K), FVCOM_source/extel_edge.F:32-229 */
static void extel_edge(orc* o, int K) {
  double* xflux = o->sc_m1[0];
  memset(xflux, 0, sizeof(double) * M1);
  for (int i = 1; i <= o->NCV; ++i) {
    int i1 = G->ntrg[i], ia = NIEC(i, 1), ib = NIEC(i, 2);
    double dij = o->d1[i1], uij = o->ua[i1], vij = o->va[i1];
    double exflux = dij * (-uij * G->dltye[i] + vij * G->dltxe[i]);
    xflux[ia] = xflux[ia] - exflux;
    xflux[ib] = xflux[ib] + exflux;
  }
  for (int i = 0; i < o->IOBCN; ++i) {
    o->xflux_obcn[i] = xflux[G->i_obc_n[i]];
    xflux[G->i_obc_n[i]] = 0.0;
  }
  for (int j = 0; j < o->NUMQBC; ++j) {
    if (G->icellq[j] > 0) {
      int j1 = N_ICELLQ(j, 1), j2 = N_ICELLQ(j, 2);
      xflux[j1] = xflux[j1] - o->qdis[j] * RDISQ(j, 1);
      xflux[j2] = xflux[j2] - o->qdis[j] * RDISQ(j, 2);
    } else {
      int jj = G->inodeq[j];
      xflux[jj] = xflux[jj] - o->qdis[j];
    }
  }
  double dtk = ALPHA_RK[K] * o->dte;
  for (int i = 0; i <= o->M; ++i) o->elf[i] = o->elrk[i] - dtk * xflux[i] / G->art1[i];
}

/* BCOND_GCN(1,K) -> BCOND_ASL (spectral), mod_obcs.F:509-592, on the
 * nodes of TYPE_OBC 1/2 */
static void bcond_asl(orc* o) {
  if (!o->p.obc_elevation_on) return;
  const double PI2 = 6.283185307179586;
  double elapsed = o->rk_time;
  for (int i = 0; i < o->IOBCN; ++i) {
    if (G->type_obc[i] > 2) continue;
    int ii = G->i_obc_n[i];
    double force = 0.0;
    for (int j = 0; j < G->ntide; ++j) {
      double phai_ij = G->phai[(size_t)j * o->IOBCN + i] * PI2 / 360.0;
      double time_flt = elapsed * PI2 / G->period[j];
      force = G->apt[(size_t)j * o->IOBCN + i] * fv_cos(time_flt - phai_ij) + force;
    }
    force = force + G->emean[i];
    o->elf[ii] = force * o->ramp;
  }
}

/* The rest of BCOND_GCN(1,K) (bcond_gcn.F:117-122): BCOND_ASL_CLP on the
 * nodes of TYPE_OBC 3/4 (mod_obcs.F:599-617), BCOND_GWI on 5/6
 * (mod_obcs.F:624-658) and BCOND_BKI on 7/8 (mod_obcs.F:665-699), the
 * radiation conditions taking the interior neighbour's new ELF, and
 * BCOND_ORE on 9/10 (mod_obcs.F:706-737) from EL, ELM1 and ELM2. */
static void bcond_obc_other(orc* o, int K) {
  for (int jn = 0; jn < o->IOBCN; ++jn) {
    const int ty = G->type_obc[jn], i1 = G->i_obc_n[jn], i2 = G->next_obc[jn];
    if (ty == 3 || ty == 4) {
      o->elf[i1] = 0.0;
    } else if (ty >= 9) {
      int i3 = G->next_obc2[jn];
      double cl = (o->elm2[i2] - o->el[i2]) / (o->el[i2] + o->elm2[i2] - 2.0 * o->elm1[i3]);
      double mu;
      if (cl >= 1.0) mu = 1.0;
      else if (cl > 0.0 && cl < 1.0) mu = cl;
      else mu = 0.0;
      o->elf[i1] = (o->elm1[i1] * (1.0 - mu) + 2.0 * mu * o->el[i2]) / (1.0 + mu);
    } else if (ty >= 5) {
      double cc = sqrt(G->grav_n[i1] * G->h[i1]) * o->dte / G->dltn_obc[jn] * ALPHA_RK[K];
      double cp = cc + 1.0;
      if (ty <= 6) o->elf[i1] = (cc * o->elf[i2] + o->elrk[i1]) / cp;
      else o->elf[i1] = (cc * o->elf[i2] + o->elrk[i1] * (1.0 - o->dte / 10800.0)) / cp;
    }
  }
}

/* FLUX_OBN(K), mod_obcs.F:1181-1215 */
static void flux_obn(orc* o, int K) {
  memset(o->fluxobn, 0, sizeof(double) * N1);
  for (int i = 0; i < o->IOBCN; ++i) {
    int j = G->i_obc_n[i];
    double flux = -(o->elf[j] - o->elrk[j]) * G->art1[j] / (ALPHA_RK[K] * o->dte) - o->xflux_obcn[i];
    if (G->nadjc_obc[i] > 0) {
      int c1 = ADJC_OBC(i, 1);
      o->fluxobn[c1] = o->fluxobn[c1] + FLUXF_OBC(i, 1) * flux;
    }
    if (G->nadjc_obc[i] > 1) {
      int c2 = ADJC_OBC(i, 2);
      o->fluxobn[c2] = o->fluxobn[c2] + FLUXF_OBC(i, 2) * flux;
    }
  }
}

/* ADVAVE_EDGE_GCN(XFLUX,YFLUX), FVCOM_source/advave_edge_gcn.F:32-1145,
 * LIMITED_NO edge loop :546-884, OBC :915-925, rivers :971-992 */
static void advave_edge_gcn(orc* o, double* xflux, double* yflux) {
  const double fact = o->p.horizontal_closure ? 1.0 : 0.0, fm1 = 1.0 - fact;
  memset(xflux, 0, sizeof(double) * N1);
  memset(yflux, 0, sizeof(double) * N1);
  memset(o->pstx, 0, sizeof(double) * N1);
  memset(o->psty, 0, sizeof(double) * N1);
  const double* ua = o->ua; const double* va = o->va;
  for (int i = 1; i <= o->NE; ++i) {
    int ia = IEC(i, 1), ib = IEC(i, 2), j1 = IENODE(i, 1), j2 = IENODE(i, 2);
    double dij = 0.5 * (o->d[j1] + o->d[j2]);
    double elij = 0.5 * (o->el[j1] + o->el[j2]);
    int k1 = NBE(ia, 1), k2 = NBE(ia, 2), k3 = NBE(ia, 3);
    double cofa1 = A1U(ia, 1) * ua[ia] + A1U(ia, 2) * ua[k1] + A1U(ia, 3) * ua[k2] + A1U(ia, 4) * ua[k3];
    double cofa2 = A2U(ia, 1) * ua[ia] + A2U(ia, 2) * ua[k1] + A2U(ia, 3) * ua[k2] + A2U(ia, 4) * ua[k3];
    double cofa5 = A1U(ia, 1) * va[ia] + A1U(ia, 2) * va[k1] + A1U(ia, 3) * va[k2] + A1U(ia, 4) * va[k3];
    double cofa6 = A2U(ia, 1) * va[ia] + A2U(ia, 2) * va[k1] + A2U(ia, 3) * va[k2] + A2U(ia, 4) * va[k3];
    double xij = G->xijc[i] - G->xc[ia], yij = G->yijc[i] - G->yc[ia];
    double uij1 = ua[ia] + cofa1 * xij + cofa2 * yij;
    double vij1 = va[ia] + cofa5 * xij + cofa6 * yij;
    k1 = NBE(ib, 1); k2 = NBE(ib, 2); k3 = NBE(ib, 3);
    double cofa3 = A1U(ib, 1) * ua[ib] + A1U(ib, 2) * ua[k1] + A1U(ib, 3) * ua[k2] + A1U(ib, 4) * ua[k3];
    double cofa4 = A2U(ib, 1) * ua[ib] + A2U(ib, 2) * ua[k1] + A2U(ib, 3) * ua[k2] + A2U(ib, 4) * ua[k3];
    double cofa7 = A1U(ib, 1) * va[ib] + A1U(ib, 2) * va[k1] + A1U(ib, 3) * va[k2] + A1U(ib, 4) * va[k3];
    double cofa8 = A2U(ib, 1) * va[ib] + A2U(ib, 2) * va[k1] + A2U(ib, 3) * va[k2] + A2U(ib, 4) * va[k3];
    xij = G->xijc[i] - G->xc[ib]; yij = G->yijc[i] - G->yc[ib];
    double uij2 = ua[ib] + cofa3 * xij + cofa4 * yij;
    double vij2 = va[ib] + cofa7 * xij + cofa8 * yij;
    double uij = 0.5 * (uij1 + uij2), vij = 0.5 * (vij1 + vij2);
    double un_tmp = -uij * G->dltyc[i] + vij * G->dltxc[i];
    double viscof1 = G->art[ia] * sqrt(sq(cofa1) + sq(cofa6) + 0.5 * sq(cofa2 + cofa5));
    double viscof2 = G->art[ib] * sqrt(sq(cofa3) + sq(cofa8) + 0.5 * sq(cofa4 + cofa7));
    double viscof = (fact * 0.5 * (viscof1 * G->cc_hvc[ia] + viscof2 * G->cc_hvc[ib]) +
                     fm1 * 0.5 * (G->cc_hvc[ia] + G->cc_hvc[ib])) / o->p.hprnu;
    double txxij = (cofa1 + cofa3) * viscof;
    double tyyij = (cofa6 + cofa8) * viscof;
    double txyij = 0.5 * (cofa2 + cofa4 + cofa5 + cofa7) * viscof;
    double fxx = dij * (txxij * G->dltyc[i] - txyij * G->dltxc[i]);
    double fyy = dij * (txyij * G->dltyc[i] - tyyij * G->dltxc[i]);
    double sgn = fsign(1.0, un_tmp);
    double xadv = dij * un_tmp * ((1.0 - sgn) * uij2 + (1.0 + sgn) * uij1) * 0.5;
    double yadv = dij * un_tmp * ((1.0 - sgn) * vij2 + (1.0 + sgn) * vij1) * 0.5;
    double isbc_tmp = (double)G->isbc[i];
    xflux[ia] = xflux[ia] + (xadv + fxx * G->epor[ia]) * (1.0 - isbc_tmp) * (double)G->iucp[ia];
    yflux[ia] = yflux[ia] + (yadv + fyy * G->epor[ia]) * (1.0 - isbc_tmp) * (double)G->iucp[ia];
    xflux[ib] = xflux[ib] - (xadv + fxx * G->epor[ib]) * (1.0 - isbc_tmp) * (double)G->iucp[ib];
    yflux[ib] = yflux[ib] - (yadv + fyy * G->epor[ib]) * (1.0 - isbc_tmp) * (double)G->iucp[ib];
    o->pstx[ia] = o->pstx[ia] - G->grav_e[ia] * o->d1[ia] * elij * G->dltyc[i];
    o->psty[ia] = o->psty[ia] + G->grav_e[ia] * o->d1[ia] * elij * G->dltxc[i];
    o->pstx[ib] = o->pstx[ib] + G->grav_e[ib] * o->d1[ib] * elij * G->dltyc[i];
    o->psty[ib] = o->psty[ib] - G->grav_e[ib] * o->d1[ib] * elij * G->dltxc[i];
  }
  for (int i = 1; i <= o->N; ++i) {
    if (G->isbce[i] == 2) {
      xflux[i] = (xflux[i] + o->fluxobn[i] * ua[i]) * (double)G->iucp[i];
      yflux[i] = (yflux[i] + o->fluxobn[i] * va[i]) * (double)G->iucp[i];
    }
  }
  for (int k = 0; k < o->NUMQBC; ++k) {
    if (G->icellq[k] > 0) {
      int i1 = G->icellq[k];
      double vlctyq = o->qdis[k] / G->qarea[k];
      double temp = o->qdis[k] * vlctyq;
      xflux[i1] = xflux[i1] - temp * cos(G->angleq[k]);
      yflux[i1] = yflux[i1] - temp * sin(G->angleq[k]);
      continue;
    }
    int j = G->inodeq[k];
    int i1 = NBVE(j, 1), i2 = NBVE(j, NTVE(j));
    double vlctyq = o->qdis[k] / G->qarea[k];
    xflux[i1] = xflux[i1] - 0.5 * o->qdis[k] * vlctyq * cos(G->angleq[k]);
    yflux[i1] = yflux[i1] - 0.5 * o->qdis[k] * vlctyq * sin(G->angleq[k]);
    xflux[i2] = xflux[i2] - 0.5 * o->qdis[k] * vlctyq * cos(G->angleq[k]);
    yflux[i2] = yflux[i2] - 0.5 * o->qdis[k] * vlctyq * sin(G->angleq[k]);
  }
}

/* EXTUV_EDGE(K), FVCOM_source/extuv_edge.F:33-196 (H1(0)=H1(1) quirk kept) */
static void extuv_edge(orc* o, int K) {
  const double* h1 = G->h1;
  for (int i = 1; i <= o->N; ++i) {
    double resx = o->adx2d[i] + o->advua[i] + o->drx2d[i] + o->pstx[i] - G->cor[i] * o->va[i] * o->d1[i] * G->art[i] -
                  (o->wusurf2[i] + o->wubot[i]) * G->art[i];
    double resy = o->ady2d[i] + o->advva[i] + o->dry2d[i] + o->psty[i] + G->cor[i] * o->ua[i] * o->d1[i] * G->art[i] -
                  (o->wvsurf2[i] + o->wvbot[i]) * G->art[i];
    o->uaf[i] = (o->uark[i] * (h1[i] + o->elrk1[i]) - ALPHA_RK[K] * o->dte * resx / G->art[i]) / (h1[i] + o->elf1[i]);
    o->vaf[i] = (o->vark[i] * (h1[i] + o->elrk1[i]) - ALPHA_RK[K] * o->dte * resy / G->art[i]) / (h1[i] + o->elf1[i]);
  }
  for (int i = 1; i <= o->N; ++i) {
    o->uaf[i] = o->uaf[i] / (1.0 + G->cc_sponge[i] * o->uaf[i] * o->uaf[i]);
    o->vaf[i] = o->vaf[i] / (1.0 + G->cc_sponge[i] * o->vaf[i] * o->vaf[i]);
  }
}

/* EXTERNAL_STEP, FVCOM_source/external_step.F:30-443 */
void orc_external_step(orc* o) {
  if (o->p.iramp != 0) {
    double tmp = (double)(o->iint - 1) + (double)o->iext / (double)o->p.isplit;
    o->ramp = fv_tanh(tmp / (double)o->p.iramp);
  } else o->ramp = 1.0;
  bcond_forcing_ext(o);
  memcpy(o->elrk1, o->el1, sizeof(double) * N1);
  memcpy(o->elrk, o->el, sizeof(double) * M1);
  memcpy(o->uark, o->ua, sizeof(double) * N1);
  memcpy(o->vark, o->va, sizeof(double) * N1);
  for (int K = 1; K <= 4; ++K) {
    o->rk_time = o->ext_time + o->dte * (ALPHA_RK[K] - 1.0);
    extel_edge(o, K);
    bcond_asl(o);
    bcond_obc_other(o, K);
    /* the RK blend of the elevation-specified nodes (external_step.F:262-266) */
    for (int i = 0; i < o->IOBCN; ++i) {
      if (G->type_obc[i] > 2) continue;
      int j = G->i_obc_n[i];
      o->elf[j] = o->elrk[j] + ALPHA_RK[K] * (o->elf[j] - o->elrk[j]);
    }
    n2e2d(o, o->elf, o->elf1);
    flux_obn(o, K);
    advave_edge_gcn(o, o->advua, o->advva);
    extuv_edge(o, K);
    for (int i = 1; i <= o->N; ++i) wall_project(o, i, &o->uaf[i], &o->vaf[i]);  /* BCOND_GCN(2,K) */
    if (o->has_ore) {   /* ASSIGN_ELM1_TO_ELM2 (external_step.F:333, mod_obcs.F:187-188) */
      memcpy(o->elm2, o->elm1, sizeof(double) * M1);
      memcpy(o->elm1, o->el, sizeof(double) * M1);
    }
    if (o->p.type_tsobc == 4) {   /* its T/S part (mod_obcs.F:190-196) */
      const size_t mk = M1 * o->KB;
      memcpy(o->t1m2, o->t1m1, sizeof(double) * mk); memcpy(o->t1m1, o->t1, sizeof(double) * mk);
      memcpy(o->s1m2, o->s1m1, sizeof(double) * mk); memcpy(o->s1m1, o->s1, sizeof(double) * mk);
    }
    memcpy(o->el, o->elf, sizeof(double) * M1);
    memcpy(o->el1, o->elf1, sizeof(double) * N1);
    n2e2d(o, o->el, o->el1);
    for (int i = 0; i <= o->M; ++i) o->d[i] = G->h[i] + o->el[i];
    for (int i = 0; i <= o->N; ++i) o->d1[i] = G->h1[i] + o->el1[i];
    memcpy(o->ua, o->uaf, sizeof(double) * N1);
    memcpy(o->va, o->vaf, sizeof(double) * N1);
    memcpy(o->dtfa, o->d, sizeof(double) * M1);
    /* DEPTH_CHECK, FVCOM_source/depth_check.F */
    for (int i = 1; i <= o->M; ++i) if (o->d[i] < o->p.min_depth) o->depth_fail = 1;
    if (K == 3) {
      for (int i = 0; i <= o->N; ++i) {
        o->uard[i] = o->uard[i] + o->ua[i] * o->d1[i];
        o->vard[i] = o->vard[i] + o->va[i] * o->d1[i];
      }
      for (int i = 0; i <= o->M; ++i) o->egf[i] = o->egf[i] + o->el[i] / (double)o->p.isplit;
    }
    if (K == 4) {
      for (int i = 0; i < o->IOBCN; ++i) {
        int j = G->i_obc_n[i];
        double tmp = -(o->elf[j] - o->elrk[j]) * G->art1[j] / o->dte - o->xflux_obcn[i];
        o->uard_obcn[i] = o->uard_obcn[i] + tmp / (double)o->p.isplit;
      }
    }
  }
}

/* -------------------------------------------------------- 3-D routines */
/* ADJUST2D3D, FVCOM_source/adjust2d3d.F:28-116 */
static void adjust2d3d(orc* o, int type) {
  const int KBM1 = o->KBM1;
  if (type == 1) {
    for (int i = 1; i <= o->N; ++i) {
      double uac = 0.0, vac = 0.0;
      for (int k = 1; k <= KBM1; ++k) {
        uac = uac + EL3(o->u, i, k) * EL3(G->dz1, i, k);
        vac = vac + EL3(o->v, i, k) * EL3(G->dz1, i, k);
      }
      for (int k = 1; k <= KBM1; ++k) {
        EL3(o->u, i, k) = EL3(o->u, i, k) + 1.0 * (o->ua[i] - uac);
        EL3(o->v, i, k) = EL3(o->v, i, k) + 1.0 * (o->va[i] - vac);
      }
    }
  } else {
    for (int i = 0; i <= o->N; ++i) {
      o->uard[i] = o->uard[i] / (double)o->p.isplit;
      o->vard[i] = o->vard[i] / (double)o->p.isplit;
    }
    for (int i = 1; i <= o->N; ++i) {
      double utmp = 0.0, vtmp = 0.0;
      for (int k = 1; k <= KBM1; ++k) {
        utmp = utmp + EL3(o->u, i, k) * EL3(G->dz1, i, k);
        vtmp = vtmp + EL3(o->v, i, k) * EL3(G->dz1, i, k);
      }
      utmp = utmp * o->dt1[i];
      vtmp = vtmp * o->dt1[i];
      for (int k = 1; k <= KBM1; ++k) {
        EL3(o->u, i, k) = EL3(o->u, i, k) - (utmp - o->uard[i]) / o->dt1[i];
        EL3(o->v, i, k) = EL3(o->v, i, k) - (vtmp - o->vard[i]) / o->dt1[i];
      }
    }
  }
}

/* BOTTOM_ROUGHNESS (BR_ORIG), FVCOM_source/brough.F:41-215 */
static void bottom_roughness(orc* o) {
  const double vk2 = .160;
  const int KB = o->KB, KBM1 = o->KBM1;
  for (int i = 1; i <= o->N; ++i) {
    double zb = EL3(G->zz1, i, KBM1) - EL3(G->z1, i, KB);
    if (o->dt1[i] > 3.0) o->cbc[i] = vk2 / sq(fv_log(zb * o->dt1[i] / G->cc_z0b[i]));
    else o->cbc[i] = vk2 / sq(fv_log(zb * 3.0 / G->cc_z0b[i]));
    if (o->cbc[i] < o->p.cbcmin) o->cbc[i] = o->p.cbcmin;
  }
  for (int i = 1; i <= o->N; ++i) {
    double ub = EL3(o->u, i, KBM1), vb = EL3(o->v, i, KBM1);
    double btps = o->cbc[i] * sqrt(ub * ub + vb * vb);
    o->wubot[i] = -btps * ub;
    o->wvbot[i] = -btps * vb;
  }
}

/* ADVECTION_EDGE_GCN(ADVX,ADVY), FVCOM_source/advection_edge_gcn.F:29-734 */
static void advection_edge_gcn(orc* o, double* xflux, double* yflux) {
  const double fact = o->p.horizontal_closure ? 1.0 : 0.0, fm1 = 1.0 - fact;
  const int KBM1 = o->KBM1;
  memset(xflux, 0, sizeof(double) * N1 * o->KB);
  memset(yflux, 0, sizeof(double) * N1 * o->KB);
  for (int i = 1; i <= o->NE; ++i) {
    int ia = IEC(i, 1), ib = IEC(i, 2), j1 = IENODE(i, 1), j2 = IENODE(i, 2);
    int k1 = NBE(ia, 1), k2 = NBE(ia, 2), k3 = NBE(ia, 3), k4 = NBE(ib, 1), k5 = NBE(ib, 2), k6 = NBE(ib, 3);
    for (int k = 1; k <= KBM1; ++k) {
      double xija = G->xijc[i] - G->xc[ia], yija = G->yijc[i] - G->yc[ia];
      double xijb = G->xijc[i] - G->xc[ib], yijb = G->yijc[i] - G->yc[ib];
      double dij = 0.5 * (o->dt[j1] * ND(G->dz, j1, k) + o->dt[j2] * ND(G->dz, j2, k));
      const double* U = o->u; const double* V = o->v;
      double cofa1 = A1U(ia, 1) * EL3(U, ia, k) + A1U(ia, 2) * EL3(U, k1, k) + A1U(ia, 3) * EL3(U, k2, k) + A1U(ia, 4) * EL3(U, k3, k);
      double cofa2 = A2U(ia, 1) * EL3(U, ia, k) + A2U(ia, 2) * EL3(U, k1, k) + A2U(ia, 3) * EL3(U, k2, k) + A2U(ia, 4) * EL3(U, k3, k);
      double cofa5 = A1U(ia, 1) * EL3(V, ia, k) + A1U(ia, 2) * EL3(V, k1, k) + A1U(ia, 3) * EL3(V, k2, k) + A1U(ia, 4) * EL3(V, k3, k);
      double cofa6 = A2U(ia, 1) * EL3(V, ia, k) + A2U(ia, 2) * EL3(V, k1, k) + A2U(ia, 3) * EL3(V, k2, k) + A2U(ia, 4) * EL3(V, k3, k);
      double uij1 = EL3(U, ia, k) + cofa1 * xija + cofa2 * yija;
      double vij1 = EL3(V, ia, k) + cofa5 * xija + cofa6 * yija;
      double cofa3 = A1U(ib, 1) * EL3(U, ib, k) + A1U(ib, 2) * EL3(U, k4, k) + A1U(ib, 3) * EL3(U, k5, k) + A1U(ib, 4) * EL3(U, k6, k);
      double cofa4 = A2U(ib, 1) * EL3(U, ib, k) + A2U(ib, 2) * EL3(U, k4, k) + A2U(ib, 3) * EL3(U, k5, k) + A2U(ib, 4) * EL3(U, k6, k);
      double cofa7 = A1U(ib, 1) * EL3(V, ib, k) + A1U(ib, 2) * EL3(V, k4, k) + A1U(ib, 3) * EL3(V, k5, k) + A1U(ib, 4) * EL3(V, k6, k);
      double cofa8 = A2U(ib, 1) * EL3(V, ib, k) + A2U(ib, 2) * EL3(V, k4, k) + A2U(ib, 3) * EL3(V, k5, k) + A2U(ib, 4) * EL3(V, k6, k);
      double uij2 = EL3(U, ib, k) + cofa3 * xijb + cofa4 * yijb;
      double vij2 = EL3(V, ib, k) + cofa7 * xijb + cofa8 * yijb;
      double uij = 0.5 * (uij1 + uij2), vij = 0.5 * (vij1 + vij2);
      double un_tmp = vij * G->dltxc[i] - uij * G->dltyc[i];
      double viscof1 = G->art[ia] * sqrt(sq(cofa1) + sq(cofa6) + 0.5 * sq(cofa2 + cofa5));
      double viscof2 = G->art[ib] * sqrt(sq(cofa3) + sq(cofa8) + 0.5 * sq(cofa4 + cofa7));
      double viscof = (fact * 0.5 * (viscof1 * G->cc_hvc[ia] + viscof2 * G->cc_hvc[ib]) +
                       fm1 * 0.5 * (G->cc_hvc[ia] + G->cc_hvc[ib])) / o->p.hprnu;
      double txxij = (cofa1 + cofa3) * viscof, tyyij = (cofa6 + cofa8) * viscof;
      double txyij = 0.5 * (cofa2 + cofa4 + cofa5 + cofa7) * viscof;
      double fxx = dij * (txxij * G->dltyc[i] - txyij * G->dltxc[i]);
      double fyy = dij * (txyij * G->dltyc[i] - tyyij * G->dltxc[i]);
      double sgn = fsign(1.0, un_tmp);
      double xadv = dij * un_tmp * ((1.0 - sgn) * uij2 + (1.0 + sgn) * uij1) * 0.5;
      double yadv = dij * un_tmp * ((1.0 - sgn) * vij2 + (1.0 + sgn) * vij1) * 0.5;
      double isbc_tmp = (double)G->isbc[i];
      double tpa = (double)(1 - G->isbc[i]) * G->epor[ia];
      double tpb = (double)(1 - G->isbc[i]) * G->epor[ib];
      EL3(xflux, ia, k) = EL3(xflux, ia, k) + xadv * tpa + (fxx + 3.0 * fxx * isbc_tmp) * G->epor[ia];
      EL3(yflux, ia, k) = EL3(yflux, ia, k) + yadv * tpa + (fyy + 3.0 * fyy * isbc_tmp) * G->epor[ia];
      EL3(xflux, ib, k) = EL3(xflux, ib, k) - xadv * tpb - (fxx + 3.0 * fxx * isbc_tmp) * G->epor[ib];
      EL3(yflux, ib, k) = EL3(yflux, ib, k) - yadv * tpb - (fyy + 3.0 * fyy * isbc_tmp) * G->epor[ib];
    }
  }
  for (int i = 1; i <= o->N; ++i)
    if (G->isbce[i] == 2)
      for (int k = 1; k <= KBM1; ++k) { EL3(xflux, i, k) = 0.0; EL3(yflux, i, k) = 0.0; }
  for (int ii = 0; ii < o->NUMQBC; ++ii) {
    if (G->icellq[ii] > 0) {
      int i1 = G->icellq[ii];
      for (int k = 1; k <= KBM1; ++k) {
        double vlctyq = o->qdis[ii] / G->qarea[ii];
        double temp = o->qdis[ii] * VQDIST(ii, k) * VQDIST(ii, k) * vlctyq / EL3(G->dz1, i1, k);
        EL3(xflux, i1, k) = EL3(xflux, i1, k) - temp * cos(G->angleq[ii]);
        EL3(yflux, i1, k) = EL3(yflux, i1, k) - temp * sin(G->angleq[ii]);
      }
      continue;
    }
    int j = G->inodeq[ii], i1 = NBVE(j, 1), i2 = NBVE(j, NTVE(j));
    for (int k = 1; k <= KBM1; ++k) {
      double vlctyq = o->qdis[ii] / G->qarea[ii];
      double temp = 0.5 * o->qdis[ii] * VQDIST(ii, k) * VQDIST(ii, k) * vlctyq / ND(G->dz, j, k);
      EL3(xflux, i1, k) = EL3(xflux, i1, k) - temp * cos(G->angleq[ii]);
      EL3(xflux, i2, k) = EL3(xflux, i2, k) - temp * cos(G->angleq[ii]);
      EL3(yflux, i1, k) = EL3(yflux, i1, k) - temp * sin(G->angleq[ii]);
      EL3(yflux, i2, k) = EL3(yflux, i2, k) - temp * sin(G->angleq[ii]);
    }
  }
}

/* BAROPG, FVCOM_source/baropg.F:33-166 */
static void baropg(orc* o) {
  const int N = o->N, KB = o->KB, KBM1 = o->KBM1;
  for (int k = 1; k <= KBM1; ++k)
    for (int i = 0; i <= o->M; ++i) ND(o->rho1, i, k) = ND(o->rho1, i, k) - ND(o->rmean1, i, k);
  for (int k = 1; k <= KB; ++k)
    for (int i = 0; i <= N; ++i) EL3(o->rho, i, k) = EL3(o->rho, i, k) - EL3(o->rmean, i, k);
  memset(o->drhox, 0, sizeof(double) * N1 * KB);
  memset(o->drhoy, 0, sizeof(double) * N1 * KB);
  for (int k = 1; k <= KB; ++k) { EL3(o->rmean, 0, k) = 0.0; EL3(o->rho, 0, k) = 0.0; }
  /* per element: RIJK(3,K) prefix sums DRIJK1 and DRIJK2 */
  double rijk[3][ORC_COL], drijk1[3][ORC_COL], drijk2[ORC_COL];
  for (int i = 1; i <= N; ++i) {
    for (int j = 1; j <= 3; ++j) {
      int jj1 = j + 1 - ((j + 1) / 4) * 3, jj2 = j + 2 - ((j + 2) / 4) * 3;
      for (int k = 1; k <= KBM1; ++k)
        rijk[j - 1][k] = 0.5 * (ND(o->rho1, NV(i, jj1), k) + ND(o->rho1, NV(i, jj2), k));
      drijk1[j - 1][1] = rijk[j - 1][1] * (-EL3(G->zz1, i, 1));
      for (int k = 2; k <= KBM1; ++k) {
        drijk1[j - 1][k] = 0.5 * (rijk[j - 1][k - 1] + rijk[j - 1][k]) * (EL3(G->zz1, i, k - 1) - EL3(G->zz1, i, k));
        drijk1[j - 1][k] = drijk1[j - 1][k] + drijk1[j - 1][k - 1];
      }
    }
    drijk2[1] = 0.0;
    for (int k = 2; k <= KBM1; ++k) {
      drijk2[k] = 0.5 * (EL3(G->zz1, i, k - 1) + EL3(G->zz1, i, k)) * (EL3(o->rho, i, k) - EL3(o->rho, i, k - 1));
      drijk2[k] = drijk2[k - 1] + drijk2[k];
    }
    for (int k = 1; k <= KBM1; ++k) {
      for (int j = 1; j <= 3; ++j) {
        int jj1 = j + 1 - ((j + 1) / 4) * 3, jj2 = j + 2 - ((j + 2) / 4) * 3;
        int n1_ = NV(i, jj1), n2_ = NV(i, jj2);
        double dij = 0.5 * (o->dt[n1_] + o->dt[n2_]);
        double drho1 = (G->vy[n1_] - G->vy[n2_]) * drijk1[j - 1][k] * o->dt1[i];
        double drho2 = (G->vy[n1_] - G->vy[n2_]) * dij * drijk2[k];
        EL3(o->drhox, i, k) = EL3(o->drhox, i, k) + drho1 + drho2;
        drho1 = (G->vx[n2_] - G->vx[n1_]) * drijk1[j - 1][k] * o->dt1[i];
        drho2 = (G->vx[n2_] - G->vx[n1_]) * dij * drijk2[k];
        EL3(o->drhoy, i, k) = EL3(o->drhoy, i, k) + drho1 + drho2;
      }
    }
  }
  for (int k = 1; k <= KBM1; ++k)
    for (int i = 0; i <= N; ++i) {
      EL3(o->drhox, i, k) = EL3(o->drhox, i, k) * o->dt1[i] * EL3(G->dz1, i, k) * G->grav_e[i] * o->ramp;
      EL3(o->drhoy, i, k) = EL3(o->drhoy, i, k) * o->dt1[i] * EL3(G->dz1, i, k) * G->grav_e[i] * o->ramp;
    }
  for (int k = 1; k <= KBM1; ++k)
    for (int i = 0; i <= o->M; ++i) ND(o->rho1, i, k) = ND(o->rho1, i, k) + ND(o->rmean1, i, k);
  for (int k = 1; k <= KB; ++k)
    for (int i = 0; i <= N; ++i) EL3(o->rho, i, k) = EL3(o->rho, i, k) + EL3(o->rmean, i, k);
}

/* VERTVL_EDGE, FVCOM_source/vertvl_edge.F:73-560 */
static void vertvl_edge(orc* o) {
  const int M = o->M, KB = o->KB, KBM1 = o->KBM1;
  double* xflux = o->sc_mk[0];
  memset(xflux, 0, sizeof(double) * M1 * KB);
  for (int i = 1; i <= o->NCV; ++i) {
    int i1 = G->ntrg[i], ia = NIEC(i, 1), ib = NIEC(i, 2);
    for (int k = 1; k <= KBM1; ++k) {
      double dij = o->dt1[i1] * EL3(G->dz1, i1, k);
      double uij = EL3(o->u, i1, k), vij = EL3(o->v, i1, k);
      double exflux = dij * (-uij * G->dltye[i] + vij * G->dltxe[i]);
      ND(xflux, ia, k) = ND(xflux, ia, k) - exflux;
      ND(xflux, ib, k) = ND(xflux, ib, k) + exflux;
    }
  }
  for (int i = 1; i <= M; ++i)
    for (int k = 1; k <= KBM1; ++k)
      if (G->isonb[i] == 2) ND(xflux, i, k) = 0.0;
  for (int j = 0; j < o->NUMQBC; ++j) {
    if (G->icellq[j] > 0) {
      int j1 = N_ICELLQ(j, 1), j2 = N_ICELLQ(j, 2);
      for (int k = 1; k <= KBM1; ++k) {
        ND(xflux, j1, k) = ND(xflux, j1, k) - o->qdis[j] * RDISQ(j, 1) * VQDIST(j, k);
        ND(xflux, j2, k) = ND(xflux, j2, k) - o->qdis[j] * RDISQ(j, 2) * VQDIST(j, k);
      }
      continue;
    }
    int jj = G->inodeq[j];
    for (int k = 1; k <= KBM1; ++k) ND(xflux, jj, k) = ND(xflux, jj, k) - o->qdis[j] * VQDIST(j, k);
  }
  memset(o->wts, 0, sizeof(double) * M1 * KB);
  for (int i = 1; i <= M; ++i) {
    double wbottom = 0.0;
    for (int k = 1; k <= KBM1; ++k)
      ND(o->wts, i, k + 1) = ND(o->wts, i, k) + ND(xflux, i, k) / G->art1[i] + ND(G->dz, i, k) * (o->d[i] - o->dt[i]) / o->dti;
    if (fabs(ND(o->wts, i, KB) - wbottom) > 1.0E-8) {
      if (G->isonb[i] != 2) {
        double tmp1 = o->elf[i] * (double)KBM1 - (ND(o->wts, i, KB) - wbottom) * o->dti / ND(G->dz, i, 1);
        tmp1 = tmp1 / (double)KBM1;
        o->dtfa[i] = tmp1 + G->h[i];
        for (int k = 2; k <= KB; ++k)
          ND(o->wts, i, k) = ND(o->wts, i, k) - (double)((float)(k - 1) / (float)KBM1) * (ND(o->wts, i, KB) - wbottom);  /* FLOAT()/FLOAT() is default real */
      }
    }
  }
  for (int i = 1; i <= o->N; ++i)
    for (int k = 1; k <= KB; ++k)
      EL3(o->w, i, k) = ONE_THIRD * (ND(o->wts, NV(i, 1), k) + ND(o->wts, NV(i, 2), k) + ND(o->wts, NV(i, 3), k));
}

/* VISCOF_H, FVCOM_source/viscofh.F:33-127 */
static void viscof_h(orc* o) {
  for (int k = 1; k <= o->KBM1; ++k)
    for (int i = 1; i <= o->M; ++i) {
      double pupx = 0.0, pupy = 0.0, pvpx = 0.0, pvpy = 0.0;
      int nt = NTVE(i), j = 1, i1 = NBVE(i, j);
      double uu = EL3(o->u, i1, k), vv = EL3(o->v, i1, k);
      pupx = pupx + uu * DLTYECEC(i, j); pupy = pupy + uu * DLTXECEC(i, j);
      pvpx = pvpx + vv * DLTYECEC(i, j); pvpy = pvpy + vv * DLTXECEC(i, j);
      if (G->isonb[i] != 0) {
        pupx = pupx + uu * DLTYNEC(i, j); pupy = pupy + uu * DLTXNEC(i, j);
        pvpx = pvpx + vv * DLTYNEC(i, j); pvpy = pvpy + vv * DLTXNEC(i, j);
      }
      for (j = 2; j <= nt - 1; ++j) {
        i1 = NBVE(i, j); uu = EL3(o->u, i1, k); vv = EL3(o->v, i1, k);
        pupx = pupx + uu * DLTYECEC(i, j); pupy = pupy + uu * DLTXECEC(i, j);
        pvpx = pvpx + vv * DLTYECEC(i, j); pvpy = pvpy + vv * DLTXECEC(i, j);
      }
      j = nt; i1 = NBVE(i, j); uu = EL3(o->u, i1, k); vv = EL3(o->v, i1, k);
      pupx = pupx + uu * DLTYECEC(i, j); pupy = pupy + uu * DLTXECEC(i, j);
      pvpx = pvpx + vv * DLTYECEC(i, j); pvpy = pvpy + vv * DLTXECEC(i, j);
      if (G->isonb[i] != 0) {
        pupx = pupx + uu * (-DLTYNEC(i, j)); pupy = pupy + uu * (-DLTXNEC(i, j));
        pvpx = pvpx + vv * (-DLTYNEC(i, j)); pvpy = pvpy + vv * (-DLTXNEC(i, j));
      }
      pupx = pupx / G->art1[i]; pupy = pupy / G->art1[i];
      pvpx = pvpx / G->art1[i]; pvpy = pvpy / G->art1[i];
      double tmp1 = sq(pupx) + sq(pvpy);
      double tmp2 = 0.5 * sq(pupy + pvpx);
      ND(o->viscofh, i, k) = sqrt(tmp1 + tmp2) * G->art1[i];
    }
}

/* ADV_UV_EDGE_GCN, FVCOM_source/adv_uv_edge_gcn.F:31-1198 */
static void adv_uv_edge_gcn(orc* o) {
  const double fact = o->p.horizontal_closure ? 1.0 : 0.0, fm1 = 1.0 - fact;
  const int N = o->N, KB = o->KB, KBM1 = o->KBM1;
  double* xflux = o->sc_nk[0]; double* yflux = o->sc_nk[1];
  double* pstx_tm = o->sc_nk[2]; double* psty_tm = o->sc_nk[3];
  const size_t nk = N1 * KB;
  memset(o->viscofm, 0, sizeof(double) * nk);
  memset(xflux, 0, sizeof(double) * nk); memset(yflux, 0, sizeof(double) * nk);
  memset(pstx_tm, 0, sizeof(double) * nk); memset(psty_tm, 0, sizeof(double) * nk);
  const double* U = o->u; const double* V = o->v;
  for (int i = 1; i <= o->NE; ++i) {
    int ia = IEC(i, 1), ib = IEC(i, 2), j1 = IENODE(i, 1), j2 = IENODE(i, 2);
    double elij = 0.5 * (o->egf[j1] + o->egf[j2]);
    int k1 = NBE(ia, 1), k2 = NBE(ia, 2), k3 = NBE(ia, 3), k4 = NBE(ib, 1), k5 = NBE(ib, 2), k6 = NBE(ib, 3);
    for (int k = 1; k <= KBM1; ++k) {
      double dij = 0.5 * (o->dt[j1] * ND(G->dz, j1, k) + o->dt[j2] * ND(G->dz, j2, k));
      double xija = G->xijc[i] - G->xc[ia], yija = G->yijc[i] - G->yc[ia];
      double xijb = G->xijc[i] - G->xc[ib], yijb = G->yijc[i] - G->yc[ib];
      double cofa1 = A1U(ia, 1) * EL3(U, ia, k) + A1U(ia, 2) * EL3(U, k1, k) + A1U(ia, 3) * EL3(U, k2, k) + A1U(ia, 4) * EL3(U, k3, k);
      double cofa2 = A2U(ia, 1) * EL3(U, ia, k) + A2U(ia, 2) * EL3(U, k1, k) + A2U(ia, 3) * EL3(U, k2, k) + A2U(ia, 4) * EL3(U, k3, k);
      double cofa5 = A1U(ia, 1) * EL3(V, ia, k) + A1U(ia, 2) * EL3(V, k1, k) + A1U(ia, 3) * EL3(V, k2, k) + A1U(ia, 4) * EL3(V, k3, k);
      double cofa6 = A2U(ia, 1) * EL3(V, ia, k) + A2U(ia, 2) * EL3(V, k1, k) + A2U(ia, 3) * EL3(V, k2, k) + A2U(ia, 4) * EL3(V, k3, k);
      double uij1 = EL3(U, ia, k) + cofa1 * xija + cofa2 * yija;
      double vij1 = EL3(V, ia, k) + cofa5 * xija + cofa6 * yija;
      double cofa3 = A1U(ib, 1) * EL3(U, ib, k) + A1U(ib, 2) * EL3(U, k4, k) + A1U(ib, 3) * EL3(U, k5, k) + A1U(ib, 4) * EL3(U, k6, k);
      double cofa4 = A2U(ib, 1) * EL3(U, ib, k) + A2U(ib, 2) * EL3(U, k4, k) + A2U(ib, 3) * EL3(U, k5, k) + A2U(ib, 4) * EL3(U, k6, k);
      double cofa7 = A1U(ib, 1) * EL3(V, ib, k) + A1U(ib, 2) * EL3(V, k4, k) + A1U(ib, 3) * EL3(V, k5, k) + A1U(ib, 4) * EL3(V, k6, k);
      double cofa8 = A2U(ib, 1) * EL3(V, ib, k) + A2U(ib, 2) * EL3(V, k4, k) + A2U(ib, 3) * EL3(V, k5, k) + A2U(ib, 4) * EL3(V, k6, k);
      double uij2 = EL3(U, ib, k) + cofa3 * xijb + cofa4 * yijb;
      double vij2 = EL3(V, ib, k) + cofa7 * xijb + cofa8 * yijb;
      double uij = 0.5 * (uij1 + uij2), vij = 0.5 * (vij1 + vij2);
      double exflux = dij * (-uij * G->dltyc[i] + vij * G->dltxc[i]);
      double viscof1 = G->art[ia] * sqrt(sq(cofa1) + sq(cofa6) + 0.5 * sq(cofa2 + cofa5));
      double viscof2 = G->art[ib] * sqrt(sq(cofa3) + sq(cofa8) + 0.5 * sq(cofa4 + cofa7));
      double viscof = (fact * 0.5 * (viscof1 * G->cc_hvc[ia] + viscof2 * G->cc_hvc[ib]) +
                       fm1 * 0.5 * (G->cc_hvc[ia] + G->cc_hvc[ib])) / o->p.hprnu;
      EL3(o->viscofm, ia, k) = EL3(o->viscofm, ia, k) + viscof;
      EL3(o->viscofm, ib, k) = EL3(o->viscofm, ib, k) + viscof;
      double txxij = (cofa1 + cofa3) * viscof, tyyij = (cofa6 + cofa8) * viscof;
      double txyij = 0.5 * (cofa2 + cofa4 + cofa5 + cofa7) * viscof;
      double fxx = dij * (txxij * G->dltyc[i] - txyij * G->dltxc[i]);
      double fyy = dij * (txyij * G->dltyc[i] - tyyij * G->dltxc[i]);
      double sgn = fsign(1.0, exflux);
      double xadv = exflux * ((1.0 - sgn) * uij2 + (1.0 + sgn) * uij1) * 0.5;
      double yadv = exflux * ((1.0 - sgn) * vij2 + (1.0 + sgn) * vij1) * 0.5;
      double isbc_tmp = (double)G->isbc[i];
      double tpa = (double)(1 - G->isbc[i]) * G->epor[ia];
      double tpb = (double)(1 - G->isbc[i]) * G->epor[ib];
      EL3(xflux, ia, k) = EL3(xflux, ia, k) + xadv * tpa + (fxx + 3.0 * fxx * isbc_tmp) * G->epor[ia];
      EL3(yflux, ia, k) = EL3(yflux, ia, k) + yadv * tpa + (fyy + 3.0 * fyy * isbc_tmp) * G->epor[ia];
      EL3(xflux, ib, k) = EL3(xflux, ib, k) - xadv * tpb - (fxx + 3.0 * fxx * isbc_tmp) * G->epor[ib];
      EL3(yflux, ib, k) = EL3(yflux, ib, k) - yadv * tpb - (fyy + 3.0 * fyy * isbc_tmp) * G->epor[ib];
      EL3(pstx_tm, ia, k) = EL3(pstx_tm, ia, k) - G->grav_e[ia] * o->dt1[ia] * EL3(G->dz1, ia, k) * elij * G->dltyc[i];
      EL3(psty_tm, ia, k) = EL3(psty_tm, ia, k) + G->grav_e[ia] * o->dt1[ia] * EL3(G->dz1, ia, k) * elij * G->dltxc[i];
      EL3(pstx_tm, ib, k) = EL3(pstx_tm, ib, k) + G->grav_e[ib] * o->dt1[ib] * EL3(G->dz1, ib, k) * elij * G->dltyc[i];
      EL3(psty_tm, ib, k) = EL3(psty_tm, ib, k) - G->grav_e[ib] * o->dt1[ib] * EL3(G->dz1, ib, k) * elij * G->dltxc[i];
    }
  }
  for (int i = 1; i <= N; ++i)
    for (int k = 1; k <= KBM1; ++k) {
      EL3(xflux, i, k) = EL3(xflux, i, k) + EL3(pstx_tm, i, k);
      EL3(yflux, i, k) = EL3(yflux, i, k) + EL3(psty_tm, i, k);
    }
  const double* W = o->w; const double* DZ1 = G->dz1;
  for (int i = 1; i <= N; ++i)
    for (int k = 1; k <= KBM1; ++k) {
      double xfluxv, yfluxv;
      if (k == 1) {
        xfluxv = -EL3(W, i, k + 1) * (EL3(U, i, k) * EL3(DZ1, i, k + 1) + EL3(U, i, k + 1) * EL3(DZ1, i, k)) / (EL3(DZ1, i, k) + EL3(DZ1, i, k + 1));
        yfluxv = -EL3(W, i, k + 1) * (EL3(V, i, k) * EL3(DZ1, i, k + 1) + EL3(V, i, k + 1) * EL3(DZ1, i, k)) / (EL3(DZ1, i, k) + EL3(DZ1, i, k + 1));
      } else if (k == KBM1) {
        xfluxv = EL3(W, i, k) * (EL3(U, i, k) * EL3(DZ1, i, k - 1) + EL3(U, i, k - 1) * EL3(DZ1, i, k)) / (EL3(DZ1, i, k) + EL3(DZ1, i, k - 1));
        yfluxv = EL3(W, i, k) * (EL3(V, i, k) * EL3(DZ1, i, k - 1) + EL3(V, i, k - 1) * EL3(DZ1, i, k)) / (EL3(DZ1, i, k) + EL3(DZ1, i, k - 1));
      } else {
        xfluxv = EL3(W, i, k) * (EL3(U, i, k) * EL3(DZ1, i, k - 1) + EL3(U, i, k - 1) * EL3(DZ1, i, k)) / (EL3(DZ1, i, k) + EL3(DZ1, i, k - 1)) -
                 EL3(W, i, k + 1) * (EL3(U, i, k) * EL3(DZ1, i, k + 1) + EL3(U, i, k + 1) * EL3(DZ1, i, k)) / (EL3(DZ1, i, k) + EL3(DZ1, i, k + 1));
        yfluxv = EL3(W, i, k) * (EL3(V, i, k) * EL3(DZ1, i, k - 1) + EL3(V, i, k - 1) * EL3(DZ1, i, k)) / (EL3(DZ1, i, k) + EL3(DZ1, i, k - 1)) -
                 EL3(W, i, k + 1) * (EL3(V, i, k) * EL3(DZ1, i, k + 1) + EL3(V, i, k + 1) * EL3(DZ1, i, k)) / (EL3(DZ1, i, k) + EL3(DZ1, i, k + 1));
      }
      EL3(xflux, i, k) = EL3(xflux, i, k) + xfluxv * G->art[i] + EL3(o->drhox, i, k) - G->cor[i] * EL3(V, i, k) * o->dt1[i] * EL3(DZ1, i, k) * G->art[i];
      EL3(yflux, i, k) = EL3(yflux, i, k) + yfluxv * G->art[i] + EL3(o->drhoy, i, k) + G->cor[i] * EL3(U, i, k) * o->dt1[i] * EL3(DZ1, i, k) * G->art[i];
    }
  for (int i = 1; i <= N; ++i)
    if (G->isbce[i] == 2)
      for (int k = 1; k <= KBM1; ++k) { EL3(xflux, i, k) = 0.0; EL3(yflux, i, k) = 0.0; }
  for (int ii = 0; ii < o->NUMQBC; ++ii) {
    if (G->icellq[ii] > 0) {
      int i1 = G->icellq[ii];
      for (int k = 1; k <= KBM1; ++k) {
        double vlctyq = o->qdis[ii] / G->qarea[ii];
        double temp = o->qdis[ii] * VQDIST(ii, k) * VQDIST(ii, k) * vlctyq / EL3(G->dz1, i1, k);
        EL3(xflux, i1, k) = EL3(xflux, i1, k) - temp * cos(G->angleq[ii]);
        EL3(yflux, i1, k) = EL3(yflux, i1, k) - temp * sin(G->angleq[ii]);
      }
      continue;
    }
    int j = G->inodeq[ii], i1 = NBVE(j, 1), i2 = NBVE(j, NTVE(j));
    for (int k = 1; k <= KBM1; ++k) {
      double vlctyq = o->qdis[ii] / G->qarea[ii];
      double temp = 0.5 * o->qdis[ii] * VQDIST(ii, k) * VQDIST(ii, k) * vlctyq / ND(G->dz, j, k);
      EL3(xflux, i1, k) = EL3(xflux, i1, k) - temp * cos(G->angleq[ii]);
      EL3(xflux, i2, k) = EL3(xflux, i2, k) - temp * cos(G->angleq[ii]);
      EL3(yflux, i1, k) = EL3(yflux, i1, k) - temp * sin(G->angleq[ii]);
      EL3(yflux, i2, k) = EL3(yflux, i2, k) - temp * sin(G->angleq[ii]);
    }
  }
  for (int i = 1; i <= N; ++i)
    for (int k = 1; k <= KBM1; ++k) {
      EL3(o->uf, i, k) = EL3(U, i, k) * o->dt1[i] / o->d1[i] - o->dti * EL3(xflux, i, k) / G->art[i] / (o->d1[i] * EL3(DZ1, i, k));
      EL3(o->vf, i, k) = EL3(V, i, k) * o->dt1[i] / o->d1[i] - o->dti * EL3(yflux, i, k) / G->art[i] / (o->d1[i] * EL3(DZ1, i, k));
    }
  for (int k = 1; k <= KB; ++k)
    for (int i = 0; i <= N; ++i) EL3(o->viscofm, i, k) = EL3(o->viscofm, i, k) / G->art[i];
}

/* VDIF_UV, FVCOM_source/vdif_uv.F:34-189 (sponge, bottom stress WUBOT) */
static void vdif_uv(orc* o) {
  const int N = o->N, KB = o->KB, KBM1 = o->KBM1, KBM2 = o->KBM2;
  const double dti = o->dti, umol = o->p.umol;
  double a[ORC_COL], c[ORC_COL], vhu[ORC_COL], vhv[ORC_COL], vhpu[ORC_COL], vhpv[ORC_COL];
  for (int i = 1; i <= N; ++i) {
    if (!(o->d1[i] > 0.0)) continue;
    double d1 = o->d1[i];
    for (int k = 1; k <= KB; ++k) c[k] = EL3(o->km1, i, k);
    for (int k = 2; k <= KBM1; ++k) {
      a[k - 1] = -dti * (c[k] + umol) / (EL3(G->dz1, i, k - 1) * EL3(G->dzz1, i, k - 1) * d1 * d1);
      c[k] = -dti * (c[k] + umol) / (EL3(G->dz1, i, k) * EL3(G->dzz1, i, k - 1) * d1 * d1);
    }
    int it1 = G->isonb[NV(i, 1)], it2 = G->isonb[NV(i, 2)], it3 = G->isonb[NV(i, 3)];
    double tmp = (it1 == 2 || it2 == 2 || it3 == 2) ? 0.0 : 1.0;
    vhu[1] = a[1] / (a[1] - 1.);
    vhv[1] = a[1] / (a[1] - 1.);
    vhpu[1] = (-dti * o->wusurf[i] * tmp / (-EL3(G->dz1, i, 1) * d1) - EL3(o->uf, i, 1)) / (a[1] - 1.);
    vhpv[1] = (-dti * o->wvsurf[i] * tmp / (-EL3(G->dz1, i, 1) * d1) - EL3(o->vf, i, 1)) / (a[1] - 1.);
    for (int k = 2; k <= KBM2; ++k) {
      vhpu[k] = 1.0 / (a[k] + c[k] * (1. - vhu[k - 1]) - 1.);
      vhpv[k] = 1.0 / (a[k] + c[k] * (1. - vhv[k - 1]) - 1.);
      vhu[k] = a[k] * vhpu[k];
      vhv[k] = a[k] * vhpv[k];
      vhpu[k] = (c[k] * vhpu[k - 1] - EL3(o->uf, i, k)) * vhpu[k];
      vhpv[k] = (c[k] * vhpv[k - 1] - EL3(o->vf, i, k)) * vhpv[k];
    }
    double ub = EL3(o->u, i, KBM1), vb = EL3(o->v, i, KBM1);
    o->tps[i] = o->cbc[i] * sqrt(ub * ub + vb * vb);
    EL3(o->uf, i, KBM1) = (c[KBM1] * vhpu[KBM2] - EL3(o->uf, i, KBM1)) /
                          (o->tps[i] * dti / (-EL3(G->dz1, i, KBM1) * d1) - 1. - (vhu[KBM2] - 1.) * c[KBM1]);
    EL3(o->vf, i, KBM1) = (c[KBM1] * vhpv[KBM2] - EL3(o->vf, i, KBM1)) /
                          (o->tps[i] * dti / (-EL3(G->dz1, i, KBM1) * d1) - 1. - (vhv[KBM2] - 1.) * c[KBM1]);
    for (int k = 2; k <= KBM1; ++k) {
      int ki = KB - k;
      EL3(o->uf, i, ki) = (vhu[ki] * EL3(o->uf, i, ki + 1) + vhpu[ki]);
      EL3(o->vf, i, ki) = (vhv[ki] * EL3(o->vf, i, ki + 1) + vhpv[ki]);
    }
  }
  for (int k = 1; k <= KBM1; ++k)
    for (int i = 1; i <= N; ++i) {
      double uu = EL3(o->uf, i, k), vv = EL3(o->vf, i, k);
      EL3(o->uf, i, k) = uu / (1.0 + G->cc_sponge[i] * uu * uu);
      EL3(o->vf, i, k) = vv / (1.0 + G->cc_sponge[i] * vv * vv);
    }
  for (int i = 1; i <= N; ++i)
    if (o->d1[i] > 0.0) {
      o->wubot[i] = -o->tps[i] * EL3(o->u, i, KBM1);
      o->wvbot[i] = -o->tps[i] * EL3(o->v, i, KBM1);
    }
}

/* WREAL, FVCOM_source/wreal.F:32-140 */
static void wreal(orc* o) {
  for (int i = 1; i <= o->N; ++i) {
    int j1 = NV(i, 1), j2 = NV(i, 2), j3 = NV(i, 3);
    double dddx = AWX(i, 1) * o->d[j1] + AWX(i, 2) * o->d[j2] + AWX(i, 3) * o->d[j3];
    double dddy = AWY(i, 1) * o->d[j1] + AWY(i, 2) * o->d[j2] + AWY(i, 3) * o->d[j3];
    double dedx = AWX(i, 1) * o->elf[j1] + AWX(i, 2) * o->elf[j2] + AWX(i, 3) * o->elf[j3];
    double dedy = AWY(i, 1) * o->elf[j1] + AWY(i, 2) * o->elf[j2] + AWY(i, 3) * o->elf[j3];
    double etf1aa = ONE_THIRD * (o->el[j1] + o->el[j2] + o->el[j3]);
    for (int k = 1; k <= o->KBM1; ++k) {
      double ww1 = 0.5 * (EL3(o->w, i, k) + EL3(o->w, i, k + 1)) + EL3(o->u, i, k) * (EL3(G->zz1, i, k) * dddx + dedx) +
                   EL3(o->v, i, k) * (EL3(G->zz1, i, k) * dddy + dedy);
      double ww2 = (EL3(G->zz1, i, k) + 1.) * (etf1aa - o->et1[i]) / o->dti;
      EL3(o->ww, i, k) = ww1 + ww2;
    }
  }
}

/* Ring gradient of a node field at one level, shared by ADV_T/ADV_S
 * (adv_t.F:219-376) and ADV_Q (adv_q.F:185-230). */
static void ring_grad(orc* o, int i, const double* f, const double* fm, double fmc, int k,
                      double* px, double* py, double* pxd, double* pyd) {
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  for (int j = 1; j <= NTSN(i) - 1; ++j) {
    int i1 = NBSN(i, j), i2 = NBSN(i, j + 1);
    double ffd = fm ? 0.5 * (ND(f, i1, k) + ND(f, i2, k) - ND(fm, i1, k) - ND(fm, i2, k))
                    : 0.5 * (ND(f, i1, k) + ND(f, i2, k) - fmc - fmc);
    double ff1 = 0.5 * (ND(f, i1, k) + ND(f, i2, k));
    a = a + ff1 * DLTYTRIE(i, j);
    b = b + ff1 * DLTXTRIE(i, j);
    c = c + ffd * DLTYTRIE(i, j);
    d = d + ffd * DLTXTRIE(i, j);
  }
  *px = a / G->art2[i]; *py = b / G->art2[i]; *pxd = c / G->art2[i]; *pyd = d / G->art2[i];
}

static void ring_minmax(orc* o, int i, const double* f, int k, double* mn, double* mx) {
  double lo = ND(f, NBSN(i, 1), k), hi = lo;
  for (int j = 2; j <= NTSN(i) - 1; ++j) {
    double x = ND(f, NBSN(i, j), k);
    if (x < lo) lo = x;
    if (x > hi) hi = x;
  }
  double c = ND(f, i, k);
  *mn = fmin(lo, c); *mx = fmax(hi, c);
}

/* ADV_Q(Q,QF), FVCOM_source/adv_q.F:35-418 */
static void adv_q(orc* o, const double* q, double* qf) {
  const double fact = o->p.horizontal_closure ? 1.0 : 0.0, fm1 = 1.0 - fact;
  const int M = o->M, KB = o->KB, KBM1 = o->KBM1;
  const double qmean1 = (double)1.E-8f;
  double* xflux = o->sc_mk[0];
  double* uq = o->sc_nk[0]; double* vq = o->sc_nk[1];
  memset(qf, 0, sizeof(double) * M1 * KB);
  memset(xflux, 0, sizeof(double) * M1 * KB);
  memset(uq, 0, sizeof(double) * N1 * KB); memset(vq, 0, sizeof(double) * N1 * KB);
  for (int k = 2; k <= KBM1; ++k)
    for (int i = 1; i <= o->N; ++i) {
      EL3(uq, i, k) = (EL3(o->u, i, k) * EL3(G->dz1, i, k - 1) + EL3(o->u, i, k - 1) * EL3(G->dz1, i, k)) / (EL3(G->dz1, i, k) + EL3(G->dz1, i, k - 1));
      EL3(vq, i, k) = (EL3(o->v, i, k) * EL3(G->dz1, i, k - 1) + EL3(o->v, i, k - 1) * EL3(G->dz1, i, k)) / (EL3(G->dz1, i, k) + EL3(G->dz1, i, k - 1));
    }
  double* pqpx = o->sc_m1[0]; double* pqpy = o->sc_m1[1]; double* pqpxd = o->sc_m1[2]; double* pqpyd = o->sc_m1[3];
  double* viscoff = o->sc_m1[4];
  for (int k = 2; k <= KBM1; ++k) {
    for (int i = 1; i <= M; ++i) ring_grad(o, i, q, NULL, qmean1, k, &pqpx[i], &pqpy[i], &pqpxd[i], &pqpyd[i]);
    for (int i = 1; i <= M; ++i)
      viscoff[i] = (ND(o->viscofh, i, k) * ND(G->dz, i, k - 1) + ND(o->viscofh, i, k - 1) * ND(G->dz, i, k)) / (ND(G->dz, i, k) + ND(G->dz, i, k - 1));
    for (int i = 1; i <= o->NCV; ++i) {
      int ia = NIEC(i, 1), ib = NIEC(i, 2), i1 = G->ntrg[i];
      double dtij = o->dt1[i1] * EL3(G->dzz1, i1, k - 1);
      double uvn = EL3(vq, i1, k) * G->dltxe[i] - EL3(uq, i1, k) * G->dltye[i];
      double fij1 = ND(q, ia, k) + DLTXNCVE(i, 1) * pqpx[ia] + DLTYNCVE(i, 1) * pqpy[ia];
      double fij2 = ND(q, ib, k) + DLTXNCVE(i, 2) * pqpx[ib] + DLTYNCVE(i, 2) * pqpy[ib];
      double q1min, q1max, q2min, q2max;
      ring_minmax(o, ia, q, k, &q1min, &q1max);
      ring_minmax(o, ib, q, k, &q2min, &q2max);
      if (fij1 < q1min) fij1 = q1min;
      if (fij1 > q1max) fij1 = q1max;
      if (fij2 < q2min) fij2 = q2min;
      if (fij2 > q2max) fij2 = q2max;
      double un = uvn;
      double viscof = (fact * 0.5 * (viscoff[ia] * G->nn_hvc[ia] + viscoff[ib] * G->nn_hvc[ib]) + fm1 * 0.5 * (G->nn_hvc[ia] + G->nn_hvc[ib])) / o->p.hprnu;
      double txx = 0.5 * (pqpxd[ia] + pqpxd[ib]) * viscof;
      double tyy = 0.5 * (pqpyd[ia] + pqpyd[ib]) * viscof;
      double fxx = -dtij * txx * G->dltye[i];
      double fyy = dtij * tyy * G->dltxe[i];
      double sgn = fsign(1.0, un);
      double exflux = -un * dtij * ((1.0 + sgn) * fij2 + (1.0 - sgn) * fij1) * 0.5 + fxx + fyy;
      ND(xflux, ia, k) = ND(xflux, ia, k) + exflux;
      ND(xflux, ib, k) = ND(xflux, ib, k) - exflux;
    }
  }
  for (int k = 2; k <= KBM1; ++k)
    for (int i = 1; i <= M; ++i) {
      double temp = ND(o->wts, i, k - 1) * ND(q, i, k - 1) - ND(o->wts, i, k + 1) * ND(q, i, k + 1);
      ND(xflux, i, k) = ND(xflux, i, k) + temp * G->art1[i] * ND(G->dzz, i, k - 1) / (ND(G->dz, i, k - 1) + ND(G->dz, i, k));
    }
  for (int i = 1; i <= M; ++i)
    for (int k = 2; k <= KBM1; ++k)
      ND(qf, i, k) = (ND(q, i, k) - ND(xflux, i, k) / G->art1[i] * (o->dti / (o->dt[i] * ND(G->dzz, i, k - 1)))) * (o->dt[i] / o->d[i]);
}

static int fct_skip_node(orc* o, int i) {
  for (int j = 0; j < o->IOBCN; ++j) if (i == G->i_obc_n[j]) return 1;
  for (int j = 0; j < o->NUMQBC; ++j) {
    if (G->icellq[j] > 0) { if (i == N_ICELLQ(j, 1) || i == N_ICELLQ(j, 2)) return 1; }
    else if (i == G->inodeq[j]) return 1;
  }
  return 0;
}

/* FCT_Q2 / FCT_Q2L, FVCOM_source/fct_q2.F:31-104, fct_q2l.F */
static void fct_q(orc* o, const double* q, double* qf) {
  const int KBM1 = o->KBM1;
  const double* DZZ = G->dzz;
  for (int i = 1; i <= o->M; ++i) {
    if (fct_skip_node(o, i)) continue;
    for (int k = 2; k <= KBM1; ++k) {
      double qmax = ND(q, NBSN(i, 1), k), qmin = qmax;
      for (int j = 2; j <= NTSN(i); ++j) { double x = ND(q, NBSN(i, j), k); qmax = fmax(qmax, x); qmin = fmin(qmin, x); }
      if (k == 2) {
        double a = (ND(q, i, k) * ND(DZZ, i, k) + ND(q, i, k + 1) * ND(DZZ, i, k - 1)) / (ND(DZZ, i, k) + ND(DZZ, i, k - 1));
        qmax = fmax(qmax, a); qmin = fmin(qmin, a);
      } else if (k == KBM1) {
        double a = (ND(q, i, k) * ND(DZZ, i, k - 2) + ND(q, i, k - 1) * ND(DZZ, i, k - 1)) / (ND(DZZ, i, k - 1) + ND(DZZ, i, k - 2));
        qmax = fmax(qmax, a); qmin = fmin(qmin, a);
      } else {
        double a = (ND(q, i, k) * ND(DZZ, i, k - 2) + ND(q, i, k - 1) * ND(DZZ, i, k - 1)) / (ND(DZZ, i, k - 1) + ND(DZZ, i, k - 2));
        double b = (ND(q, i, k) * ND(DZZ, i, k) + ND(q, i, k + 1) * ND(DZZ, i, k - 1)) / (ND(DZZ, i, k) + ND(DZZ, i, k - 1));
        qmax = fmax(qmax, fmax(a, b)); qmin = fmin(qmin, fmin(a, b));
      }
      if (qmin - ND(qf, i, k) > 0.0) ND(qf, i, k) = qmin;
      if (ND(qf, i, k) - qmax > 0.0) ND(qf, i, k) = qmax;
    }
  }
}

/* VDIF_Q (MY2.5), FVCOM_source/vdif_q.F:34-551 */
static void vdif_q(orc* o) {
  const int M = o->M, KB = o->KB, KBM1 = o->KBM1;
  const double A1 = 0.92, B1 = 16.6, A2 = 0.74, B2 = 10.1, C1 = 0.08, E1 = 1.80, E2 = 1.33, SEF = 1.00;
  const double KAPPA = 0.40, SMALL = 1.E-8, GH_MAX = .0233, GHC = -6.0, CBCNST = 100.0, SURFL = 2.E+5;
  const double GH_MIN = -.281;
  const double dti = o->dti, umol = o->p.umol;
  const double const1 = pow(16.6, .6666667) * SEF;
  const double cvhp1 = pow(15.8 * CBCNST, 0.6666667);
  const double p9 = (double)0.9f, p1 = (double)0.1f;
  double* Q2 = o->q2; double* Q2L = o->q2l; double* Q2F = o->q2f; double* Q2LF = o->q2lf;
  double* L = o->l; double* KM = o->km; double* KH = o->kh; double* KQ = o->kq;
  const double* DZ = G->dz; const double* DZZ = G->dzz; const double* Z = G->z;
  double gh[ORC_COL], sm[ORC_COL], sh[ORC_COL], prod[ORC_COL], dtef[ORC_COL], kn[ORC_COL], boygr[ORC_COL], a[ORC_COL], c[ORC_COL], vhp[ORC_COL], vh[ORC_COL], stf[ORC_COL];
  for (int k = 2; k <= KBM1; ++k)
    for (int i = 1; i <= M; ++i)
      if (ND(Q2F, i, k) <= SMALL || ND(Q2LF, i, k) <= SMALL) { ND(Q2F, i, k) = SMALL; ND(Q2LF, i, k) = SMALL; }
  for (size_t q = 0; q < M1 * KB; ++q) { Q2[q] = fabs(Q2[q] + SMALL); Q2L[q] = fabs(Q2L[q]); }
  for (int i = 1; i <= M; ++i) {
    const double d = o->d[i];
    for (int k = 0; k <= KB; ++k) { gh[k] = sm[k] = sh[k] = prod[k] = dtef[k] = kn[k] = boygr[k] = a[k] = c[k] = vhp[k] = vh[k] = 0.0; stf[k] = 1.0; }
    for (int k = 2; k <= KBM1; ++k) {
      a[k] = -0.5 * dti * (ND(KQ, i, k + 1) + ND(KQ, i, k) + 2.0 * umol) / (ND(DZZ, i, k - 1) * ND(DZ, i, k) * d * d);
      c[k] = -0.5 * dti * (ND(KQ, i, k - 1) + ND(KQ, i, k) + 2.0 * umol) / (ND(DZZ, i, k - 1) * ND(DZ, i, k - 1) * d * d);
    }
    double wus = 0.0, wvs = 0.0;
    for (int j = 1; j <= NTVE(i); ++j) { wus = wus + o->wusurf[NBVE(i, j)]; wvs = wvs + o->wvsurf[NBVE(i, j)]; }
    wus = wus / (double)NTVE(i); wvs = wvs / (double)NTVE(i);
    double utau2 = sqrt(sq(wus) + sq(wvs));
    vhp[1] = cvhp1 * utau2;
    vh[1] = 0.0;
    double l0 = SURFL * utau2 / G->grav_n[i];
    double wub = 0.0, wvb = 0.0;
    for (int j = 1; j <= NTVE(i); ++j) { wub = wub + o->wubot[NBVE(i, j)]; wvb = wvb + o->wvbot[NBVE(i, j)]; }
    wub = wub / (double)NTVE(i); wvb = wvb / (double)NTVE(i);
    ND(Q2F, i, KB) = sqrt(sq(wub) + sq(wvb)) * const1;
    for (int k = 2; k <= KBM1; ++k) {
      boygr[k] = G->grav_n[i] * (ND(o->rho1, i, k - 1) - ND(o->rho1, i, k)) / (ND(DZZ, i, k - 1) * d);
      prod[k] = ND(KM, i, k) * 0.0 * fv_pow15(.5 * (-boygr[k] + fabs(boygr[k])));
    }
    for (int k = 2; k <= KBM1; ++k) {
      if (d > 0.0) {
        double uk = 0.0, vk = 0.0, ukm = 0.0, vkm = 0.0;
        for (int j = 1; j <= NTVE(i); ++j) {
          int e = NBVE(i, j);
          uk = uk + EL3(o->u, e, k); vk = vk + EL3(o->v, e, k);
          ukm = ukm + EL3(o->u, e, k - 1); vkm = vkm + EL3(o->v, e, k - 1);
        }
        double nt = (double)NTVE(i);
        uk = uk / nt; vk = vk / nt; ukm = ukm / nt; vkm = vkm / nt;
        prod[k] = prod[k] + ND(KM, i, k) * SEF * (sq(uk - ukm) + sq(vk - vkm)) / sq(ND(DZZ, i, k - 1) * d);
        prod[k] = prod[k] + ND(KH, i, k) * boygr[k];
      }
    }
    for (int k = 2; k <= KBM1; ++k) ND(L, i, k) = ND(Q2L, i, k) / ND(Q2, i, k);
    for (int k = 2; k <= KBM1; ++k) {
      if (ND(Z, i, k) > -0.5) ND(L, i, k) = fmax(ND(L, i, k), KAPPA * l0);
      if (boygr[k] < 0.0) {
        double lmax = sqrt(fabs(GH_MIN) * ND(Q2, i, k) / (-boygr[k] + SMALL));
        ND(L, i, k) = fmin(ND(L, i, k), lmax);
        ND(Q2L, i, k) = ND(Q2, i, k) * ND(L, i, k);
      }
    }
    for (int k = 1; k <= KBM1; ++k) gh[k] = (sq(ND(L, i, k)) / ND(Q2, i, k)) * boygr[k];
    for (int k = 1; k <= KB; ++k) gh[k] = fmin(gh[k], GH_MAX);
    ND(L, i, 1) = KAPPA * l0;
    ND(L, i, KB) = 0.0;
    gh[1] = 0.0; gh[KB] = 0.0;
    for (int k = 1; k <= KB; ++k) {
      stf[k] = 1.0;
      if (gh[k] < 0.0) stf[k] = 1.0 - p9 * fv_pow15(gh[k] / GHC);
      if (gh[k] < GHC) stf[k] = p1;
    }
    for (int k = 1; k <= KB; ++k) dtef[k] = sqrt(ND(Q2, i, k)) / (B1 * ND(L, i, k) + SMALL) * stf[k];
    for (int k = 2; k <= KBM1; ++k) {
      vhp[k] = 1. / (a[k] + c[k] * (1. - vh[k - 1]) - (2. * dti * dtef[k] + 1.));
      vh[k] = a[k] * vhp[k];
      vhp[k] = (-2. * dti * prod[k] + c[k] * vhp[k - 1] - ND(Q2F, i, k)) * vhp[k];
    }
    for (int k = 1; k <= KBM1; ++k) { int ki = KB - k; ND(Q2F, i, ki) = vh[ki] * ND(Q2F, i, ki + 1) + vhp[ki]; }
    vh[1] = 0.0; vhp[1] = 0.0;
    for (int k = 2; k <= KBM1; ++k) {
      if (d > 0.0) {
        dtef[k] = dtef[k] * (1. + E2 * sq((1. / fabs(ND(Z, i, k) - ND(Z, i, 1)) + 1. / fabs(ND(Z, i, k) - ND(Z, i, KB))) * ND(L, i, k) / (d * KAPPA)));
        vhp[k] = 1. / (a[k] + c[k] * (1. - vh[k - 1]) - (dti * dtef[k] + 1.));
        vh[k] = a[k] * vhp[k];
        vhp[k] = (dti * (-prod[k] * ND(L, i, k) * E1) + c[k] * vhp[k - 1] - ND(Q2LF, i, k)) * vhp[k];
      }
    }
    for (int k = 1; k <= KBM1; ++k) { int ki = KB - k; ND(Q2LF, i, ki) = vh[ki] * ND(Q2LF, i, ki + 1) + vhp[ki]; }
    for (int k = 2; k <= KBM1; ++k)
      if (ND(Q2F, i, k) <= SMALL || ND(Q2LF, i, k) <= SMALL) { ND(Q2F, i, k) = SMALL; ND(Q2LF, i, k) = SMALL; }
    const double coef4 = 18.0 * A1 * A1 + 9.0 * A1 * A2, coef5 = 9.0 * A1 * A2;
    for (int k = 1; k <= KB; ++k) {
      double coef1 = A2 * (1.0 - 6.0 * A1 / B1 * stf[k]);
      double coef2 = 3.0 * A2 * B2 / stf[k] + 18.0 * A1 * A2;
      double coef3 = A1 * (1.0 - 3.0 * C1 - 6.0 * A1 / B1 * stf[k]);
      sh[k] = coef1 / (1.0 - coef2 * gh[k]);
      sm[k] = (coef3 + sh[k] * coef4 * gh[k]) / (1. - coef5 * gh[k]);
    }
    for (int k = 1; k <= KB; ++k) {
      kn[k] = ND(L, i, k) * sqrt(ND(Q2, i, k));
      ND(KQ, i, k) = 0.5 * (kn[k] * KAPPA * sm[k] + ND(KQ, i, k));
      ND(KM, i, k) = 0.5 * (kn[k] * sm[k] + ND(KM, i, k));
      ND(KH, i, k) = 0.5 * (kn[k] * sh[k] * o->p.vprnu + ND(KH, i, k));
    }
  }
  /* node 0: SM/SH stay zero (never computed there), L(0,1)=KAPPA*L0(0)=0 */
  ND(L, 0, 1) = 0.0;
  ND(L, 0, KB) = 0.0;
  for (int k = 1; k <= KB; ++k) {
    double kn0 = ND(L, 0, k) * sqrt(ND(Q2, 0, k));
    ND(KQ, 0, k) = 0.5 * (kn0 * KAPPA * 0.0 + ND(KQ, 0, k));
    ND(KM, 0, k) = 0.5 * (kn0 * 0.0 + ND(KM, 0, k));
    ND(KH, 0, k) = 0.5 * (kn0 * 0.0 * o->p.vprnu + ND(KH, 0, k));
  }
}

/* ADV_T / ADV_S (the tracer CV scheme), FVCOM_source/adv_t.F:34-1077,
 * adv_s.F:34-1029; which = 1 temperature, 2 salinity */
static void adv_ts(orc* o, int which) {
  const double fact = o->p.horizontal_closure ? 1.0 : 0.0, fm1 = 1.0 - fact;
  const int M = o->M, KB = o->KB, KBM1 = o->KBM1;
  const double* f1 = which == 1 ? o->t1 : o->s1;
  const double* fmean = which == 1 ? o->tmean1 : o->smean1;
  const double* fdis = which == 1 ? G->tdis : G->sdis;
  double* ff = which == 1 ? o->tf1 : o->sf1;
  double* xflux = o->sc_mk[0]; double* xflux_adv = o->sc_mk[1];
  memset(xflux, 0, sizeof(double) * M1 * KB);
  memset(xflux_adv, 0, sizeof(double) * M1 * KB);
  double* ptpx = o->sc_m1[0]; double* ptpy = o->sc_m1[1]; double* ptpxd = o->sc_m1[2]; double* ptpyd = o->sc_m1[3];
  for (int k = 1; k <= KBM1; ++k) {
    for (int i = 1; i <= M; ++i) ring_grad(o, i, f1, fmean, 0.0, k, &ptpx[i], &ptpy[i], &ptpxd[i], &ptpyd[i]);
    if (k == KBM1)
      for (int i = 1; i <= M; ++i) { o->pfpxb[i] = ptpx[i]; o->pfpyb[i] = ptpy[i]; }
    if (k == KBM1)
      for (int i = 1; i <= M; ++i) o->ah_bottom[i] = (fact * ND(o->viscofh, i, k) + fm1) * G->nn_hvc[i];
    for (int i = 1; i <= o->NCV; ++i) {
      int ia = NIEC(i, 1), ib = NIEC(i, 2), i1 = G->ntrg[i];
      double dtij = o->dt1[i1] * EL3(G->dz1, i1, k);
      double uvn = EL3(o->v, i1, k) * G->dltxe[i] - EL3(o->u, i1, k) * G->dltye[i];
      double fij1 = ND(f1, ia, k) + DLTXNCVE(i, 1) * ptpx[ia] + DLTYNCVE(i, 1) * ptpy[ia];
      double fij2 = ND(f1, ib, k) + DLTXNCVE(i, 2) * ptpx[ib] + DLTYNCVE(i, 2) * ptpy[ib];
      double t1min, t1max, t2min, t2max;
      ring_minmax(o, ia, f1, k, &t1min, &t1max);
      ring_minmax(o, ib, f1, k, &t2min, &t2max);
      if (fij1 < t1min) fij1 = t1min;
      if (fij1 > t1max) fij1 = t1max;
      if (fij2 < t2min) fij2 = t2min;
      if (fij2 > t2max) fij2 = t2max;
      double un = uvn;
      double viscof = fact * 0.5 * (ND(o->viscofh, ia, k) * G->nn_hvc[ia] + ND(o->viscofh, ib, k) * G->nn_hvc[ib]) +
                      fm1 * 0.5 * (G->nn_hvc[ia] + G->nn_hvc[ib]);
      double txx = 0.5 * (ptpxd[ia] + ptpxd[ib]) * viscof;
      double tyy = 0.5 * (ptpyd[ia] + ptpyd[ib]) * viscof;
      double fxx = -dtij * txx * G->dltye[i];
      double fyy = dtij * tyy * G->dltxe[i];
      double sgn = fsign(1.0, un);
      double exflux = -un * dtij * ((1.0 + sgn) * fij2 + (1.0 - sgn) * fij1) * 0.5 + fxx + fyy;
      ND(xflux, ia, k) = ND(xflux, ia, k) + exflux;
      ND(xflux, ib, k) = ND(xflux, ib, k) - exflux;
      ND(xflux_adv, ia, k) = ND(xflux_adv, ia, k) + (exflux - fxx - fyy);
      ND(xflux_adv, ib, k) = ND(xflux_adv, ib, k) - (exflux - fxx - fyy);
    }
  }
  for (int k = 1; k <= KBM1; ++k)
    for (int i = 0; i < o->IOBCN; ++i) o->xflux_obc[(size_t)(k - 1) * o->IOBCN + i] = ND(xflux_adv, G->i_obc_n[i], k);
  const double* DZ = G->dz;
  for (int i = 1; i <= M; ++i)
    for (int k = 1; k <= KBM1; ++k) {
      double temp;
      if (k == 1)
        temp = -ND(o->wts, i, k + 1) * (ND(f1, i, k) * ND(DZ, i, k + 1) + ND(f1, i, k + 1) * ND(DZ, i, k)) / (ND(DZ, i, k) + ND(DZ, i, k + 1));
      else if (k == KBM1)
        temp = ND(o->wts, i, k) * (ND(f1, i, k) * ND(DZ, i, k - 1) + ND(f1, i, k - 1) * ND(DZ, i, k)) / (ND(DZ, i, k) + ND(DZ, i, k - 1));
      else
        temp = ND(o->wts, i, k) * (ND(f1, i, k) * ND(DZ, i, k - 1) + ND(f1, i, k - 1) * ND(DZ, i, k)) / (ND(DZ, i, k) + ND(DZ, i, k - 1)) -
               ND(o->wts, i, k + 1) * (ND(f1, i, k) * ND(DZ, i, k + 1) + ND(f1, i, k + 1) * ND(DZ, i, k)) / (ND(DZ, i, k) + ND(DZ, i, k + 1));
      if (G->isonb[i] == 2) ND(xflux, i, k) = temp * G->art1[i];
      else ND(xflux, i, k) = ND(xflux, i, k) + temp * G->art1[i];
    }
  if (!o->p.river_ts_specified) {
    for (int j = 0; j < o->NUMQBC; ++j) {
      double stpoint = fdis[j];
      if (G->icellq[j] > 0) {
        int j1 = N_ICELLQ(j, 1), j2 = N_ICELLQ(j, 2);
        for (int k = 1; k <= KBM1; ++k) {
          ND(xflux, j1, k) = ND(xflux, j1, k) - o->qdis[j] * RDISQ(j, 1) * VQDIST(j, k) * stpoint;
          ND(xflux, j2, k) = ND(xflux, j2, k) - o->qdis[j] * RDISQ(j, 2) * VQDIST(j, k) * stpoint;
        }
        continue;
      }
      int jj = G->inodeq[j];
      for (int k = 1; k <= KBM1; ++k) ND(xflux, jj, k) = ND(xflux, jj, k) - o->qdis[j] * VQDIST(j, k) * stpoint;
    }
  }
  for (int i = 1; i <= M; ++i)
    for (int k = 1; k <= KBM1; ++k) {
      if (which == 1) ND(xflux, i, k) = ND(xflux, i, k) - 0.0 * G->art1[i];  /* RF = 0: no heating */
      ND(ff, i, k) = (ND(f1, i, k) - ND(xflux, i, k) / G->art1[i] * (o->dti / (o->dt[i] * ND(DZ, i, k)))) * (o->dt[i] / o->dtfa[i]);
    }
}

/* FCT_T / FCT_S, FVCOM_source/fct_t.F:33-108, fct_s.F */
static void fct_ts(orc* o, const double* f1, double* ff) {
  const int KBM1 = o->KBM1;
  const double* DZ = G->dz;
  for (int i = 1; i <= o->M; ++i) {
    if (fct_skip_node(o, i)) continue;
    for (int k = 1; k <= KBM1; ++k) {
      double tmax = ND(f1, NBSN(i, 1), k), tmin = tmax;
      for (int j = 2; j <= NTSN(i); ++j) { double x = ND(f1, NBSN(i, j), k); tmax = fmax(tmax, x); tmin = fmin(tmin, x); }
      if (k == 1) {
        double a = (ND(f1, i, k) * ND(DZ, i, k + 1) + ND(f1, i, k + 1) * ND(DZ, i, k)) / (ND(DZ, i, k) + ND(DZ, i, k + 1));
        tmax = fmax(tmax, a); tmin = fmin(tmin, a);
      } else if (k == KBM1) {
        double a = (ND(f1, i, k) * ND(DZ, i, k - 1) + ND(f1, i, k - 1) * ND(DZ, i, k)) / (ND(DZ, i, k) + ND(DZ, i, k - 1));
        tmax = fmax(tmax, a); tmin = fmin(tmin, a);
      } else {
        double a = (ND(f1, i, k) * ND(DZ, i, k - 1) + ND(f1, i, k - 1) * ND(DZ, i, k)) / (ND(DZ, i, k) + ND(DZ, i, k - 1));
        double b = (ND(f1, i, k) * ND(DZ, i, k + 1) + ND(f1, i, k + 1) * ND(DZ, i, k)) / (ND(DZ, i, k) + ND(DZ, i, k + 1));
        tmax = fmax(tmax, fmax(a, b)); tmin = fmin(tmin, fmin(a, b));
      }
      if (tmin - ND(ff, i, k) > 0.0) ND(ff, i, k) = tmin;
      if (ND(ff, i, k) - tmax > 0.0) ND(ff, i, k) = tmax;
    }
  }
}

/* VDIF_TS(NCON,F), FVCOM_source/vdif_ts.F:42-385 (no surface heat flux) */
static void vdif_ts(orc* o, int ncon, double* f) {
  const int M = o->M, KB = o->KB, KBM1 = o->KBM1, KBM2 = o->KBM2;
  const double dti = o->dti, umolpr = o->p.umol * 1.E0;
  const double* DZ = G->dz; const double* DZZ = G->dzz;
  double af[ORC_COL], cf[ORC_COL], vhf[ORC_COL], vhpf[ORC_COL], ffc[ORC_COL];
  for (int i = 1; i <= M; ++i) {
    const double d = o->d[i];
    double khbottom = 0.0;
    for (int k = 2; k <= KBM1; ++k) {
      double fkh = ND(o->kh, i, k);
      if (k == KBM1) khbottom = fkh;
      af[k - 1] = -dti * (fkh + umolpr) / (ND(DZ, i, k - 1) * ND(DZZ, i, k - 1) * d * d);
      cf[k] = -dti * (fkh + umolpr) / (ND(DZ, i, k) * ND(DZZ, i, k - 1) * d * d);
    }
    vhf[1] = af[1] / (af[1] - 1.);
    vhpf[1] = -dti * (0.0 + 0.0 - 0.0 + 0.0 - 0.0) / (-ND(DZ, i, 1) * d) - ND(f, i, 1);
    vhpf[1] = vhpf[1] / (af[1] - 1.);
    for (int k = 2; k <= KBM2; ++k) {
      vhpf[k] = 1. / (af[k] + cf[k] * (1. - vhf[k - 1]) - 1.);
      vhf[k] = af[k] * vhpf[k];
      vhpf[k] = (cf[k] * vhpf[k - 1] - ND(f, i, k) + dti * (0.0 - 0.0) / (d * ND(DZ, i, k))) * vhpf[k];
    }
    for (int k = 1; k <= KBM1; ++k) ffc[k] = ND(f, i, k);
    if (G->isonb[i] != 2) {
      double tmp1 = o->pfpxb[i] * cos(G->sita_gd[i]) + o->pfpyb[i] * sin(G->sita_gd[i]);
      double tmp2 = o->ah_bottom[i] * G->phpn[i];
      double tmp3 = khbottom + umolpr + o->ah_bottom[i] * G->phpn[i] * G->phpn[i];
      double tmp = tmp1 * tmp2 / tmp3 * (khbottom + umolpr);
      if (o->p.noflux_bot) {
        if (ncon == 1) { if (tmp1 > 0.0) tmp = 0.0; }
        else { if (tmp1 < 0.0) tmp = 0.0; tmp = -tmp; }
      } else tmp = 0.0;
      ffc[KBM1] = (cf[KBM1] * vhpf[KBM2] - ffc[KBM1] - 0.0 + dti * (0.0 - 0.0 - tmp) / (d * ND(DZ, i, KBM1))) /
                  (cf[KBM1] * (1. - vhf[KBM2]) - 1.);
      for (int k = 2; k <= KBM1; ++k) { int ki = KB - k; ffc[ki] = (vhf[ki] * ffc[ki + 1] + vhpf[ki]); }
    }
    for (int k = 1; k <= KBM1; ++k) ND(f, i, k) = ffc[k];
  }
}

/* BCOND_TS(NCON2), FVCOM_source/bcond_ts.F:28-268 with the perturbation
 * forms TYPE_TSOBC 1-3 of BCOND_T/S_PERTURBATION (mod_obcs.F:1223-1373) */
static void bcond_ts(orc* o, int ncon) {
  const int KBM1 = o->KBM1;
  const double* DZ = G->dz;
  double* ff = ncon == 1 ? o->tf1 : o->sf1;
  const double* f1 = ncon == 1 ? o->t1 : o->s1;
  if (o->p.river_ts_specified) {
    for (int i = 0; i < o->NUMQBC; ++i) {
      if (G->icellq[i] > 0) {
        int j11 = N_ICELLQ(i, 1), j22 = N_ICELLQ(i, 2);
        for (int k = 1; k <= KBM1; ++k) {
          ND(o->tf1, j11, k) = G->tdis[i]; ND(o->sf1, j11, k) = G->sdis[i];
          ND(o->tf1, j22, k) = G->tdis[i]; ND(o->sf1, j22, k) = G->sdis[i];
        }
        continue;
      }
      int j11 = G->inodeq[i];
      for (int k = 1; k <= KBM1; ++k) { ND(o->tf1, j11, k) = G->tdis[i]; ND(o->sf1, j11, k) = G->sdis[i]; }
    }
  }
  if (!o->p.obc_on) return;
  double ttmp[ORC_COL];
  for (int i = 0; i < o->IOBCN; ++i) {
    int j = G->i_obc_n[i], j1 = G->next_obc[i];
    double t2d = 0.0, t2d_next = 0.0, xflux2d = 0.0;
    for (int k = 1; k <= KBM1; ++k) {
      t2d = t2d + ND(f1, j, k) * ND(DZ, j, k);
      t2d_next = t2d_next + ND(ff, j1, k) * ND(DZ, j1, k);
      xflux2d = xflux2d + o->xflux_obc[(size_t)(k - 1) * o->IOBCN + i];
    }
    if (o->uard_obcn[i] > 0.0) {
      double tmp = xflux2d + t2d * o->uard_obcn[i];
      double t2d_obc = (t2d * o->dt[j] - tmp * o->dti / G->art1[j]) / o->d[j];
      double cc = sqrt(G->grav_n[j] * G->h[j]) * o->dti / G->dltn_obc[i], cp = cc + 1.0;
      if (o->p.type_tsobc == 4) {
        /* Orlanski radiation of the perturbation, BCOND_T/S_PERTURBATION
         * case 4 (mod_obcs.F:1258-1285, 1336-1363) */
        const double* m1a = ncon == 1 ? o->t1m1 : o->s1m1;
        const double* m2a = ncon == 1 ? o->t1m2 : o->s1m2;
        int j2 = G->next_obc2[i];
        double t2d_next1 = 0.0, tm12d_next2 = 0.0, tm12d_next1 = 0.0, tm22d_next1 = 0.0;
        for (int k = 1; k <= KBM1; ++k) {
          t2d_next1 = t2d_next1 + ND(f1, j1, k) * ND(DZ, j1, k);
          tm12d_next2 = tm12d_next2 + ND(m1a, j2, k) * ND(DZ, j2, k);
          tm12d_next1 = tm12d_next1 + ND(m1a, j, k) * ND(DZ, j, k);
          tm22d_next1 = tm22d_next1 + ND(m2a, j1, k) * ND(DZ, j1, k);
        }
        for (int k = 1; k <= KBM1; ++k) {
          double cl = ((ND(m2a, j1, k) - tm22d_next1) - (ND(f1, j1, k) - t2d_next1)) /
                      ((ND(f1, j1, k) - t2d_next1) + (ND(m2a, j1, k) - tm22d_next1) -
                       2.0 * (ND(m1a, j2, k) - tm12d_next2));
          double mu;
          if (cl >= 1.0) mu = 1.0;
          else if (cl > 0.0 && cl < 1.0) mu = cl;
          else mu = 0.0;
          ttmp[k] = ((ND(m1a, j, k) - tm12d_next1) * (1.0 - mu) + 2.0 * mu * (ND(f1, j1, k) - t2d_next1)) / (1.0 + mu);
        }
      } else {
        for (int k = 1; k <= KBM1; ++k) {
          double pert_next = ND(ff, j1, k) - t2d_next, pert = ND(f1, j, k) - t2d;
          if (o->p.type_tsobc == 1) ttmp[k] = ND(ff, j1, k) - t2d_next;
          else if (o->p.type_tsobc == 2) ttmp[k] = (cc * pert_next + pert) / cp;
          else ttmp[k] = (cc * pert_next + pert * (1.0 - o->dti / 10800.0)) / cp;
        }
      }
      for (int k = 1; k <= KBM1; ++k) ND(ff, j, k) = t2d_obc + ttmp[k];
      for (int k = 1; k <= KBM1; ++k) {
        double tmax = ND(f1, NBSN(j, 1), k), tmin = tmax;
        for (int jj = 2; jj <= NTSN(j); ++jj) { double x = ND(f1, NBSN(j, jj), k); tmax = fmax(tmax, x); tmin = fmin(tmin, x); }
        if (k == 1) {
          double a = (ND(f1, j, k) * ND(DZ, j, k + 1) + ND(f1, j, k + 1) * ND(DZ, j, k)) / (ND(DZ, j, k) + ND(DZ, j, k + 1));
          tmax = fmax(tmax, a); tmin = fmin(tmin, a);
        } else if (k == KBM1) {
          double a = (ND(f1, j, k) * ND(DZ, j, k - 1) + ND(f1, j, k - 1) * ND(DZ, j, k)) / (ND(DZ, j, k) + ND(DZ, j, k - 1));
          tmax = fmax(tmax, a); tmin = fmin(tmin, a);
        } else {
          double a = (ND(f1, j, k) * ND(DZ, j, k - 1) + ND(f1, j, k - 1) * ND(DZ, j, k)) / (ND(DZ, j, k) + ND(DZ, j, k - 1));
          double b = (ND(f1, j, k) * ND(DZ, j, k + 1) + ND(f1, j, k + 1) * ND(DZ, j, k)) / (ND(DZ, j, k) + ND(DZ, j, k + 1));
          tmax = fmax(tmax, fmax(a, b)); tmin = fmin(tmin, fmin(a, b));
        }
        if (tmin - ND(ff, j, k) > 0.0) ND(ff, j, k) = tmin;
        if (ND(ff, j, k) - tmax > 0.0) ND(ff, j, k) = tmax;
      }
    } else {
      for (int k = 1; k <= KBM1; ++k) ND(ff, j, k) = ND(f1, j, k);
    }
  }
  for (int k = 1; k <= KBM1; ++k) { EL3(o->t, 0, k) = 0.0; EL3(o->s, 0, k) = 0.0; }
}

/* ADJUST_TS, FVCOM_source/adjust_ts.F:150-172, called when RIVER_TS_SETTING
 * is 'calculated' (internal_step.F:1374-1376): a river node's T1/S1 on every
 * layer is raised to the river's TDIS/SDIS, then N2E3D of both. */
static void adjust_ts(orc* o) {
  if (o->NUMQBC <= 0) return;
  for (int i = 0; i < o->NUMQBC; ++i) {
    if (G->icellq[i] > 0) {
      int j1 = N_ICELLQ(i, 1), j2 = N_ICELLQ(i, 2);
      for (int k = 1; k <= o->KBM1; ++k) {
        ND(o->t1, j1, k) = fmax(ND(o->t1, j1, k), G->tdis[i]);
        ND(o->t1, j2, k) = fmax(ND(o->t1, j2, k), G->tdis[i]);
        ND(o->s1, j1, k) = fmax(ND(o->s1, j1, k), G->sdis[i]);
        ND(o->s1, j2, k) = fmax(ND(o->s1, j2, k), G->sdis[i]);
      }
      continue;
    }
    int j = G->inodeq[i];
    for (int k = 1; k <= o->KBM1; ++k) {
      ND(o->t1, j, k) = fmax(ND(o->t1, j, k), G->tdis[i]);
      ND(o->s1, j, k) = fmax(ND(o->s1, j, k), G->sdis[i]);
    }
  }
  n2e3d(o, o->t1, o->t);
  n2e3d(o, o->s1, o->s);
}

/* DENS2, FVCOM_source/eqs_of_state.F:254-403 */
static void dens2(orc* o) {
  for (int k = 1; k <= o->KB; ++k)
    for (int i = 0; i <= o->M; ++i) {
      double s = ND(o->s1, i, k), t = ND(o->t1, i, k);
      double r = s * s * s * 6.76786136E-6 - s * s * 4.8249614E-4 + s * 8.14876577E-1 - 0.22584586E0;
      r = r * (t * t * t * 1.667E-8 - t * t * 8.164E-7 + t * 1.803E-5);
      r = r + 1.0 - t * t * t * 1.0843E-6 + t * t * 9.8185E-5 - t * 4.786E-3;
      r = r * (s * s * s * 6.76786136E-6 - s * s * 4.8249614E-4 + s * 8.14876577E-1 + 3.895414E-2);
      r = r - (t - 3.98) * (t - 3.98) * (t + 283.0) / (503.57 * (t + 67.26));
      ND(o->rho1, i, k) = r * 1.e-3;
    }
  for (int i = 0; i <= o->M; ++i) ND(o->rho1, i, o->KB) = 0.0;
  for (int k = 1; k <= o->KB; ++k) ND(o->rho1, 0, k) = 0.0;
  n2e3d(o, o->rho1, o->rho);
}

/* INTERNAL_STEP, FVCOM_source/internal_step.F:29-1607 */
void orc_internal_step(orc* o) {
  const int KB = o->KB;
  o->iint += 1;
  o->int_time += o->dti;
  if (o->p.iramp != 0) o->ramp = fv_tanh((double)o->iint / (double)o->p.iramp);
  else o->ramp = 1.0;
  adjust2d3d(o, 1);
  for (int i = 1; i <= o->N; ++i)                                   /* BCOND_GCN(5,0) */
    for (int k = 1; k <= o->KBM1; ++k) wall_project(o, i, &EL3(o->u, i, k), &EL3(o->v, i, k));
  bcond_forcing_int(o);                                             /* BCOND_GCN(8,0) */
  bottom_roughness(o);
  advection_edge_gcn(o, o->advx, o->advy);
  if (!o->p.barotropic) baropg(o);
  for (int i = 0; i <= o->N; ++i) { o->adx2d[i] = o->ady2d[i] = o->drx2d[i] = o->dry2d[i] = 0.0; }
  for (int k = 1; k <= o->KBM1; ++k)
    for (int i = 1; i <= o->N; ++i) {
      o->adx2d[i] = o->adx2d[i] + EL3(o->advx, i, k);
      o->ady2d[i] = o->ady2d[i] + EL3(o->advy, i, k);
      o->drx2d[i] = o->drx2d[i] + EL3(o->drhox, i, k);
      o->dry2d[i] = o->dry2d[i] + EL3(o->drhoy, i, k);
    }
  advave_edge_gcn(o, o->advua, o->advva);
  for (int i = 0; i <= o->N; ++i) { o->adx2d[i] = o->adx2d[i] - o->advua[i]; o->ady2d[i] = o->ady2d[i] - o->advva[i]; }
  memset(o->uard, 0, sizeof(double) * N1);
  memset(o->vard, 0, sizeof(double) * N1);
  memset(o->egf, 0, sizeof(double) * M1);
  memset(o->uard_obcn, 0, sizeof(double) * (o->IOBCN ? o->IOBCN : 1));
  for (int iext = 1; iext <= o->p.isplit; ++iext) {
    o->iext = iext;
    /* ExtTime = ExtTime + IMDTE; the TIME type is exact, so use the closed form */
    o->ext_time = o->p.time0 + (double)(o->iint - 1) * o->dti + (double)iext * o->dte;
    orc_external_step(o);
  }
  adjust2d3d(o, 2);
  vertvl_edge(o);
  viscof_h(o);
  adv_uv_edge_gcn(o);
  vdif_uv(o);
  for (int i = 1; i <= o->N; ++i)                                   /* BCOND_GCN(3,0) */
    for (int k = 1; k <= o->KBM1; ++k) wall_project(o, i, &EL3(o->uf, i, k), &EL3(o->vf, i, k));
  wreal(o);
  if (o->p.vertical_closure) {
    adv_q(o, o->q2, o->q2f);
    adv_q(o, o->q2l, o->q2lf);
    if (o->p.scalar_positivity) { fct_q(o, o->q2, o->q2f); fct_q(o, o->q2l, o->q2lf); }
    vdif_q(o);
    memcpy(o->q2, o->q2f, sizeof(double) * M1 * KB);
    memcpy(o->q2l, o->q2lf, sizeof(double) * M1 * KB);
  } else {
    for (size_t q = 0; q < M1 * KB; ++q) { o->km[q] = o->p.umol; o->kh[q] = o->p.umol * o->p.vprnu; }
  }
  n2e3d(o, o->km, o->km1);
  if (o->p.temperature_active) {
    adv_ts(o, 1);
    if (o->p.scalar_positivity) fct_ts(o, o->t1, o->tf1);
    vdif_ts(o, 1, o->tf1);
    bcond_ts(o, 1);
    memcpy(o->t1, o->tf1, sizeof(double) * M1 * KB);
    n2e3d(o, o->t1, o->t);
  }
  if (o->p.salinity_active) {
    adv_ts(o, 2);
    if (o->p.scalar_positivity) fct_ts(o, o->s1, o->sf1);
    vdif_ts(o, 2, o->sf1);
    bcond_ts(o, 2);
    memcpy(o->s1, o->sf1, sizeof(double) * M1 * KB);
    n2e3d(o, o->s1, o->s);
  }
  if (!o->p.river_ts_specified) adjust_ts(o);
  if (!o->p.barotropic) dens2(o);
  memcpy(o->u, o->uf, sizeof(double) * N1 * KB);
  memcpy(o->v, o->vf, sizeof(double) * N1 * KB);
  memcpy(o->et, o->el, sizeof(double) * M1);
  memcpy(o->dt, o->d, sizeof(double) * M1);
  memcpy(o->et1, o->el1, sizeof(double) * N1);
  memcpy(o->dt1, o->d1, sizeof(double) * N1);
}

void orc_run(orc* o, long long nsteps) {
  for (long long s = 0; s < nsteps; ++s) orc_internal_step(o);
}

/* Single external step on the current state for 2-D parity tests: sets the
 * clock as INTERNAL_STEP would for IINT+1 and calls EXTERNAL_STEP once. */
void orc_external_only(orc* o, int iext) {
  long long save = o->iint;
  o->iint = save + 1;
  o->iext = iext;
  o->ext_time = o->p.time0 + (double)save * o->dti + (double)iext * o->dte;
  if (o->p.iramp != 0) o->ramp = fv_tanh((double)o->iint / (double)o->p.iramp);
  else o->ramp = 1.0;
  for (int j = 0; j < o->NUMQBC; ++j) o->qdis[j] = G->qdis0[j] * o->ramp;
  orc_external_step(o);
  o->iint = save;
}

/* The FP64 budgets used by the conservation tests. */
void orc_budgets(orc* o, double* volume, double* salt, double* heat) {
  double v = 0, s = 0, h = 0;
  for (int i = 1; i <= o->M; ++i) {
    v += G->art1[i] * o->d[i];
    for (int k = 1; k <= o->KBM1; ++k) {
      double w = G->art1[i] * o->d[i] * ND(G->dz, i, k);
      s += w * ND(o->s1, i, k);
      h += w * ND(o->t1, i, k);
    }
  }
  *volume = v; *salt = s; *heat = h;
}

/* Exported for the accuracy tests of the deterministic elementary functions. */
double orc_fv_cos(double x) { return fv_cos(x); }
double orc_fv_tanh(double x) { return fv_tanh(x); }
double orc_fv_log(double x) { return fv_log(x); }
double orc_fv_expm1(double x) { return fv_expm1(x); }

/* Test hook: one routine of the step on the oracle's current state, so that
 * the tests can pin it by itself against closed forms.  BCOND_GCN(2/3/5)
 * are the wall projections of the step's own loops above.  Returns 1 for an
 * unknown name. */
int orc_routine(orc* o, const char* name, int arg) {
  if (strcmp(name, "viscof_h") == 0) viscof_h(o);
  else if (strcmp(name, "wreal") == 0) wreal(o);
  else if (strcmp(name, "vdif_q") == 0) vdif_q(o);
  else if (strcmp(name, "bcond_ts") == 0) bcond_ts(o, arg);
  else if (strcmp(name, "adjust2d3d") == 0) adjust2d3d(o, arg);
  else if (strcmp(name, "bottom_roughness") == 0) bottom_roughness(o);
  else if (strcmp(name, "baropg") == 0) baropg(o);
  else if (strcmp(name, "vdif_uv") == 0) vdif_uv(o);
  else if (strcmp(name, "advave") == 0) advave_edge_gcn(o, o->advua, o->advva);
  else if (strcmp(name, "advection") == 0) advection_edge_gcn(o, o->advx, o->advy);
  else if (strcmp(name, "adv_uv") == 0) adv_uv_edge_gcn(o);
  else if (strcmp(name, "adv_q") == 0) adv_q(o, o->q2, o->q2f);
  else if (strcmp(name, "adv_ts") == 0) adv_ts(o, arg);
  else if (strcmp(name, "extuv") == 0) extuv_edge(o, arg);
  else if (strcmp(name, "vertvl") == 0) vertvl_edge(o);
  else if (strcmp(name, "fct_ts") == 0) { if (arg == 1) fct_ts(o, o->t1, o->tf1); else fct_ts(o, o->s1, o->sf1); }
  else if (strcmp(name, "fct_q") == 0) { if (arg == 1) fct_q(o, o->q2, o->q2f); else fct_q(o, o->q2l, o->q2lf); }
  else if (strcmp(name, "adjust_ts") == 0) adjust_ts(o);
  else if (strcmp(name, "bcond_gcn2") == 0)
    for (int i = 1; i <= o->N; ++i) wall_project(o, i, &o->uaf[i], &o->vaf[i]);
  else if (strcmp(name, "bcond_gcn5") == 0)
    for (int i = 1; i <= o->N; ++i)
      for (int k = 1; k <= o->KBM1; ++k) wall_project(o, i, &EL3(o->u, i, k), &EL3(o->v, i, k));
  else return 1;
  return 0;
}
