/*
 * fvgpu.h -- C-ABI of the MI355X-native FVCOM mode-split time step.
 *
 * Two shared libraries sit behind this header:
 *
 *   libfvgrid.so  host C++: seeded synthetic cases, the grid topology and
 *                 metrics of the reference's GRID_METRICS
 *                 (FVCOM_source/grid_metrics.F:55-134) and the cold-start
 *                 state (FVCOM_source/mod_startup.F:1062-1181, 1332-1408).
 *   libfvgpu.so   HIP/CDNA4: the device-resident state and the time step
 *                 that replaces INTERNAL_STEP (FVCOM_source/internal_step.F:29)
 *                 and EXTERNAL_STEP (FVCOM_source/external_step.F:30).
 *
 * Array conventions follow the reference's ALL_VARS module
 * (FVCOM_source/mod_main.F:1093-2165): every entity dimension keeps the
 * dummy entry 0 ("outside", NBE=0 / IEC(:,2)=0 point there), indices stored
 * in the arrays are 1-based, 3-D arrays are A(0:MT,1:KB) column-major
 * (entity index fastest, then layer), multi-component arrays such as
 * NV(0:NT,3) or A1U(0:NT,4) are component-major.  Ragged rings NBSN/NBVE
 * (mod_main.F:1783,1791) are stored as CSR in the reference's ring order.
 *
 * Every entry point returns 0 on success and a non-zero status otherwise;
 * fv_last_error() (libfvgrid) and fvgpu_last_error() (libfvgpu) return the
 * message of the last failure on this thread.
 */
#ifndef FVGPU_H
#define FVGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Run parameters (namelist values of NML_INTEGRATION / NML_PHYSICS /  */
/* NML_RIVER_TYPE / NML_OPEN_BOUNDARY_CONTROL, FVCOM_source/mod_main.F */
/* :121-942).                                                          */
/* ------------------------------------------------------------------ */
typedef struct fv_params {
  double  dte;                 /* EXTSTEP_SECONDS                         */
  int32_t isplit;              /* ISPLIT, DTI = ISPLIT*DTE                */
  int32_t iramp;               /* IRAMP (0: RAMP = 1)                     */
  double  umol;                /* VERTICAL_MIXING_COEFFICIENT             */
  double  vprnu;               /* VERTICAL_PRANDTL_NUMBER                 */
  double  hprnu;               /* HORIZONTAL_PRANDTL_NUMBER               */
  double  cbcmin;              /* BOTTOM_ROUGHNESS_MINIMUM                */
  double  min_depth;           /* MIN_DEPTH (DEPTH_CHECK)                 */
  int32_t horizontal_closure;  /* HORIZONTAL_MIXING_TYPE 'closure'=1      */
  int32_t vertical_closure;    /* VERTICAL_MIXING_TYPE 'closure'=1        */
  int32_t scalar_positivity;   /* SCALAR_POSITIVITY_CONTROL (FCT on)      */
  int32_t noflux_bot;          /* NOFLUX_BOT_CONDITION                    */
  int32_t barotropic;          /* BAROTROPIC                              */
  int32_t temperature_active;  /* TEMPERATURE_ACTIVE                      */
  int32_t salinity_active;     /* SALINITY_ACTIVE                         */
  int32_t wind_on;             /* WIND_ON, WIND_TYPE='speed'              */
  int32_t river_ts_specified;  /* RIVER_TS_SETTING == 'specified'         */
  int32_t type_tsobc;          /* OBC_TS_TYPE 1..4                        */
  int32_t obc_on;              /* OBC_ON                                  */
  int32_t obc_elevation_on;    /* OBC_ELEVATION_FORCING_ON (spectral)     */
  int64_t iint0;               /* IINT before the first step              */
  double  time0;               /* IntTime at IINT0, seconds after SpecTime*/
} fv_params;

/* Extents of the model arrays (see fv_extent_len in libfvgrid). */
enum fv_extent {
  FVX_M1 = 0,   /* 0:M                       */
  FVX_M2,       /* CSR offsets over nodes     */
  FVX_N1,       /* 0:N                        */
  FVX_N1x3,     /* (0:N,3)                    */
  FVX_N1x4,     /* (0:N,4)                    */
  FVX_NE1,      /* 0:NE                       */
  FVX_NE1x2,    /* (0:NE,2)                   */
  FVX_CV1,      /* 0:NCV                      */
  FVX_CV1x2,    /* (0:NCV,2)                  */
  FVX_M1K,      /* (0:M,KB)                   */
  FVX_N1K,      /* (0:N,KB)                   */
  FVX_RSN,      /* NBSN ring entries          */
  FVX_RVE,      /* NBVE ring entries          */
  FVX_OBC,      /* IOBCN                      */
  FVX_OBCx2,    /* (IOBCN,2)                  */
  FVX_OBCxT,    /* (IOBCN,NTIDE)              */
  FVX_TIDE,     /* NTIDE                      */
  FVX_RIV,      /* NUMQBC                     */
  FVX_RIVxK,    /* (NUMQBC,KBM1)              */
  FVX_RIVx2,    /* (NUMQBC,2)                 */
  FVX_COUNT
};

/*
 * X(ctype, NAME, extent): the grid, metric, forcing and state arrays of a
 * model, named as in the reference's ALL_VARS.  State arrays hold the
 * initial (or snapshot) values on the host.
 */
#define FV_GRID_FIELDS(X)                                                   \
  /* nodes */                                                               \
  X(double,  vx,        FVX_M1)   X(double,  vy,        FVX_M1)             \
  X(double,  h,         FVX_M1)   X(double,  art1,      FVX_M1)             \
  X(double,  art2,      FVX_M1)   X(double,  grav_n,    FVX_M1)             \
  X(double,  nn_hvc,    FVX_M1)   X(double,  sita_gd,   FVX_M1)             \
  X(double,  phpn,      FVX_M1)   X(int32_t, isonb,     FVX_M1)             \
  X(int32_t, ntsn,      FVX_M1)   X(int32_t, ntve,      FVX_M1)             \
  X(int32_t, nbsn_off,  FVX_M2)   X(int32_t, nbve_off,  FVX_M2)             \
  /* node rings */                                                          \
  X(int32_t, nbsn,      FVX_RSN)  X(double,  dltxtrie,  FVX_RSN)            \
  X(double,  dltytrie,  FVX_RSN)                                            \
  X(int32_t, nbve,      FVX_RVE)  X(int32_t, nbvt,      FVX_RVE)            \
  X(double,  dltxecec,  FVX_RVE)  X(double,  dltyecec,  FVX_RVE)            \
  X(double,  dltxnec,   FVX_RVE)  X(double,  dltynec,   FVX_RVE)            \
  /* elements */                                                            \
  X(double,  xc,        FVX_N1)   X(double,  yc,        FVX_N1)             \
  X(double,  art,       FVX_N1)   X(double,  h1,        FVX_N1)             \
  X(double,  cor,       FVX_N1)   X(double,  cc_hvc,    FVX_N1)             \
  X(double,  cc_sponge, FVX_N1)   X(double,  cc_z0b,    FVX_N1)             \
  X(double,  grav_e,    FVX_N1)   X(double,  epor,      FVX_N1)             \
  X(double,  alpha,     FVX_N1)   X(int32_t, isbce,     FVX_N1)             \
  X(int32_t, iucp,      FVX_N1)   X(int32_t, nv,        FVX_N1x3)           \
  X(int32_t, nbe,       FVX_N1x3) X(double,  a1u,       FVX_N1x4)           \
  X(double,  a2u,       FVX_N1x4) X(double,  awx,       FVX_N1x3)           \
  X(double,  awy,       FVX_N1x3)                                           \
  /* edges */                                                               \
  X(int32_t, iec,       FVX_NE1x2) X(int32_t, ienode,   FVX_NE1x2)          \
  X(int32_t, isbc,      FVX_NE1)  X(double,  xijc,      FVX_NE1)            \
  X(double,  yijc,      FVX_NE1)  X(double,  dltxc,     FVX_NE1)            \
  X(double,  dltyc,     FVX_NE1)                                            \
  /* control-volume sub-edges */                                            \
  X(int32_t, niec,      FVX_CV1x2) X(int32_t, ntrg,     FVX_CV1)            \
  X(double,  dltxe,     FVX_CV1)  X(double,  dltye,     FVX_CV1)            \
  X(double,  dltxncve,  FVX_CV1x2) X(double, dltyncve,  FVX_CV1x2)          \
  /* sigma coordinates */                                                   \
  X(double,  z,         FVX_M1K)  X(double,  zz,        FVX_M1K)            \
  X(double,  dz,        FVX_M1K)  X(double,  dzz,       FVX_M1K)            \
  X(double,  z1,        FVX_N1K)  X(double,  zz1,       FVX_N1K)            \
  X(double,  dz1,       FVX_N1K)  X(double,  dzz1,      FVX_N1K)            \
  /* open boundary; TYPE_OBC per node (mod_obcs.F:30-39): 1/2 elevation   \
   * specified (ASL), 3/4 zero elevation (ASL_CLP), 5/6 gravity-wave       \
   * radiation (GWI), 7/8 Blumberg-Kantha (BKI), 9/10 Orlanski (ORE); odd  \
   * types make the node's adjacent cells linear (IUCP = 0).  Built with   \
   * type 1.  NEXT_OBC2: the interior neighbour of NEXT_OBC most normal to \
   * the boundary (mod_obcs.F:1072-1101)                                   */ \
  X(int32_t, i_obc_n,   FVX_OBC)  X(int32_t, next_obc,  FVX_OBC)            \
  X(int32_t, type_obc,  FVX_OBC)  X(int32_t, next_obc2, FVX_OBC)            \
  X(int32_t, nadjc_obc, FVX_OBC)  X(int32_t, adjc_obc,  FVX_OBCx2)          \
  X(double,  fluxf_obc, FVX_OBCx2) X(double, dltn_obc,  FVX_OBC)            \
  X(double,  emean,     FVX_OBC)  X(double,  apt,       FVX_OBCxT)          \
  X(double,  phai,      FVX_OBCxT) X(double, period,    FVX_TIDE)           \
  /* rivers: a node river (RIVER_INFLOW_LOCATION = 'node') has its node   \
   * in INODEQ and ICELLQ = 0; an edge river ('edge') has its cell in      \
   * ICELLQ, INODEQ = 0, the nodes of its solid edge in N_ICELLQ and their \
   * shares of the discharge in RDISQ (mod_force.F:3515-3568)              */ \
  X(int32_t, inodeq,    FVX_RIV)  X(double,  qarea,     FVX_RIV)            \
  X(int32_t, icellq,    FVX_RIV)  X(int32_t, n_icellq,  FVX_RIVx2)          \
  X(double,  rdisq,     FVX_RIVx2)                                          \
  X(double,  angleq,    FVX_RIV)  X(double,  qdis0,     FVX_RIV)            \
  X(double,  tdis,      FVX_RIV)  X(double,  sdis,      FVX_RIV)            \
  X(double,  vqdist,    FVX_RIVxK)                                          \
  /* wind speed at elements (WIND_TYPE = 'speed'): UUWIND/VVWIND at       \
   * IntTime of the coming internal step, UUWIND0/VVWIND0 at its start     \
   * (IntTime - DTI); the external steps take the linear interpolation at  \
   * ExtTime (UPDATE_WIND, bcond_gcn.F:878-895), the internal surface     \
   * stress the IntTime value (bcond_gcn.F:661-678)                        */ \
  X(double,  uuwind,    FVX_N1)   X(double,  vvwind,    FVX_N1)             \
  X(double,  uuwind0,   FVX_N1)   X(double,  vvwind0,   FVX_N1)

/* Prognostic / persistent state carried between internal steps. */
#define FV_STATE_FIELDS(X)                                                  \
  X(double, el,   FVX_M1)  X(double, et,   FVX_M1)                          \
  X(double, d,    FVX_M1)  X(double, dt,   FVX_M1)                          \
  X(double, ua,   FVX_N1)  X(double, va,   FVX_N1)                          \
  X(double, el1,  FVX_N1)  X(double, et1,  FVX_N1)                          \
  X(double, d1,   FVX_N1)  X(double, dt1,  FVX_N1)                          \
  X(double, u,    FVX_N1K) X(double, v,    FVX_N1K)                         \
  X(double, w,    FVX_N1K) X(double, ww,   FVX_N1K)                         \
  X(double, t,    FVX_N1K) X(double, s,    FVX_N1K)                         \
  X(double, rho,  FVX_N1K) X(double, rmean, FVX_N1K)                        \
  X(double, km1,  FVX_N1K)                                                  \
  X(double, t1,   FVX_M1K) X(double, s1,   FVX_M1K)                         \
  X(double, rho1, FVX_M1K) X(double, rmean1, FVX_M1K)                       \
  X(double, tmean1, FVX_M1K) X(double, smean1, FVX_M1K)                     \
  X(double, q2,   FVX_M1K) X(double, q2l,  FVX_M1K)                         \
  X(double, l,    FVX_M1K) X(double, km,   FVX_M1K)                         \
  X(double, kh,   FVX_M1K) X(double, kq,   FVX_M1K)                         \
  X(double, wts,  FVX_M1K)                                                  \
  /* the elevation one and two RK stages back (ASSIGN_ELM1_TO_ELM2,        \
   * mod_obcs.F:178-199), kept when an Orlanski OBC node needs them        */ \
  X(double, elm1, FVX_M1)  X(double, elm2, FVX_M1)

#define FV_ALL_FIELDS(X) FV_GRID_FIELDS(X) FV_STATE_FIELDS(X)

/* Host model: grid + metrics + forcing + state, owned by libfvgrid. */
typedef struct fv_model {
  int32_t m, n, ne, ncv, ncv_i, kb, kbm1, kbm2;
  int32_t iobcn, ntide, numqbc, nring_sn, nring_ve;
  fv_params p;
#define FV_DECL_FIELD(T, NAME, EXT) T* NAME;
  FV_ALL_FIELDS(FV_DECL_FIELD)
#undef FV_DECL_FIELD
} fv_model;

/* Synthetic case description (the seeded configurations). */
typedef struct fv_case {
  int32_t nx, ny;            /* structured node counts along x and y      */
  double  dx;                /* node spacing (m)                          */
  double  jitter;            /* uniform node displacement, fraction of dx */
  int32_t kb;                /* sigma levels (KB)                         */
  int32_t depth_kind;        /* 0: linear 5 m (west) .. 20 m (east);
                                1: 5+195*(x/Lx)^1.5 plus seamounts        */
  int32_t n_seamounts;
  int32_t ntide;             /* 1: M2; 3: M2,S2,K1 (seeded phases)        */
  double  tide_amp[3];
  int32_t river;             /* 1: one river at the west-end node         */
  double  river_q, river_t, river_s;
  int32_t wind;              /* 1: seeded Gaussian random field           */
  double  wind_mean, wind_std, wind_corr;
  int32_t ts_kind;           /* 0: T=15, S=30*x/Lx ;
                                1: T=20-10*(z/h)... S=33+2*(z/h)          */
  double  el_pert_std, el_pert_corr;  /* initial elevation perturbation   */
  double  coriolis;          /* f-plane Coriolis parameter (s^-1)         */
  double  sponge_radius;     /* sponge around every OBC node (m), 0: none  */
  double  sponge_coef;       /* its damping coefficient at the node        */
  uint64_t seed;
  fv_params p;
} fv_case;

/* ------------------------------- libfvgrid ------------------------------ */
const char* fv_last_error(void);

/* Fill *c with the defaults of a named configuration:
 *   0 estuary channel 81x21 @500 m, KB=11 (3200 elements)
 *   1 jittered 1001x501 @200 m, KB=41, seamounts, M2/S2/K1, wind
 *   3 ensemble member 501x251 @200 m, KB=21                              */
int  fv_case_defaults(int32_t config, fv_case* c);
/* Build grid, metrics, forcing and cold-start state of a synthetic case. */
int  fv_model_build(const fv_case* c, fv_model** out);
/* Build from an explicit mesh: NV(1:N,1:3) clockwise, 1-based, row-major
 * triples; OBC node list 1-based.  Depth h(1:M).  Forcing arrays zero.   */
int  fv_model_from_mesh(int32_t m, int32_t n, const double* x, const double* y,
                        const int32_t* nv_rows, const double* h, int32_t kb,
                        int32_t iobcn, const int32_t* obc_nodes,
                        const fv_params* p, fv_model** out);
void fv_model_free(fv_model* mdl);
int64_t fv_extent_len(const fv_model* mdl, int32_t extent);
/* Reverse Cuthill-McKee numbering of a mesh (neighbours at neighbouring
 * addresses for the gathers; the reference keeps the grid file's numbering): nodes by
 * RCM of the node graph from a pseudo-peripheral node per component,
 * elements by their lowest/middle/highest new node index.  node_new[j-1]
 * and elem_new[e-1] receive the 1-based new index of old node j and old
 * element e; a host renumbers its mesh with them before
 * fv_model_from_mesh, so that the gather kernels read neighbouring
 * entities from neighbouring addresses.                                   */
int  fv_rcm_order(int32_t m, int32_t n, const int32_t* nv_rows, int32_t* node_new, int32_t* elem_new);
/* Look up one array by its ALL_VARS name (lower case).  *dtype: 0 double,
 * 1 int32.  Returns non-zero if the name is unknown.                     */
int  fv_model_field(const fv_model* mdl, const char* name, void** ptr,
                    int64_t* len, int32_t* dtype);
/* Number of named fields, and the i-th name (for iteration).             */
int32_t     fv_model_nfields(void);
const char* fv_model_field_name(int32_t i);
/* Spectral tide of ntide constituents, the same amplitude and phase at
 * every OBC node (BCOND_ASL, mod_obcs.F:552-582).                        */
int  fv_model_set_tide(fv_model* mdl, int32_t ntide, const double* period,
                       const double* amp, const double* phase_deg);
/* Add a river at a solid-boundary node (RIVER_INFLOW_LOCATION='node',
 * mod_force.F:3470-3511): discharge q (m^3/s), temperature t, salinity s,
 * vertical distribution proportional to DZ.                             */
int  fv_model_set_river(fv_model* mdl, int32_t node, double q, double t, double s);
/* Batched instances: the disjoint union of nparts models (same KB, NTIDE,
 * tidal periods and run parameters) as one model, instance p's entities
 * renumbered after those of instances 0..p-1 (no counterpart in the
 * reference, whose EnKF members run one after another through restart
 * files, FVCOM_source/fvcom.F:2006-2128).  Every instance of the union
 * advances exactly as it would alone.                                    */
int  fv_model_union(const fv_model* const* parts, int32_t nparts, fv_model** out);
/* Sponge layer (SETUP_SPONGE, FVCOM_source/mod_setup.F:458-499): for each
 * of the n sponge nodes (1-based) with radius r and coefficient c, every
 * element within r gets CC_SPONGE = max(CC_SPONGE, c*(1 - dist/r)).      */
int  fv_model_set_sponge(fv_model* mdl, int32_t n, const int32_t* nodes, const double* radius,
                         const double* coef);
/* Edge river (RIVER_INFLOW_LOCATION = 'edge', mod_force.F:3515-3568) on
 * boundary cell `cell` (1-based, ISBCE = 1, exactly one solid edge): its
 * discharge enters through the solid edge, split between the edge's two
 * nodes by their control-volume areas; momentum enters the cell.         */
int  fv_model_set_river_edge(fv_model* mdl, int32_t cell, double q, double t, double s);
/* TYPE_OBC of the n = IOBCN open-boundary nodes, each 1..10 (mod_obcs.F:
 * 30-39), and IUCP recomputed: 0 in
 * the adjacent cells of odd-type nodes, 1 elsewhere (mod_obcs.F:959-962,
 * 993-996).  Replaces the reference's TYPE_OBC column of the OBC file.   */
int  fv_model_set_obc_type(fv_model* mdl, int32_t n, const int32_t* types);
/* Cold start: EL=0, rest, DENS2 density, RMEAN/TMEAN/SMEAN, default
 * turbulence (mod_startup.F:1062-1181, 1332-1408).                       */
int  fv_model_cold_start(fv_model* mdl);

/* ------------------------------- libfvgpu ------------------------------- */
typedef struct fvgpu_handle fvgpu_handle;

/* Message of the last failed libfvgpu call on this thread. */
const char* fvgpu_last_error(void);

/* Diagnostics gathered on the device (SHUTDOWN_CHECK, DEPTH_CHECK and the
 * FP64 budgets; FVCOM_source/mod_utils.F:320-432, depth_check.F).        */
typedef struct fv_diag {
  int32_t nonfinite;      /* 1 if any of EL, UA, VA, U, V, T1, S1 is not finite */
  int32_t depth_fail;     /* 1 if MIN(D) < MIN_DEPTH was seen                   */
  double  min_depth;      /* minimum total depth D over nodes                    */
  double  volume;         /* sum ART1*D over nodes (m^3), compensated sum        */
  double  salt;           /* sum ART1*D*DZ*S1 over node layers (psu m^3)         */
  double  heat;           /* sum ART1*D*DZ*T1 over node layers (degC m^3)        */
  int64_t iint;           /* internal steps completed (IINT)                     */
} fv_diag;

/* Precision modes. */
enum fv_precision { FV_FP64 = 0, FV_MIXED = 1 };

/* Upload grid, metrics, forcing and state; the device owns the copies.
 * Replaces ALLOC_VARS + GRID_METRICS + STARTUP for the device state.     */
int  fvgpu_create(const fv_model* mdl, int32_t precision, fvgpu_handle** out);
/* One call of INTERNAL_STEP (internal_step.F:29), ISPLIT EXTERNAL_STEPs
 * included.  Advances IINT by one.                                       */
int  fvgpu_internal_step(fvgpu_handle* h);
/* nsteps internal steps; with use_graph!=0 each step is one hipGraph
 * launch captured on first use.                                          */
int  fvgpu_run(fvgpu_handle* h, int64_t nsteps, int32_t use_graph);
/* One call of EXTERNAL_STEP (external_step.F:30) on the current state,
 * for parity tests of the 2-D mode alone.  iext is IEXT (1..ISPLIT).     */
int  fvgpu_external_step(fvgpu_handle* h, int32_t iext);
/* Copy one state or grid array (ALL_VARS name, lower case) to/from the host
 * (double precision on the host side whatever the device precision).     */
int  fvgpu_get_state(fvgpu_handle* h, const char* name, double* host, int64_t len);
int  fvgpu_put_state(fvgpu_handle* h, const char* name, const double* host, int64_t len);
int  fvgpu_diagnostics(fvgpu_handle* h, fv_diag* out);
/* EnKF ensemble update A <- A*X5 (the update that ends ANALYSIS,
 * mod_enkfassim.F:2746-3010, applied by ENKF_ASS_EVE to every member's
 * state, mod_enkfassim.F:308-753) for one state array of a batched union
 * (fv_model_union), in place on the device: member i (0-based, nrens <= 64)
 * owns nodes node_off[i]+1 .. node_off[i]+m and elements elem_off[i]+1 ..
 * elem_off[i]+n of the union; x5 is nrens x nrens, column-major
 * (X5(i,j) at x5[i + j*nrens]).  Each new value is the sum over i in
 * increasing order of A(p,i)*X5(i,j).                                    */
int  fvgpu_ensemble_update(fvgpu_handle* h, const char* name, int32_t nrens, const int64_t* node_off,
                           const int64_t* elem_off, int64_t m, int64_t n, const double* x5);
/* fvgpu_ensemble_update without any device allocation: X5 and the member
 * offsets are staged in a fixed 33,280-byte device array of the library
 * (one per GPU, looked up by fvgpu_create) instead of a buffer the handle
 * allocates on its first update; the same kernel and results.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  nrens <= 64, as for fvgpu_ensemble_update.  Every call on
 * one GPU shares the staging array: calls that overlap on one GPU race.   */
int  fvgpu_ensemble_update_noalloc(fvgpu_handle* h, const char* name, int32_t nrens, const int64_t* node_off,
                                   const int64_t* elem_off, int64_t m, int64_t n, const double* x5);
/* Time-varying forcing between internal steps, taking effect from the
 * next step.  The reference interpolates its forcing files in time inside
 * the step (UPDATE_RIVERS at IntTime, bcond_gcn.F:633-641; UPDATE_WIND and
 * ASIMPLE_DRAG at IntTime and at each ExtTime, bcond_gcn.F:661-665,
 * 878-882); here the caller supplies the values at the coming step's
 * IntTime:
 *   "wind"  UUWIND(0:N) then VVWIND(0:N), element wind speed (m/s) at the
 *           end of the coming internal step; the wind set before (or the
 *           model's, at first) becomes the wind at its start.  The
 *           internal surface stress uses the end value; each external step
 *           uses the speed interpolated linearly in time to its ExtTime
 *           (fraction IEXT/ISPLIT of the step), converted by PSIMPLE_DRAG
 *           (mod_bulk.F:58-90).  The model must have been created with
 *           wind on.
 *   "qdis"  river discharge per river (m^3/s, before RAMP)
 *   "tdis", "sdis"  river temperature / salinity per river               */
int  fvgpu_set_forcing(fvgpu_handle* h, const char* name, const double* host, int64_t len);
/* Kernel-level timing of the last fvgpu_run: average device time of the
 * named kernel family in microseconds, measured with HIP events on the
 * handle's stream.  Returns non-zero if the name is unknown.             */
int  fvgpu_kernel_time(fvgpu_handle* h, const char* name, double* avg_us, int64_t* count);
/* Enumerate the timed kernel families (idx = 0, 1, ... until non-zero):
 * name (valid until the next profiling call), average us and launches.  */
int  fvgpu_kernel_stat(fvgpu_handle* h, int32_t idx, const char** name, double* avg_us, int64_t* count);
int  fvgpu_set_profiling(fvgpu_handle* h, int32_t on);
/* Restrict the timing events to one kernel family (NULL or "": all).     */
int  fvgpu_profile_only(fvgpu_handle* h, const char* name);
/* Time one in `every` launches of the timed families (1: every launch);
 * the averages are then over that sample.                                */
int  fvgpu_profile_every(fvgpu_handle* h, int32_t every);
int  fvgpu_synchronize(fvgpu_handle* h);
/* Measurement aid: average time of a plain device copy of n doubles
 * (k_copy, the same launch shape as the step's copies) over reps launches;
 * calibrates HBM byte counters and gives the achievable copy bandwidth.   */
int  fvgpu_stream_copy(fvgpu_handle* h, int64_t n, int32_t reps, double* avg_us);
void fvgpu_destroy(fvgpu_handle* h);

/* Workgroup size (64, 128 or 256) of one kernel family of the step.      */
int  fvgpu_set_block(fvgpu_handle* h, const char* family, int32_t block);
/* Concurrency of the step's independent chains on a second stream of the
 * handle, a bit mask: 1 the S pass beside the T pass and Q2L advection
 * beside Q2, 2 the 2-D edge terms beside the elevation update of every RK
 * stage (default 3, both); 0 runs every chain on the main stream, so that
 * each kernel family's timing is its own.  No counterpart in the
 * reference; results do not depend on it.                                */
int  fvgpu_set_concurrency(fvgpu_handle* h, int32_t mode);
/* Autotuner: times every kernel family of the step with workgroups of 64,
 * 128 and 256 (`steps` internal steps per candidate), keeps the fastest,
 * and restores the model state afterwards.  `report` receives
 * "family=size:time;..." (may be NULL).                                  */
int  fvgpu_autotune(fvgpu_handle* h, int32_t steps, char* report, int64_t report_len);

/* ---------------------------- learned closure ---------------------------
 * A 5-64-64-2 MLP (SiLU hidden layers, softplus outputs) standing in for
 * the MY2.5 stability functions and length scale that set KM/KH at the end
 * of VDIF_Q (FVCOM_source/vdif_q.F:470-491).  No counterpart exists in the
 * reference; the closure is an option of this library.  Parameters are an
 * FP32 vector of fvmlp_nparams() values: W1[64][5] b1[64] W2[64][64] b2[64]
 * W3[2][64] b3[2].  Samples are 8 floats of normalised inputs (N^2, S^2,
 * z/h, u*, D, 1, 0, 0) and 2 floats of targets (KM, KH).                  */
/* mode 0: MY2.5 (default); mode 1: the MLP with the per-column MY2.5
 * fallback.  The parameters are copied (as BF16 for the MFMA inference). */
int  fvgpu_set_closure(fvgpu_handle* h, int32_t mode, const float* params);
/* Columns that fell back to MY2.5 in the last internal step.            */
int  fvgpu_closure_fallbacks(fvgpu_handle* h, int64_t* fallback_columns, int64_t* columns);
/* Features and the current KM/KH of every node and interface 2..KBM1 of
 * the current state (x: n x 8, y: n x 2); with x or y NULL only *n is set. */
int  fvgpu_closure_samples(fvgpu_handle* h, float* x, float* y, int64_t max_samples, int64_t* n);

/* The trainer: device-resident samples, FP32 master weights, AdamW.       */
typedef struct fvmlp fvmlp;
int32_t fvmlp_nparams(void);
int  fvmlp_create(int64_t n, const float* x, const float* y, int32_t bmax, fvmlp** out);
int  fvmlp_set_params(fvmlp* t, const float* params);   /* also resets AdamW */
int  fvmlp_get_params(fvmlp* t, float* params);
/* Mean loss and its gradient over the samples idx[0..B) (NULL: 0..B-1).  */
int  fvmlp_grad(fvmlp* t, const int32_t* idx, int32_t B, float* grad, float* loss);
/* `steps` AdamW steps on seeded minibatches drawn from samples [0, ntrain);
 * *last_loss is the last minibatch's mean loss.                          */
int  fvmlp_train(fvmlp* t, int32_t steps, int32_t batch, int64_t ntrain, float lr, float wd, uint64_t seed,
                 float* last_loss);
/* Predicted (KM, KH) of samples [lo, lo+n): FP32 SIMT or BF16 MFMA.       */
int  fvmlp_infer(fvmlp* t, int64_t lo, int64_t n, int32_t use_mfma, float* y);
void fvmlp_destroy(fvmlp* t);
int  fvgpu_device_count(int32_t* n);

#ifdef __cplusplus
}
#endif
#endif /* FVGPU_H */
