/*
 * fvref.h -- CPU oracle for fosite's unsplit finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker.  The product (fosite_amd) never links,
 * loads or calls it.
 *
 * Storage follows the reference's marray layout (mesh/marray_base.f90:255-300):
 * one 2-D mesh with 2 ghost layers per side in x and y, cell index i fastest,
 * then j, then the field (state vectors) or face/field (face arrays).
 */
#ifndef FVREF_H
#define FVREF_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvr fvr;

typedef struct fvr_cfg {
  int physics;        /* 16 EULER_ISOTHERM, 17 EULER   (physics_base.f90:496-498) */
  double gamma;
  int geometry;       /* 1 CARTESIAN, 31 LOGCYLINDRICAL (geometry_base.f90:345-357) */
  int inum, jnum;
  double xmin, xmax, ymin, ymax, gparam;
  int bc[4];          /* W,E,S,N: 1 NO_GRADIENTS, 2 PERIODIC, 13 ABSORBING (boundary_base.f90:110-130) */
  int limiter;        /* 1 minmod 2 mc 3 sweby 4 superbee 5 ospre 7 vanleer 8 nolimit */
  double theta;
  int recvars;        /* 1 PRIMITIVE, 2 CONSERVATIVE */
  int method, order;  /* 2 RK_FEHLBERG, 3 CASH_KARP, 4 DORMAND_PRINCE, 5 SSPRK */
  double cfl, tol_rel, tol_abs[4], stoptime, dtlimit, beta;
  int fargo;          /* 0 off, 1 mean azimuthal velocity per ring (dynamic), 2 user-supplied */
  int gravity;        /* 1: central point mass at the origin */
  double mass, GN;
  int outbound;       /* accretion through boundary 1 W, 2 E, 3 S, 4 N; 0 off
                         (gravity_pointmass.f90:285-299) */
  double acclimit;    /* accretion limit per unit mass and time; < 0 off */
  double pmin, rhomin;/* CheckData limits, <= 0: TINY (timedisc_base.f90:353-359) */
  int count;          /* datafile count, <= 0: 1 (fileio_base.f90:420-470) */
  double starttime;   /* timedisc_base.f90:295 */
  double omega;       /* angular speed of the rotating frame (mesh_base.f90:338) */
  int rotframe;       /* rotating-frame source (sources_rotframe.f90): 0 off, 1 on,
                         2 on with disable_centaccel; rotation centre at the origin */
  int vismodel;       /* viscosity source (sources_viscosity.f90): 0 off, 1 MOLECULAR, 2 ALPHA */
  double dynconst, bulkconst, cvis; /* dynamic viscosity constant (alpha), bulk viscosity
                         (MOLECULAR), viscous Courant number (sources_viscosity.f90:145-162) */
} fvr_cfg;

typedef struct fvr_scalars {
  int iter, n_adj, brk, dtcause;
  double time, dt, dtold, maxerrold;
} fvr_scalars;

/* returns NULL on an unsupported configuration */
fvr *fvr_create(const fvr_cfg *cfg);
void fvr_destroy(fvr *s);
int fvr_ncells(const fvr *s);
int fvr_vnum(const fvr *s);
/* mesh arrays as the reference holds them, all cells incl. ghosts:
 * id 0 dAxdydz(W) 1 dAxdydz(E) 2 dAydzdx(S) 3 dAydzdx(N) 4 dydzdV 5 dzdxdV
 * 6 dxdydV 7 dlx 8 dly 9 cxyx 10 cyxy 11 czxz 12 czyz 13..16 reconstruction
 * offsets W,E,S,N 17 radius(bcenter) 18..21 radius(faces W,E,S,N)
 * 22 posvec_x 23 posvec_y 24 bcenter x 25 bcenter y 26 accel_x 27 accel_y
 * 28 29 centrifugal acceleration of the rotating frame, x and y */
int fvr_mesh(const fvr *s, int id, double *out);
void fvr_set_sound_speeds(fvr *s, const double *bccsound, const double *fcsound);
void fvr_set_background_velocity(fvr *s, const double *w);
/* primitive initial data, all cells, then Convert2Conservative */
void fvr_set_pvar(fvr *s, const double *pvar);
/* fosite%FirstStep: returns the break flag */
int fvr_first_step(fvr *s);
/* fosite%Step: returns 1 if the run ended (stop time or break) */
int fvr_step(fvr *s);
void fvr_get(const fvr *s, double *cvar, double *pvar, fvr_scalars *sc);
/* Timedisc%w: FARGO background velocity per x-ring (NI values) */
void fvr_get_background_velocity(const fvr *s, double *w);
/* accumulated boundary fluxes Fluxes%bxflux / byflux in the reference's
 * layout (fluxes_base.f90:142-143): bx(JGMIN:JGMAX,1,2,VNUM) = NJ*2*V values,
 * by(1,IGMIN:IGMAX,2,VNUM) = NI*2*V values (side 1 west/south, 2 east/north) */
void fvr_get_bflux(const fvr *s, double *bx, double *by);
/* point mass (gravity_pointmass.f90:79-85): out[0..3] = mass, accrate,
 * massloss, mdot */
void fvr_get_pointmass(const fvr *s, double *out);
/* one ComputeRHS of the current (cvar,pvar) pair without touching the step
 * state's history; rhs out (all cells) -- used for unit-level parity */
void fvr_compute_rhs(fvr *s, double *rhs_out);
/* test aid: FP32 storage emulation (see fvref.c) */
void fvr_round_state_f32(fvr *s);

#ifdef __cplusplus
}
#endif
#endif
