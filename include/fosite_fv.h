/*
 * fosite_fv.h -- C ABI of the MI355X finite-volume step (libfosite_fv.so).
 *
 * One context = one fosite simulation (Mesh + Physics + Fluxes + Boundary +
 * Sources + Timedisc) whose state lives on the current HIP device.  All entry
 * points return 0 on success and a negative code on failure; the message is
 * available from fv_last_error().  The reference has no C interface: each
 * entry point below replaces the Fortran type-bound procedure cited beside it
 * (paths relative to the fosite source tree).
 *
 * Arrays cross the boundary in the reference's marray layout
 * (mesh/marray_base.f90:255-300): ghost cells included (2 per side in x and
 * y), cell index i fastest, then j, then the field (state vectors) or the
 * face (W,E,S,N) for face arrays.  Fields are ordered density, x-momentum,
 * y-momentum[, energy] (conservative) or density, x-velocity, y-velocity
 * [, pressure] (primitive), physics/physics_euler.f90:2603-2613.
 */
#ifndef FOSITE_FV_H
#define FOSITE_FV_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fv_ctx fv_ctx;

/* Configuration: the Dict keys of a fosite setup program (e.g.
 * tests/riemann2d.f90:586-670), flattened.  Integer ids are the reference's. */
typedef struct fv_config {
  int physics;      /* 16 EULER_ISOTHERM, 17 EULER (physics/physics_base.f90:496-498) */
  double gamma;     /* physics/physics_euler.f90:152 */
  int geometry;     /* 1 CARTESIAN, 31 LOGCYLINDRICAL (mesh/geometry_base.f90:345-357) */
  int inum, jnum;   /* resolution, knum = 1 (mesh/mesh_base.f90:350-352) */
  double xmin, xmax, ymin, ymax, gparam;
  int bc[4];        /* western, eastern, southern, northern:
                       1 NO_GRADIENTS, 2 PERIODIC, 13 ABSORBING (boundary/boundary_base.f90:110-130) */
  int limiter;      /* 1 minmod, 2 mc, 3 sweby, 4 superbee, 5 ospre, 7 van leer, 8 no limit
                       (fluxes/reconstruction_linear.f90:71-78) */
  double theta;     /* limiter parameter for mc and sweby */
  int recvars;      /* 1 PRIMITIVE, 2 CONSERVATIVE (fluxes/reconstruction_base.f90:117,
                       fluxes/fluxes_base.f90:378-384) */
  int method, order;/* 2 RK_FEHLBERG (order 3, 5), 3 CASH_KARP (5), 4 DORMAND_PRINCE (5),
                       5 SSPRK (3, 5) (timedisc/timedisc_base.f90:187-191) */
  double cfl, tol_rel, tol_abs[4], stoptime, dtlimit, beta; /* timedisc_base.f90:315-352;
                       dtlimit <= 0 selects the default EPSILON*stoptime */
  int fargo;        /* 0 off, 1 per-ring mean azimuthal velocity recomputed every step
                     (timedisc/timedisc_base.f90:2243-2326),
                     2 user-supplied background velocity (mesh/mesh_base.f90:372-423) */
  int gravity;      /* 1: point mass at the origin, Newtonian
                       (sources/gravity_pointmass.f90:132-301) */
  double mass, GN;  /* point mass and gravitational constant (GEOMETRICAL units: GN = 1) */
  int dtype;        /* 0 = FP64, the plain drop-in (bit-exact with the reference where no
                       libm call is on the path); 1 = FP32 precision form: FP32 state and
                       arithmetic, compensated RK accumulation, FP64 error norm and CFL
                       reduction (no reference counterpart) */
  int outbound;     /* point-mass accretion through boundary 1 WEST, 2 EAST, 3 SOUTH, 4 NORTH,
                       0 off (sources/gravity_pointmass.f90:285-299, 383-428).  Zero-initialised
                       means off; the reference's own default on cylindrical meshes is WEST,
                       so a setup that relies on it sets 1 here. */
  double acclimit;  /* accretion limit per unit mass and time, < 0 off
                       (sources/gravity_pointmass.f90:190, 412-416) */
  double pmin, rhomin; /* CheckData limits and the Euler pressure floor of ComputeRHS
                       (timedisc/timedisc_base.f90:353-359, 1150-1164, 1666-1742);
                       <= 0 selects the default TINY(1.0) */
  int count;        /* datafile count: outputs at starttime + k*|stoptime - starttime|/count
                       (io/fileio_base.f90:420-470, 487, 591-600); <= 0 selects 1 */
  double starttime; /* Timedisc%time at setup (timedisc/timedisc_base.f90:295); a restart
                       sets it to the dataset's time (examples/restart.f90:103-106) */
  double omega;     /* Mesh%omega, angular speed of the rotating frame (mesh/mesh_base.f90:338) */
  int rotframe;     /* rotating-frame source term (sources/sources_rotframe.f90:73-247), centre
                       of rotation at the origin (the default rotation_center, mesh_base.f90:344):
                       0 off, 1 centrifugal and Coriolis accelerations, 2 Coriolis only
                       (disable_centaccel = 1) */
  int vismodel;     /* viscosity source (sources/sources_viscosity.f90:73-577): 0 off,
                       1 MOLECULAR (constant dynamic and bulk viscosity), 2 ALPHA
                       (Shakura-Sunyaev), 3 BETA (Duschl); ALPHA and BETA use omega above;
                       FP64 state, single contexts (not in batches) */
  double dynconst, bulkconst; /* dynamic viscosity (MOLECULAR), alpha (ALPHA) or beta (BETA),
                       default 0.1;
                       bulk viscosity (MOLECULAR), the reference's default being -2/3 dynconst
                       (sources_viscosity.f90:152-158) */
  double cvis;      /* viscous Courant number, default 0.5 (sources_viscosity.f90:165) */
} fv_config;

typedef struct fv_scalars {
  int iter, n_adj, brk, dtcause;
  double time, dt, dtold, maxerrold;
} fv_scalars;

/* Mesh metric query, host only (no context, no device): the array the
 * reference's Mesh holds for configuration `cfg`, all cells incl. ghosts.
 * Ids: 0 dAxdydz(W) 1 dAxdydz(E) 2 dAydzdx(S) 3 dAydzdx(N) 4 dydzdV 5 dzdxdV
 * 6 dxdydV 7 dlx 8 dly 9 cxyx 10 cyxy 11 czxz 12 czyz 13..16 reconstruction
 * offsets W,E,S,N 17 radius(bcenter) 18..21 radius(faces W,E,S,N)
 * 22 posvec_x 23 posvec_y 24 bcenter x 25 bcenter y 26 accel_x 27 accel_y
 * 28 FARGO type-1 wfactor (timedisc/timedisc_base.f90:428-447).
 * Replaces mesh/mesh_midpoint.f90:73-146 (InitMesh_midpoint) and
 * sources/gravity_pointmass.f90:263-282 (static acceleration). */
int fv_mesh_array(const fv_config *cfg, int id, double *out);

/* fosite%Setup (fosite.f90:152-255): validates the configuration, builds the
 * mesh metric on the host and allocates the device-resident state. */
int fv_create(fv_ctx **ctx, const fv_config *cfg);
/* fosite%Finalize (fosite.f90:680-691) */
int fv_destroy(fv_ctx *ctx);
/* ghost-inclusive cell counts and number of conservative fields */
int fv_dims(const fv_ctx *ctx, int *ni, int *nj, int *vnum);
/* Device memory a context with this configuration allocates, in bytes
 * (no reference counterpart; used for admission by memory budget). */
int fv_device_bytes(const fv_config *cfg, double *bytes);
/* SHA-256 of the sources and compile flags the library was built from (no
 * reference counterpart; lets a caller check a prebuilt library against its
 * source tree). */
const char *fv_build_id(void);

/* physics_eulerisotherm%SetSoundSpeeds_center / _faces
 * (physics/physics_eulerisotherm.f90:288-309); isothermal physics only.
 * bccsound: cells; fcsound: cells x 4 faces. */
int fv_set_sound_speeds(fv_ctx *ctx, const double *bccsound, const double *fcsound);
/* Timedisc%w for FARGO type 2 (timedisc/timedisc_base.f90:389,420):
 * one value per x-ring, IGMIN..IGMAX. */
int fv_set_background_velocity(fv_ctx *ctx, const double *w);
/* InitData: Timedisc%pvar = pvar, then Physics%Convert2Conservative
 * (physics/physics_eulerisotherm.f90:426-449). */
int fv_set_pvar(fv_ctx *ctx, const double *pvar);
/* Timedisc%cvar (timedisc/timedisc_base.f90:87-88), the conservative state
 * as the host holds it, all cells; no conversion.  The Timedisc-level
 * binding hands the state over this way: fosite's FirstStep has by then
 * replaced pvar with Convert2Primitive(cvar) in CenterBoundary
 * (fosite.f90:307-313), which a second Convert2Conservative would not
 * invert bit for bit. */
int fv_set_cvar(fv_ctx *ctx, const double *cvar);

/* fosite%FirstStep (fosite.f90:258-348): full RHS with data checks, CFL
 * time step, initial output frame change.  Returns the break flag in *brk. */
int fv_first_step(fv_ctx *ctx, int *brk);
/* fosite%Step (fosite.f90:367-464): stop-time test, output-time clipping
 * (io/fileio_base.f90:604-620), Timedisc%IntegrationStep
 * (timedisc/timedisc_base.f90:649-750), next CFL time step
 * (timedisc/timedisc_base.f90:818-909).  *done = 1 when the run ended
 * (stop time reached or break). */
int fv_step(fv_ctx *ctx, int *done);
/* nsteps x fv_step; *taken = steps executed before the run ended */
int fv_run(fv_ctx *ctx, int nsteps, int *taken);
/* fv_step that allocates nothing.  No counterpart in the reference: for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  Its kernels' reduction scratch (per-workgroup
 * partials of the end-of-step evaluation and of the last stage, and the
 * reduction result) is a static device array of the library, one per
 * device, instead of the context's buffers; the context itself still comes
 * from fv_create.  Limit: 16384 workgroups per RHS launch (64 x 8 cells each,
 * 8.4 M cells); larger meshes return an error.  Calls on one device share
 * the array: calls that overlap on one device (other host threads, other
 * streams) race.  Fails while fv_profile is enabled. */
int fv_step_noalloc(fv_ctx *ctx, int *done);

/* The same path cut at the Timedisc object, for a host that keeps fosite's
 * own fosite%Step, Datafile and output (a timedisc_base extension whose
 * IntegrationStep, ComputeRHS and CalcTimestep call these):
 *
 * fv_first_rhs: the device part of fosite%FirstStep (fosite.f90:305-321):
 * CalcBackgroundVelocity for FARGO type 1, Timedisc%ComputeRHS with every
 * data check, Timedisc%CalcTimestep; no output frame change.  *brk =
 * Timedisc%break. */
int fv_first_rhs(fv_ctx *ctx, int *brk);
/* Timedisc%IntegrationStep (timedisc/timedisc_base.f90:649-750) with the
 * host's time step: dt and dtcause as Datafile%AdjustTimestep left them
 * (io/fileio_base.f90:604-620).  No stop test, output step or next CFL step
 * (fosite%Step does those).  *brk = Timedisc%break. */
int fv_integration_step(fv_ctx *ctx, double dt, int dtcause, int *brk);
/* Timedisc%CalcTimestep (timedisc/timedisc_base.f90:818-909): the CFL time
 * step of the last end-of-step evaluation, and its cause. */
int fv_calc_timestep(fv_ctx *ctx, double *dt, int *dtcause);
/* Physics%AddBackgroundVelocityY on Timedisc%pvar/cvar as WriteDataset
 * applies it before writing (io/fileio_base.f90:523-533,
 * physics/physics_eulerisotherm.f90:1263-1303): a no-op when the state is in
 * the real frame already, or without FARGO. */
int fv_add_background_velocity(fv_ctx *ctx);

/* Device -> host copy of Timedisc%cvar and Timedisc%pvar, all cells (either
 * pointer may be NULL).  pvar is the one the last host-controlled ComputeRHS
 * left (timedisc_base.f90:1098-1164: after the PMIN clamp, ghost cells from
 * CenterBoundary), in the state's FARGO frame; every output step is such a
 * step.  After device-controlled steps, or before FirstStep, pvar is
 * Convert2Primitive(cvar). */
int fv_get_state(fv_ctx *ctx, double *cvar, double *pvar);
int fv_get_scalars(const fv_ctx *ctx, fv_scalars *out);
/* Timedisc%w, the FARGO background velocity per x-ring (NI values, ghost
 * rings included; timedisc/timedisc_base.f90:389,420): the value set by
 * fv_set_background_velocity (fargo 2), or the one CalcBackgroundVelocity
 * computed at the start of the last step (fargo 1,
 * timedisc/timedisc_base.f90:2243-2326). */
int fv_get_background_velocity(fv_ctx *ctx, double *w);
/* Fluxes%bxflux and Fluxes%byflux (fluxes/fluxes_base.f90:142-143): the
 * fluxes through the mesh boundaries accumulated over every accepted step,
 * sum_m -dt*b_high(m)*rhs_m at the boundary rows
 * (timedisc/timedisc_rkfehlberg.f90:216-259, timedisc/timedisc_base.f90:
 * 1224-1252, 937-957, 981-1000), in the reference's layout:
 * bx(JGMIN:JGMAX,1,2,VNUM) = NJ*2*VNUM values, by(1,IGMIN:IGMAX,2,VNUM) =
 * NI*2*VNUM values, side 1 west/south, side 2 east/north; ghost entries are
 * zero.  Either pointer may be NULL. */
int fv_get_boundary_flux(fv_ctx *ctx, double *bx, double *by);
/* Fluxes%GetBoundaryFlux (fluxes/fluxes_base.f90:220-300, serial build):
 * direction 1 WEST, 2 EAST, 3 SOUTH, 4 NORTH (mesh/mesh_base.f90:102-105);
 * bflux[VNUM] = SUM over the boundary's interior range, in index order. */
int fv_get_boundary_flux_total(fv_ctx *ctx, int direction, double *bflux);
/* The point mass (sources/gravity_pointmass.f90:79-85, output attributes
 * 192-195): out[0] mass, out[1] accrate, out[2] massloss, out[3] mdot (the
 * GetBoundaryFlux density sum its last update saw).  With accretion the
 * values follow UpdateGravity_single (gravity_pointmass.f90:383-428), run at
 * every ComputeRHS; without it they stay at (mass, 0, 0, 0). */
int fv_get_point_mass(fv_ctx *ctx, double *out);
/* wall time of the device work issued by the last fv_run, in ms
 * (HIP events on the context's stream) */
double fv_last_run_ms(const fv_ctx *ctx);
const char *fv_last_error(const fv_ctx *ctx);

/* Kernel timing (no reference counterpart; measurement support).
 * fv_profile(ctx, 1) resets the statistics and brackets every kernel launch
 * with HIP events on the context's stream (launches are then issued one by
 * one instead of as HIP graphs).  Kernel classes: 0 fused RHS on an RK stage
 * state, 1 end-of-step RHS (CFL, data checks, next step's first slope),
 * 2 last RK stage fused with the final combination and embedded error,
 * 3 FARGO frame change, 4 step control of the device-controlled batches (one
 * block), 5 FARGO azimuthal remap, 6 accretion update of a device-controlled
 * step (boundary-flux accumulation and UpdateGravity, two launches; no
 * algorithmic bytes counted), 7 viscous source of an evaluation (its
 * algorithmic bytes: the stage inputs read and V-1 source words written).
 * fv_kernel_stats
 * returns the number of launches, their summed device time in ms and their
 * algorithmic bytes (state words each launch must read and write, 8 bytes
 * per word, interior cells). */
int fv_profile(fv_ctx *ctx, int enable);
int fv_kernel_stats(const fv_ctx *ctx, int kernel, int *launches, double *ms, double *alg_bytes);

/* ------------------------------------------------------------------ batches
 * No reference counterpart: the reference runs one simulation per process.
 * A batch advances many independent contexts together.  Every kernel of a
 * device-controlled step is ONE launch over the tiles of all members, and
 * the steps of a round are one HIP graph launch.  Each member keeps its own
 * mesh, time step control and host scalars; results are bit-identical to
 * running the member alone with fv_run.  Members must agree in physics,
 * dtype and time integrator (method, order); limiters, boundaries, FARGO,
 * the rotating frame, point-mass accretion and sizes may differ (FARGO type 1
 * members take their steps on the host path).  A context belongs to at most one batch; fv_destroy
 * removes it from its batch.
 *
 * Tilings of the RHS kernels: 0 = 64-wide tiles in row-major order, 1 =
 * 64-wide tiles in XCD-chunked order (block b of the launch runs on XCD
 * b mod 8; each XCD gets a contiguous run of tiles, so neighbouring tiles
 * share that XCD's L2), 2 and 3 = the same with 32-wide tiles.  Every tiling
 * gives the same bits.  With autotune, the first device steps of
 * fv_batch_run are timed under each tiling in turn (2 steps per trial, 2
 * trials per tiling), and the fastest is kept. */
typedef struct fv_batch fv_batch;
typedef struct fv_batch_opts {
  int autotune;        /* 1: choose the tiling by timing (see above) */
  int tiling;          /* tiling without autotune (0..3) */
  int max_graph_steps; /* device steps per graph launch, 1..512 (0 = 512) */
  int reuse_tuning;    /* with autotune: take the choice of an earlier batch of the same class
                          in this process (no trials), and record this batch's for later ones */
} fv_batch_opts;
typedef struct fv_batch_info {
  int members;
  int tiling;          /* tiling in use */
  int tuned;           /* 1 once the autotuner has chosen by its trials, 2 when it took the choice
                          of an earlier batch of the same class (reuse_tuning: element type,
                          fields, stages, log2 of the members' total cells and of their number) */
  double tune_ms[4];   /* best time per device step of each tiling in the trials (0 = not timed) */
  long device_steps;   /* member steps taken on the device-controlled path */
  long host_steps;     /* member steps taken on the host path (first steps, halts) */
  long halts;          /* device rounds a member left for the host path */
  long rounds;         /* graph or trial launches */
  double last_run_ms;  /* wall time of the last fv_batch_run (HIP events, batch stream) */
} fv_batch_info;
int fv_batch_create(fv_batch **b, const fv_batch_opts *opts); /* empty batch (opts may be NULL) */
int fv_batch_add(fv_batch *b, fv_ctx *ctx);
int fv_batch_remove(fv_batch *b, fv_ctx *ctx);
/* nsteps Steps (as fv_run) of every member; taken[k] = steps member k took
 * (members in the order they were added; fewer if its run ended). */
int fv_batch_run(fv_batch *b, int nsteps, int *taken);
int fv_batch_get_info(const fv_batch *b, fv_batch_info *out);
int fv_batch_destroy(fv_batch *b);
const char *fv_batch_last_error(const fv_batch *b);

/* ------------------------------------------------------------------ serving
 * No reference counterpart.  A server runs many simulations to their stop
 * times by continuous batching: requests queue on the host and are admitted
 * in arrival order while their device footprint (fv_device_bytes) fits the
 * memory budget; admitted runs stay on the device and join the batch of
 * their dtype.  fv_server_step is one scheduling round: both batches advance
 * by up to round_steps steps (concurrently, from two host threads), queued
 * requests that fit are prepared meanwhile (context, initial data, first
 * step), then finished runs are copied back and released and the prepared
 * ones join.  A served result is bit-identical to the same run made alone.
 * Times are seconds on the server's steady clock (fv_server_get_stats). */
typedef struct fv_server fv_server;
typedef struct fv_server_opts {
  double hbm_budget_bytes; /* device memory all admitted runs together may hold */
  int round_steps;         /* steps per batch round, 1..512 */
  int autotune;            /* batches choose their tiling by timing */
  int backfill;            /* 0: admission strictly in arrival order; k > 0: queued runs that fit
                              may pass a head that does not, until that head has been passed
                              over in k rounds (then only the head may enter) */
} fv_server_opts;
typedef struct fv_request_info {
  int state;               /* 0 queued, 1 running, 2 done, 3 failed, 4 result taken */
  long steps;              /* steps taken in batch rounds (the first step is extra) */
  double t_submit, t_admit, t_done, device_bytes;
} fv_request_info;
typedef struct fv_server_stats {
  int submitted, completed, failed, running, queued, peak_running;
  double device_bytes, peak_device_bytes;
  long rounds;
  double busy_ms;          /* summed wall time of the batch rounds */
  double clock;            /* current server time, s */
} fv_server_stats;
int fv_server_create(fv_server **s, const fv_server_opts *opts);
/* queue one run: configuration (its stoptime ends the run) and initial data
 * as for fv_set_pvar / fv_set_sound_speeds / fv_set_background_velocity
 * (bccsound, fcsound, w may be NULL); *id numbers requests from 0. */
int fv_server_submit(fv_server *s, const fv_config *cfg, const double *pvar, const double *bccsound,
                     const double *fcsound, const double *w, int *id);
int fv_server_step(fv_server *s, int *running, int *queued);
int fv_server_poll(const fv_server *s, int id, fv_request_info *out);
/* final conservative state (all cells) and scalars of a finished run; the
 * result is released afterwards */
int fv_server_result(fv_server *s, int id, double *cvar, fv_scalars *sc);
int fv_server_get_stats(const fv_server *s, fv_server_stats *out);
int fv_server_destroy(fv_server *s);
const char *fv_server_last_error(const fv_server *s);

#ifdef __cplusplus
}
#endif
#endif
