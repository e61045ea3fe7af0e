// fv_device.h -- data shared by the host control code and the HIP kernels of
// the finite-volume step.  Plain structs passed by value as kernel arguments.
#pragma once
#include <hip/hip_runtime.h>

namespace fv {

constexpr int GH = 2;       // ghost layers per side (mesh/mesh_base.f90:462)
constexpr int MAXK = 7;     // RK stages (Dormand-Prince has 7)
constexpr int TILE_X = 64;  // tile width in cells (one wavefront per tile row)
constexpr int CTL_HIST = 512;  // steps per device-controlled batch

enum Physics { EULER_ISOTHERM = 16, EULER = 17 };
enum Geometry { CARTESIAN = 1, LOGCYLINDRICAL = 31 };
enum Bc { NO_GRADIENTS = 1, PERIODIC = 2, ABSORBING = 13 };
enum Limiter { MINMOD = 1, MONOCENT = 2, SWEBY = 3, SUPERBEE = 4, OSPRE = 5, VANLEER = 7, NOLIMIT = 8 };

// Per-ring (x index i, ghost-inclusive) metric and physics arrays.  On the
// supported geometries every metric quantity depends on i only
// (mesh/mesh_midpoint.f90:97-145 with geometry_logcylindrical.f90:310-319),
// except the y reconstruction offsets and the y component of the rotating
// frame's centrifugal acceleration, which depend on j only.
enum RingId {
  R_DAXW, R_DAYS, R_DYDZDV, R_DZDXDV, R_DXDYDV, R_DLX, R_DLY, R_CXYX, R_CYXY, R_CZXZ, R_CZYZ,
  R_RDXW, R_RDXE, R_ACCX, R_ACCY, R_CS, R_CSW, R_CSE, R_CSS, R_CSN, R_W, R_WFAC, R_CACX,
  R_DAXE, R_DAYN, R_COUNT
};
enum ColId { C_RDYS, C_RDYN, C_CACY, C_COUNT };

struct Geom {
  int NI, NJ, P;        // ghost-inclusive cells per row / column, row pitch (elements)
  long FS;              // field stride (elements)
  int inum, jnum;
  double invdx, invdy, gamma, pmin, rhomin;
  int limiter;
  double theta;
  int recon_prim;       // 1: PRIMITIVE reconstruction (reconstruction_base.f90:117, fluxes_base.f90:378-381)
  int bc[4];
  double bdydz, bdzdx;  // Mesh%dy*Mesh%dz, Mesh%dz*Mesh%dx: boundary-flux weights
  int fargo, gravity, curvilinear;
  int rotframe;         // rotating-frame source: 0 off, 1 centrifugal + Coriolis, 2 Coriolis only
  double twoomega;      // 2*Omega (sources_rotframe.f90:162); centrifugal terms in R_CACX, C_CACY
  int tile_x;           // k_rhs tile width in cells: 64 or 32
  int stages;           // RK stages of the tableau (picks the kernels' coefficient slots)
  int dtype;            // stored state: 0 FP64 (plain drop-in), 1 FP32 (precision form)
  const double *ring;   // [R_COUNT][NI]
  const double *col;    // [C_COUNT][NJ]
  const float *ringf;   // FP32 copies read by the FP32 form (dtype 1)
  const float *colf;
};

// Device-side step control (one per context).  In a device-controlled batch
// the kernels read the time step from here; a one-block control kernel at the
// end of every step reduces the per-block partials of the final combination
// (embedded error) and of the end-of-step RHS (CFL, data checks), performs the
// accept test and computes the next time step.  Every step that the device
// cannot decide alone (rejection, break, output, stop time) halts the batch
// and the host redoes that step itself.
enum HaltReason { H_NONE = 0, H_ERR = 1, H_DTLIMIT = 2, H_FLAGS = 3, H_OUTPUT = 4, H_STOP = 5, H_BATCH = 6 };
struct Ctl {
  double time, dt, dt_step, t_out, stoptime, cfl, dtlimit, maxerrold, beta, order;
  double invdt;
  double cvis, invdt_v;  // viscosity (vis = 1): its Courant number and the last step's viscous 1/dt
  int halt, reason, step, nsteps, flags, vis;
  double err_hist[CTL_HIST], dt_hist[CTL_HIST];
};

// Boundary-flux accumulation of one step (timedisc_rkfehlberg.f90:216-259):
// out = in - sum_q dt*b_high(q)*k_q over the boundary rows, k_1..k_{s-1} from
// the ghost positions of the stage rhs arrays (element type by dtype), k_s
// from ks.  in/out/ks use the boundary-flux layout of StageArgsT::bks.
struct BfluxArgs {
  const void *k[MAXK];
  const double *ks, *in;
  double *out;
  double bh[MAXK];
  double dt;            // host-controlled step; device rounds take Ctl::dt
  int nk, V, NI, NJ, P, inum, jnum, dtype;
  long FS;
  Ctl *ctl;             // a device-controlled k_bflux launch: dt and halting from here
};

// Point-mass accretion (gravity_pointmass.f90:79-85, 383-428).  A state
// buffer holds mass, accrate, massloss, mdot, then the acceleration rings
// x and y ([2][NI], sources_gravity's 0 + accel_pointmass); the context keeps
// two and every update reads one and writes the other.
enum PointMassSlot { PM_MASS, PM_ACCRATE, PM_MASSLOSS, PM_MDOT, PM_HDR };
struct AccreteArgs {
  const double *gin;
  double *gout;
  float *gfout;         // FP32 copies of the output rings (dtype 1), or null
  const double *bf;     // density row of the outflow boundary's boundary fluxes
  int lo, hi;           // its interior range (GetBoundaryFlux sums lo..hi in order)
  int NI;
  double acclimit;
  double dt;            // ComputeRHS's dt argument (host-controlled)
  int dt_ctl;           // device rounds: dt = Ctl::dt
  int zero_accrate;     // device rounds: the stage evaluations before (dmass = 0, dt > 0) left accrate = 0
  Ctl *ctl;             // device rounds: halting
};

// One right-hand-side evaluation (timedisc_base.f90:1073-1271) on the RK
// stage state S = Y - sum_q (a[q]*dt) * k[q] (timedisc_rkfehlberg.f90:277-302).
// With ynew set, the launch is the last stage of the step and also performs
// the final combination and the embedded error estimate
// (timedisc_rkfehlberg.f90:199-266, timedisc_base.f90:1012-1053) in place of
// storing the stage's rhs.
template <typename T>
struct StageArgsT {
  const T *Y;
  const T *k[MAXK];
  double a[MAXK];       // Butcher coefficients of this stage
  int nk;
  double dt;            // time step (host-controlled)
  Ctl *ctl;             // device-controlled: dt and halting come from here
  T *kout;              // rhs (interior cells); unused with ynew
  T *yout;              // transformed state incl. ghost cells, or null
  int add_bg, sub_bg;   // FARGO frame change at entry / before the fluxes
  int pmin_clamp;       // Euler CHECK_PMIN round trip (timedisc_base.f90:1150-1164)
  int want_cfl, want_check;
  double *part;         // per-block partials: [2*b] max 1/dt, [2*b+1] check flags
  // fused final combination (last stage only)
  T *ynew;
  double bh[MAXK], bl[MAXK];
  double tol_rel, tol_abs[4];
  double *part_c;       // per-block partials: [2*b] max error, [2*b+1] = 0
  // boundary fluxes (timedisc_base.f90:1224-1252): with bflux set, a stage
  // stores the boundary rows of its rhs at the ghost positions next to the
  // boundary cells of kout; the last stage (which stores no rhs) writes them
  // to bks instead, in the boundary-flux layout: x part [V][2 sides][NJ],
  // then y part [V][2 sides][NI]
  int bflux;
  double *bks;
  // device-controlled end-of-step evaluation: the step's boundary-flux
  // accumulation (bfa.out set), by the launch's first threads after their tiles
  BfluxArgs bfa;
  // FARGO type 1: the azimuthal velocity of the real frame as the next
  // CalcBackgroundVelocity sees it, (v - w) + w (timedisc_base.f90:2243-2326),
  // for interior rows of every ring (ghost rings included); [NJ][P], or null
  double *vout;
  // end-of-step evaluation: Timedisc%pvar as ComputeRHS leaves it (after the
  // PMIN clamp and the frame change; ghost cells from CenterBoundary), all
  // cells, or null
  T *pout;
  // gravitational acceleration of this evaluation, rings x then y ([2][NI]):
  // the static rings, or the current point-mass state with accretion
  const double *acc;
  const float *accf;    // FP32 copies (dtype 1)
  // viscous source of this evaluation (fields 1..V-1, [V-1][NJ][P], interior
  // cells; k_visc writes it before the launch), or null
  const double *vsrc;
};
using StageArgs = StageArgsT<double>;

// Viscosity source (sources_viscosity.f90:73-577), FP64 state only.
enum VisModel { VIS_MOLECULAR = 1, VIS_ALPHA = 2, VIS_BETA = 3 };
struct VisArgs {
  int model;
  int update;           // ALPHA: this evaluation recomputes the dynamic viscosity
                        // (UpdateViscosity's time test, sources_viscosity.f90:371),
                        // else it reads the one stored by the last update
  double dynconst, bulkconst, omega;
  const double *ephir;  // ALPHA, BETA: e_phi/r in curvilinear components, then radius^2, [3][NJ][P]
  double *dyn;          // ALPHA, BETA: dynamic viscosity of the last update, all cells, [NJ][P]
  double *src;          // k_visc output: StageArgs::vsrc
  // ALPHA with Euler physics: Physics%bccsound, the sound speed of
  // Timedisc%pvar that every ComputeRHS of the main state updates
  // (timedisc_base.f90:1166-1174); a main-state evaluation (nk = 0) writes it
  // to cs_out from its own pvar, a stage evaluation reads it from cs_in
  const double *cs_in;
  double *cs_out;
  // device-controlled steps (StageArgs::ctl set): the evaluation's time is
  // ctl->time + c_cur*ctl->dt and the previous one's ctl->time + c_prev*ctl->dt,
  // and UpdateViscosity's time test is decided there (update is unused)
  double c_cur, c_prev;
  // end-of-step evaluation of a device-controlled step: per-workgroup maxima
  // of the viscous 1/dt (as k_visc_cfl on the pvar the evaluation leaves),
  // [2*b], or null
  double *part;
};

// FargoAdvectionY on the combined state (timedisc_base.f90:1935-2069):
// comoving-frame CenterBoundary for the y-ghost rows, then per ring the
// integer shift and the minmod-limited fractional transport.
template <typename T>
struct RemapArgsT {
  const T *Y;
  T *out;               // remapped interior cells
  double dt_step;       // this%dt (host-controlled)
  Ctl *ctl;             // device-controlled: dt_step = (time+dt)-time, halting
};
using RemapArgs = RemapArgsT<double>;

enum Flag { F_INVALID = 1, F_RHOMIN = 2, F_PMIN = 4 };

// Batches of independent meshes: one launch per kernel of the step covers
// every mesh.  Each launch reads its per-mesh arguments from a table in device
// memory and maps block b to a tile through a tile list of 32-bit codes
// (mesh index, tile column, tile row).
struct RhsEntry {
  Geom g;
  StageArgs a;
};
struct RemapEntry {
  Geom g;
  RemapArgs a;
};
struct ControlEntry {
  Ctl *ctl;
  const double *part_c, *part_r;
  int nb;
};
constexpr int TILE_DISK_BITS = 9, TILE_BX_BITS = 11, TILE_BY_BITS = 12;
constexpr int MAX_BATCH = 1 << TILE_DISK_BITS;
constexpr unsigned NO_TILE = 0xFFFFFFFFu;  // empty slot of a tile list
__host__ __device__ inline unsigned tile_code(int disk, int bx, int by) {
  return (unsigned)disk | ((unsigned)bx << TILE_DISK_BITS) | ((unsigned)by << (TILE_DISK_BITS + TILE_BX_BITS));
}
__host__ __device__ inline int tile_disk(unsigned e) { return (int)(e & ((1u << TILE_DISK_BITS) - 1)); }
__host__ __device__ inline int tile_bx(unsigned e) { return (int)((e >> TILE_DISK_BITS) & ((1u << TILE_BX_BITS) - 1)); }
__host__ __device__ inline int tile_by(unsigned e) { return (int)(e >> (TILE_DISK_BITS + TILE_BX_BITS)); }

// kernel launchers (fv_kernels.hip).  The host passes state buffers through
// the FP64-typed argument structs; for g.dtype == 1 they hold FP32 arrays.
int rhs_blocks(const Geom &g);
void launch_rhs(const Geom &g, const StageArgs &a, int V, hipStream_t s);
void launch_remap(const Geom &g, const RemapArgs &a, int V, hipStream_t s);
// (w2: the w row, of the state's element type, of a second frame change
// applied in the same pass)
void launch_frame(const Geom &g, const void *src, void *dst, int sign, int V, hipStream_t s,
                  const Ctl *ctl = nullptr, const void *w2 = nullptr);
void launch_reduce(const double *part, int n, double *out, hipStream_t s);
// CalcTimestep partials (max per block) from a stored pvar, v + w when add_w
void launch_cfl_pvar(const Geom &g, const void *pv, int add_w, double *part, int nblocks, int V, hipStream_t s);
// viscous source of the evaluation a into v.src (FP64)
void launch_visc(const Geom &g, const StageArgs &a, const VisArgs &v, int V, hipStream_t s);
// viscous time-step limit (partials, max per block) from a stored FP64 pvar
// (v + w when add_w); v.update: the dynamic viscosity from that pvar
void launch_visc_cfl(const Geom &g, const VisArgs &v, const double *pv, int add_w, double *part, int nblocks, int V,
                     hipStream_t s);
// part_v/nbv: the viscous 1/dt partials of the end-of-step k_visc (or null)
void launch_control(Ctl *ctl, const double *part_c, const double *part_r, int nb, hipStream_t s,
                    const double *part_v = nullptr, int nbv = 0);
// k_visc workgroups of a mesh (tiles of 64 x 8 cells over the ghost-inclusive mesh)
int visc_blocks(const Geom &g);
void launch_bflux(const BfluxArgs &bf, hipStream_t s);
void launch_accrete(const AccreteArgs &p, hipStream_t s);
// FARGO type 1: w(i) = SUM_j v(i,j)*wfactor(i) / JNUM per ring into the ring
// arrays (FP64 and, for the precision form, FP32)
void launch_wsum(const Geom &g, const double *v, double *ring, float *ringf, hipStream_t s,
                 const Ctl *ctl = nullptr);
// batch launchers: ntiles blocks over the tile list; TX = tile width of the list
#ifndef FV_RHS_ROWS
#define FV_RHS_ROWS 8
#endif
constexpr int TY_TILE = FV_RHS_ROWS;  // k_rhs tile rows
// FARGO remap: rows per thread; a 64 x 4 block covers 64 rings x 4*REMAP_ROWS rows
#ifndef FV_REMAP_ROWS
#define FV_REMAP_ROWS 2
#endif
constexpr int REMAP_ROWS = FV_REMAP_ROWS;
__host__ __device__ inline int remap_rows_blocks(int jnum) { return (jnum + 4 * REMAP_ROWS - 1) / (4 * REMAP_ROWS); }
// mode: 0 RK stage rhs, 1 last stage with the final combination, 2 end-of-step evaluation
void launch_rhs_multi(const RhsEntry *tab, const unsigned *tiles, int ntiles, int V, int dtype, int limiter,
                      int tx, int stages, int mode, hipStream_t s);
void launch_remap_multi(const RemapEntry *tab, const unsigned *tiles, int ntiles, int V, int dtype, hipStream_t s);
void launch_control_multi(const ControlEntry *tab, int n, hipStream_t s);
// batches with point-mass accretion: the step's boundary-flux accumulation
// (bpm workgroups per member) and UpdateGravity (one workgroup per member) of
// the n accreting members in tab
void launch_bflux_multi(const BfluxArgs *tab, int n, int bpm, hipStream_t s);
void launch_accrete_multi(const AccreteArgs *tab, int n, hipStream_t s);

// Reduction scratch of the allocation-free step (fv_step_noalloc): a static
// device array, one copy per device, laid out as the reduction result (2
// values), then the per-workgroup partials of the end-of-step evaluation and
// of the last stage (2 values per workgroup each).  At most
// SCRATCH_MAX_BLOCKS workgroups per launch.
constexpr int SCRATCH_MAX_BLOCKS = 16384;
constexpr int SCRATCH_RED = 0, SCRATCH_PART_R = 32, SCRATCH_PART_C = SCRATCH_PART_R + 2 * SCRATCH_MAX_BLOCKS;
constexpr int SCRATCH_DOUBLES = SCRATCH_PART_C + 2 * SCRATCH_MAX_BLOCKS;
// device address of the scratch array on device dev (looked up once per
// device, with that device current); hipSuccess or the lookup's error
hipError_t scratch_address(int dev, double **out);

}  // namespace fv
