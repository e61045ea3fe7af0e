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
// except the y reconstruction offsets, which depend on j only.
enum RingId {
  R_DAXW, R_DAYS, R_DYDZDV, R_DZDXDV, R_DXDYDV, R_DLX, R_DLY, R_CXYX, R_CYXY, R_CZXZ, R_CZYZ,
  R_RDXW, R_RDXE, R_ACCX, R_ACCY, R_CS, R_CSW, R_CSE, R_CSS, R_CSN, R_W, R_COUNT
};
enum ColId { C_RDYS, C_RDYN, C_COUNT };

struct Geom {
  int NI, NJ, P;        // ghost-inclusive cells per row / column, row pitch (elements)
  long FS;              // field stride (elements)
  int inum, jnum;
  double invdx, invdy, gamma, pmin, rhomin;
  int limiter;
  double theta;
  int bc[4];
  int fargo, gravity, curvilinear;
  int tile_x;           // k_rhs tile width in cells: 64 or 32
  const double *ring;   // [R_COUNT][NI]
  const double *col;    // [C_COUNT][NJ]
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
  int halt, reason, step, nsteps, flags, pad;
  double err_hist[CTL_HIST], dt_hist[CTL_HIST];
};

// One right-hand-side evaluation (timedisc_base.f90:1073-1271) on the RK
// stage state S = Y - sum_q (a[q]*dt) * k[q] (timedisc_rkfehlberg.f90:277-302).
// With ynew set, the launch is the last stage of the step and also performs
// the final combination and the embedded error estimate
// (timedisc_rkfehlberg.f90:199-266, timedisc_base.f90:1012-1053) in place of
// storing the stage's rhs.
struct StageArgs {
  const double *Y;
  const double *k[MAXK];
  double a[MAXK];       // Butcher coefficients of this stage
  int nk;
  double dt;            // time step (host-controlled)
  Ctl *ctl;             // device-controlled: dt and halting come from here
  double *kout;         // rhs (interior cells); unused with ynew
  double *yout;         // transformed state incl. ghost cells, or null
  int add_bg, sub_bg;   // FARGO frame change at entry / before the fluxes
  int pmin_clamp;       // Euler CHECK_PMIN round trip (timedisc_base.f90:1150-1164)
  int want_cfl, want_check;
  double *part;         // per-block partials: [2*b] max 1/dt, [2*b+1] check flags
  // fused final combination (last stage only)
  double *ynew;
  double bh[MAXK], bl[MAXK];
  double tol_rel, tol_abs[4];
  double *part_c;       // per-block partials: [2*b] max error, [2*b+1] = 0
};

// FargoAdvectionY on the combined state (timedisc_base.f90:1935-2069):
// comoving-frame CenterBoundary for the y-ghost rows, then per ring the
// integer shift and the minmod-limited fractional transport.
struct RemapArgs {
  const double *Y;
  double *out;          // remapped interior cells
  double dt_step;       // this%dt (host-controlled)
  Ctl *ctl;             // device-controlled: dt_step = (time+dt)-time, halting
};

enum Flag { F_INVALID = 1, F_RHOMIN = 2, F_PMIN = 4 };

// kernel launchers (fv_kernels.hip)
int rhs_blocks(const Geom &g);
void launch_rhs(const Geom &g, const StageArgs &a, int V, hipStream_t s);
void launch_remap(const Geom &g, const RemapArgs &a, int V, hipStream_t s);
void launch_frame(const Geom &g, const double *src, double *dst, int sign, int V, hipStream_t s);
void launch_reduce(const double *part, int n, double *out, hipStream_t s);
void launch_control(Ctl *ctl, const double *part_c, const double *part_r, int nb, hipStream_t s);

}  // namespace fv
