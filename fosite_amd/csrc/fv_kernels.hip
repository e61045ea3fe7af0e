// fv_kernels.hip -- CDNA4 (gfx950) kernels of fosite's unsplit finite-volume
// step, FP64.
//
// Every floating-point expression keeps the reference's operation order and
// the file is compiled with -ffp-contract=off, so IEEE +,-,*,/ and sqrt give
// the reference's bits.  Citations are file:line in the fosite source tree.
//
// k_rhs (one launch per right-hand-side evaluation) is fully fused:
//   * the input state of every loaded cell is formed on the fly: the RK stage
//     state Y - sum a*dt*k (timedisc_rkfehlberg.f90:277-302);
//   * ghost cells are evaluated on the fly from interior cells with the
//     boundary rules (boundary_generic.f90:287-517, boundary_nogradients.f90,
//     boundary_periodic.f90, boundary_absorbing.f90), so no boundary pass and
//     no intermediate state array exist;
//   * FARGO frame round trip, Euler PMIN round trip, linear reconstruction,
//     face conversions, wave speeds, physical and KT fluxes, geometric,
//     gravity and FARGO sources, rhs assembly, CFL and data checks follow.
//   * in the last stage of a step, the final RK combination and the embedded
//     error norm (pointwise) replace the store of the stage's rhs.
//   Tiles of 64 x BY threads (one cell each) with the halo staged in LDS.
// k_remap: FARGO azimuthal shift of the combined state (y direction only).
// k_control: one block; in device-controlled batches it reduces the partials
// of the last stage and of the end-of-step k_rhs, runs the accept test and
// the CFL update.  Kernel boundaries order everything: no grid-wide fences.
#include <float.h>
#include <math.h>

#include "fv_device.h"

namespace fv {

#define TINYD DBL_MIN

__device__ __forceinline__ double fmin2(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double fmax2(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }
__device__ __forceinline__ bool bad(double x) { return (x != x) || (x > DBL_MAX); }

// per-ring arrays: uniform base per id, 32-bit unsigned lane offset
__device__ __forceinline__ double ring(const Geom &g, int id, int i) { return (g.ring + id * g.NI)[(unsigned)i]; }
__device__ __forceinline__ long cidx(const Geom &g, int i, int j) { return (long)j * g.P + i; }

// ---------------------------------------------------------------- physics
// Cons2Prim2d / Prim2Cons2d (physics_eulerisotherm.f90:2835-2881,
// physics_euler.f90:3056-3104)
template <int V>
__device__ __forceinline__ void c2p(const double *u, double *p, double gamma) {
  double inv = 1. / u[0];
  p[0] = u[0];
  p[1] = u[1] * inv;
  p[2] = u[2] * inv;
  if (V == 4) p[3] = (gamma - 1.) * (u[3] - 0.5 * inv * (u[1] * u[1] + u[2] * u[2]));
}
template <int V>
__device__ __forceinline__ void p2c(const double *p, double *u, double gamma) {
  u[0] = p[0];
  u[1] = p[0] * p[1];
  u[2] = p[0] * p[2];
  if (V == 4) u[3] = p[3] / (gamma - 1.) + 0.5 * p[0] * (p[1] * p[1] + p[2] * p[2]);
}
// physics_euler.f90:2682-2689
__device__ __forceinline__ double sound_speed(double gamma, double rho, double pr) {
  return sqrt(fmax2(2.0 * TINYD, gamma * pr / rho));
}

// ------------------------------------------------------------ input state
template <int V>
struct In {
  const Geom &g;
  const StageArgs &a;
  double dt;
  double c[MAXK];  // a(m,q)*dt
  int nk;

  __forceinline__ __device__ In(const Geom &g_, const StageArgs &a_) : g(g_), a(a_) {
    dt = a.ctl ? a.ctl->dt : a.dt;  // device-controlled: this step's dt
    nk = a.nk;
#pragma unroll
    for (int q = 0; q < MAXK; q++) c[q] = a.a[q] * dt;
  }

  // input state at an interior cell
  __forceinline__ __device__ void interior(int i, int j, double *s) const {
    {
      const long ix = cidx(g, i, j);
      // all loads first (uniform branches only), then the stage sum in the
      // reference's order
      double kv[MAXK][V];
#pragma unroll
      for (int l = 0; l < V; l++) s[l] = a.Y[l * g.FS + ix];
#pragma unroll
      for (int q = 0; q < MAXK; q++)
        if (q < nk) {
#pragma unroll
          for (int l = 0; l < V; l++) kv[q][l] = a.k[q][l * g.FS + ix];
        }
#pragma unroll
      for (int q = 0; q < MAXK; q++)
        if (q < nk) {
#pragma unroll
          for (int l = 0; l < V; l++) s[l] = s[l] - c[q] * kv[q][l];
        }
    }
  }

  // stage states of two interior cells (p0, p1 say which are wanted): all
  // loads of both cells are issued before any arithmetic
  __forceinline__ __device__ void interior2(bool p0, int ia, int ja, double *s0, bool p1, int ib, int jb,
                                            double *s1) const {
    const unsigned x0 = (unsigned)ja * (unsigned)g.P + (unsigned)ia;
    const unsigned x1 = (unsigned)jb * (unsigned)g.P + (unsigned)ib;
    double k0[MAXK][V], k1[MAXK][V];
    if (p0) {
#pragma unroll
      for (int l = 0; l < V; l++) s0[l] = (a.Y + l * g.FS)[x0];
#pragma unroll
      for (int q = 0; q < MAXK; q++)
        if (q < nk) {
#pragma unroll
          for (int l = 0; l < V; l++) k0[q][l] = (a.k[q] + l * g.FS)[x0];
        }
    }
    if (p1) {
#pragma unroll
      for (int l = 0; l < V; l++) s1[l] = (a.Y + l * g.FS)[x1];
#pragma unroll
      for (int q = 0; q < MAXK; q++)
        if (q < nk) {
#pragma unroll
          for (int l = 0; l < V; l++) k1[q][l] = (a.k[q] + l * g.FS)[x1];
        }
    }
#pragma unroll
    for (int q = 0; q < MAXK; q++)
      if (q < nk) {
#pragma unroll
        for (int l = 0; l < V; l++) {
          s0[l] = s0[l] - c[q] * k0[q][l];
          s1[l] = s1[l] - c[q] * k1[q][l];
        }
      }
  }

  // ComputeRHS transform (timedisc_base.f90:1098-1208): ur = cvar after the
  // frame change, CenterBoundary and PMIN (real frame), pr = pvar at that
  // point, uc = cvar handed to the fluxes
  __forceinline__ __device__ void transform(int i, const double *s, double *ur, double *pr, double *uc) const {
#pragma unroll
    for (int l = 0; l < V; l++) ur[l] = s[l];
    if (a.add_bg) ur[2] = ur[2] + ur[0] * ring(g, R_W, i);
    c2p<V>(ur, pr, g.gamma);
    if (V == 4 && a.pmin_clamp) {
      pr[3] = fmax2(pr[3], g.pmin);
      p2c<V>(pr, ur, g.gamma);
    }
#pragma unroll
    for (int l = 0; l < V; l++) uc[l] = ur[l];
    if (a.sub_bg) uc[2] = ur[2] - ur[0] * ring(g, R_W, i);
  }

  // pvar of an interior cell right after CenterBoundary's conversion
  __forceinline__ __device__ void prim_interior(int i, int j, double *p) const {
    double s[V];
    interior(i, j, s);
    if (a.add_bg) s[2] = s[2] + s[0] * ring(g, R_W, i);
    c2p<V>(s, p, g.gamma);
  }

  // absorbing boundary, isothermal physics, x direction:
  // boundary_absorbing.f90:114-150, physics_eulerisotherm.f90:1764-1849,
  // 2027-2085, 2895-2969
  __forceinline__ __device__ void absorbing_x(int east, int j, double *p) const {
    const int imin = GH, imax = GH + g.inum - 1;
    const int i1 = east ? imax : imin, i2 = east ? imax - 1 : imin + 1;
    const int iL = i1 < i2 ? i1 : i2, iR = i1 < i2 ? i2 : i1;
    const int delta = east ? 1 : -1;
    double pL[V], pR[V];
    prim_interior(iL, j, pL);
    prim_interior(iR, j, pR);
    double p1[V];  // element-wise select keeps both states in registers
#pragma unroll
    for (int l = 0; l < V; l++) p1[l] = east ? pR[l] : pL[l];
    const double cs = ring(g, R_CS, i1), u1 = p1[1];
    const double lam0 = u1 - cs, lam1 = u1, lam2 = u1 + cs;
    const double csf = ring(g, R_CSW, iR);
    const double dlnrho = log(pR[0] / pL[0]);
    const double ducs = (pR[1] - pL[1]) / csf;
    double x0 = dlnrho - ducs, x1 = pR[2] - pL[2], x2 = dlnrho + ducs;
    if (east) {
      if (lam0 <= 0.0) x0 = 0.0;
      if (lam1 <= 0.0) x1 = 0.0;
      if (lam2 <= 0.0) x2 = 0.0;
    } else {
      if (lam0 >= 0.0) x0 = 0.0;
      if (lam1 >= 0.0) x1 = 0.0;
      if (lam2 >= 0.0) x2 = 0.0;
    }
    const double cs2 = ring(g, east ? R_CSE : R_CSW, i1);
    p[0] = p1[0] * exp(delta * 0.5 * (x2 + x0));
    p[1] = u1 + delta * 0.5 * cs2 * (x2 - x0);
    p[2] = p1[2] + delta * x1;
  }

  // x-ghost (side, layer 1..2) at interior row j
  __forceinline__ __device__ void ghost_we(int east, int gl, int j, double *p) const {
    const int imin = GH, imax = GH + g.inum - 1, bc = east ? g.bc[1] : g.bc[0];
    if (bc == NO_GRADIENTS)
      prim_interior(east ? imax : imin, j, p);
    else if (bc == PERIODIC)
      prim_interior(east ? imin + gl - 1 : imax - gl + 1, j, p);
    else
      absorbing_x(east, j, p);
  }
  // y-ghost (side, layer) at interior column i
  __forceinline__ __device__ void ghost_sn(int north, int gl, int i, double *p) const {
    const int jmin = GH, jmax = GH + g.jnum - 1, bc = north ? g.bc[3] : g.bc[2];
    if (bc == NO_GRADIENTS)
      prim_interior(i, north ? jmax : jmin, p);
    else
      prim_interior(i, north ? jmin + gl - 1 : jmax - gl + 1, p);
  }

  // pvar of any ghost cell after CenterBoundary: boundary rules in the order
  // W, E, S, N, then SetCornerEdges (boundary_generic.f90:434-517)
  __forceinline__ __device__ void ghost_prim(int i, int j, double *p) const {
    const int imin = GH, imax = GH + g.inum - 1, jmin = GH, jmax = GH + g.jnum - 1;
    const bool xg = i < imin || i > imax, yg = j < jmin || j > jmax;
    if (xg && !yg) {
      const int east = i > imax;
      ghost_we(east, east ? i - imax : imin - i, j, p);
    } else if (yg && !xg) {
      const int north = j > jmax;
      ghost_sn(north, north ? j - jmax : jmin - j, i, p);
    } else {
      const int east = i > imax, north = j > jmax;
      const int ga = east ? i - imax : imin - i, gb = north ? j - jmax : jmin - j;
      const int ir = east ? imax : imin, jr = north ? jmax : jmin;
      if (ga == 2 && gb == 1) {
        ghost_we(east, 2, jr, p);
      } else if (ga == 1 && gb == 2) {
        ghost_sn(north, 2, ir, p);
      } else {
        double px[V], py[V];
        ghost_we(east, ga, jr, px);
        ghost_sn(north, gb, ir, py);
#pragma unroll
        for (int l = 0; l < V; l++) p[l] = 0.5 * (px[l] + py[l]);
      }
    }
  }

  // final ghost state: PMIN round trip, prim->cons, data check, frame change
  __forceinline__ __device__ void ghost_final(int i, int j, double *u, double *p, int &flags) const {
    ghost_prim(i, j, p);
    if (V == 4 && a.pmin_clamp) p[3] = fmax2(p[3], g.pmin);
    p2c<V>(p, u, g.gamma);
    if (a.want_check) {
#pragma unroll
      for (int l = 0; l < V; l++)
        if (bad(u[l]) || bad(p[l])) flags |= F_INVALID;
    }
    if (a.sub_bg) {
      const double w = ring(g, R_W, i);
      p[2] = p[2] - w;
      u[2] = u[2] - u[0] * w;
    }
  }
};

// CFL contribution of one cell (timedisc_base.f90:869-872 with
// physics_eulerisotherm.f90:530-552); u, v = pvar after the frame change back
__device__ __forceinline__ double cell_invdt(const Geom &g, int i, double u, double v, double cs) {
  const double mnx = fmin2(0., u - cs), mxx = fmax2(0., u + cs);
  const double mny = fmin2(0., v - cs), mxy = fmax2(0., v + cs);
  return fmax2(mxx, -mnx) / ring(g, R_DLX, i) + fmax2(mxy, -mny) / ring(g, R_DLY, i);
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax2(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_or(int v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}
// block-wide max / or; every thread gets the result
__device__ double blk_max(double v, double *red) {
  v = wave_max(v);
  const int t = threadIdx.x + threadIdx.y * blockDim.x, nw = (blockDim.x * blockDim.y + 63) / 64;
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double r = -DBL_MAX;
  for (int k = 0; k < nw; k++) r = fmax2(r, red[k]);
  return r;
}
__device__ int blk_or(int v, int *red) {
  v = wave_or(v);
  const int t = threadIdx.x + threadIdx.y * blockDim.x, nw = (blockDim.x * blockDim.y + 63) / 64;
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  int r = 0;
  for (int k = 0; k < nw; k++) r |= red[k];
  return r;
}
// ----------------------------------------------------------- reconstruction
// limiters, reconstruction_linear.f90:645-733
// LIM = 0: limiter chosen at run time from g.limiter
template <int LIM>
__device__ __forceinline__ double slope(const Geom &g, double um, double u0, double up) {
  const double a1 = u0 - um, a2 = up - u0;
  switch (LIM ? LIM : g.limiter) {
    case MINMOD:
      return (fsign(1.0, a1) * fsign(1.0, a2) > 0) ? fsign(fmin2(fabs(a1), fabs(a2)), a1) : 0.;
    case MONOCENT: {
      const double b1 = g.theta * (u0 - um), b2 = g.theta * (up - u0), b3 = 0.5 * (up - um);
      if ((fsign(1.0, b1) * fsign(1.0, b2) > 0) && (fsign(1.0, b2) * fsign(1.0, b3) > 0))
        return fsign(fmin2(fmin2(fabs(b1), fabs(b2)), fabs(b3)), b1);
      return 0.;
    }
    case SWEBY:
    case SUPERBEE: {
      const double q = (LIM ? LIM : g.limiter) == SUPERBEE ? 2.0 : g.theta;
      return (fsign(1.0, a1) * fsign(1.0, a2) > 0)
                 ? fsign(fmax2(fmin2(q * fabs(a1), fabs(a2)), fmin2(fabs(a1), q * fabs(a2))), a1)
                 : 0.;
    }
    case OSPRE:
      return 1.5 * a1 * a2 * (a1 + a2) / (a1 * (a1 + 0.5 * a2) + a2 * (a2 + 0.5 * a1) + TINYD);
    case VANLEER: {
      const double b1 = fabs(a1), b2 = fabs(a2);
      return (a1 * b2 + a2 * b1) / (b1 + b2 + TINYD);
    }
    default:
      return 0.5 * (a1 + a2);
  }
}

// KT flux across one interface from the two face states (fluxes_kt.f90:126-176,
// physics_eulerisotherm.f90:555-686, 835-930, 2742-2817; physics_euler.f90:492-600)
// d = 0: x-interface (normal momentum index 1), d = 1: y-interface (index 2).
template <int V>
__device__ __forceinline__ void kt_flux(const Geom &g, int d, const double *UL, const double *UR,
                                        double csL, double csR, double dA, double *F) {
  const int un = 1 + d, ut = 2 - d;
  double pl[V], pr[V];
  c2p<V>(UL, pl, g.gamma);
  c2p<V>(UR, pr, g.gamma);
  if (V == 4) {
    csL = sound_speed(g.gamma, pl[0], pl[3]);
    csR = sound_speed(g.gamma, pr[0], pr[3]);
  }
  const double vL = pl[un], vR = pr[un];
  const double amin = fmin2(fmin2(0.0, fmin2(0., vR - csR)), fmin2(0., vL - csL));
  const double amax = fmax2(fmax2(0.0, fmax2(0., vR + csR)), fmax2(0., vL + csL));
  double fL[V], fR[V];
  fL[0] = pl[0] * vL;
  fR[0] = pr[0] * vR;
  if (V == 3) {
    fL[un] = UL[un] * vL + pl[0] * csL * csL;
    fR[un] = UR[un] * vR + pr[0] * csR * csR;
  } else {
    fL[un] = UL[un] * vL + pl[3];
    fR[un] = UR[un] * vR + pr[3];
    fL[3] = (UL[3] + pl[3]) * vL;
    fR[3] = (UR[3] + pr[3]) * vR;
  }
  fL[ut] = UL[ut] * vL;
  fR[ut] = UR[ut] * vR;
  const double fac = dA / (amax - amin);
#pragma unroll
  for (int l = 0; l < V; l++) F[l] = fac * (amax * fL[l] - amin * fR[l] + amin * amax * (UR[l] - UL[l]));
}

// ------------------------------------------------------------ step control
// PI controller (timedisc_base.f90:759-814), used on the device only to stop
// a batch when the step could fall under dtlimit; the host redoes such steps
__device__ double pi_dtnew(const Ctl *c, double maxerr, double dtold) {
  double dttmp = maxerr > 0. ? 0.9 * dtold * exp(-log(maxerr) / c->order) : 4. * dtold;
  double dtnew = c->maxerrold > 0. ? dttmp * exp(-c->beta * log(c->maxerrold / maxerr)) : dttmp;
  if (maxerr < 1.0 && 4. * dtold < dtnew) dtnew = 4. * dtold;
  return dtnew;
}
__device__ void halt(Ctl *c, int reason) {
  c->halt = 1;
  c->reason = reason;
}
// after SolveODE: accept when the embedded error is finite and < 1
// (timedisc_base.f90:686-713, AcceptSolution 911-959)
__device__ void control_accept(Ctl *c, double err) {
  const int s = c->step;
  c->err_hist[s] = err;
  c->dt_hist[s] = c->dt;
  if (!(err < 1.0)) return halt(c, H_ERR);  // NaN, Inf or a rejection
  if (pi_dtnew(c, err, c->dt) < c->dtlimit * (1.0 + 1.0e-6)) return halt(c, H_DTLIMIT);
  c->maxerrold = err;
  const double t1 = c->time + c->dt;
  c->dt_step = t1 - c->time;
  c->time = t1;
}
// after the end-of-step RHS: data checks, output test, next CFL step,
// stop-time test and output-time clipping (fosite.f90:383-462,
// io/fileio_base.f90:604-620)
__device__ void control_next(Ctl *c, double invdt, int flags) {
  if (flags) return halt(c, H_FLAGS);
  const double t = c->time;
  if (fabs(c->t_out - t) <= 1.0E-5 * c->t_out) return halt(c, H_OUTPUT);
  double dt = c->cfl / invdt;
  if (fabs(c->stoptime - t) <= 1.0E-05 * c->stoptime) return halt(c, H_STOP);
  if ((t + dt) / c->t_out > 1.0)
    dt = c->t_out - t;
  else if ((t + 1.5 * dt) / c->t_out > 1.0)
    dt = 0.5 * (c->t_out - t);
  c->invdt = invdt;
  c->dt = dt;
  c->step = c->step + 1;
  if (c->step >= c->nsteps) halt(c, H_BATCH);
}

// -------------------------------------------------------------- fused RHS
// Tile: TX x TY_RHS output cells (TX = 64 or 32), one thread each.  The footprint (2-cell
// halo, corners excluded) is staged in LDS as the state the fluxes see;
// threads with an extra job cover the halo cells and the faces and fluxes one
// cell beyond the tile.  Phases, separated by barriers:
//   1 stage state + transform of the own and one extra footprint cell (ghost
//     cells on boundary tiles), sources, data checks and CFL of the own cell
//   2 limited slopes and face states: own cell; rows -1 and TY; columns -1, TX
//   3 KT interface fluxes: own east and north interfaces; west of column 0,
//     south of row 0
//   4 rhs assembly, or the final RK combination in the last stage; stores
// Faces live in their own LDS region; the fluxes reuse the footprint region,
// which is dead after phase 2, so three barriers separate the four phases.
// Output-aligned tiles keep every lane of the flux phases busy and divide
// power-of-two meshes exactly; at most 128 VGPRs, so that 4 waves share
// each SIMD.
#ifndef FV_RHS_ROWS
#define FV_RHS_ROWS 8
#endif
#ifndef FV_RHS_WAVES_PER_SIMD
#define FV_RHS_WAVES_PER_SIMD 4
#endif
constexpr int TY_RHS = FV_RHS_ROWS;

template <int V, int LIM, int TX>
__global__ __launch_bounds__(TX *TY_RHS, FV_RHS_WAVES_PER_SIMD) void k_rhs(Geom g, StageArgs a) {
  constexpr int TY = TY_RHS, SX = TX + 4, SY = TY + 4;
  constexpr int N_UC = V * SY * SX;
  constexpr int XW = V * TY * (TX + 1), XE = V * TY, YS = V * (TY + 1) * TX, YN = V * TX;
  constexpr int N_FACE = XW + XE + YS + YN;
  constexpr int FX = V * TY * (TX + 1), FY = V * (TY + 1) * TX;
  constexpr int N_FLUX = FX + FY;
  // region A: footprint state (phases 1-2), then fluxes (phases 3-4);
  // region B: face states (phases 2-3)
  constexpr int NA = N_UC > N_FLUX ? N_UC : N_FLUX;
  __shared__ double lds[NA + N_FACE];
  __shared__ double redd[TY];
  __shared__ int redi[TY];
  // CFL term and check flags of the ghost cells a thread evaluated (kept in
  // LDS rather than registers across the interior path)
  __shared__ double gcfl[TX * TY];
  __shared__ int gflg[TX * TY];
  // footprint state, tile-relative row r in [-2, TY+1], column c in [-2, TX+1]
  auto UC = [&](int l, int r, int c) -> double & { return lds[(l * SY + r + 2) * SX + c + 2]; };
  // faces: W of columns 0..TX, E of column -1, S of rows 0..TY, N of row -1
  double *fb = lds + NA;
  auto FW = [&](int l, int r, int c) -> double & { return fb[(l * TY + r) * (TX + 1) + c]; };
  auto FE1 = [&](int l, int r) -> double & { return fb[XW + l * TY + r]; };
  auto FS = [&](int l, int r, int c) -> double & { return fb[XW + XE + (l * (TY + 1) + r) * TX + c]; };
  auto FN1 = [&](int l, int c) -> double & { return fb[XW + XE + YS + l * TX + c]; };
  // fluxes: x interface c-1/2 (c = 0..TX) of row r; y interface r-1/2 (r = 0..TY) of column c
  auto FLX = [&](int l, int r, int c) -> double & { return lds[(l * TY + r) * (TX + 1) + c]; };
  auto FLY = [&](int l, int r, int c) -> double & { return lds[FX + (l * (TY + 1) + r) * TX + c]; };

  if (a.ctl && a.ctl->halt) return;
  const In<V> in(g, a);
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + TX * ty;
  const int i0 = GH + blockIdx.x * TX, j0 = GH + blockIdx.y * TY;
  const int imin = GH, jmin = GH, imax = GH + g.inum - 1, jmax = GH + g.jnum - 1;
  const int i = i0 + tx, j = j0 + ty;
  const bool out = i <= imax && j <= jmax;  // own cell is an interior cell

  // ---- phase 1: footprint state into LDS
  // extra footprint cell: rows -2,-1,TY,TY+1 of the own column (t < 4*TX),
  // then columns -2,-1,TX,TX+1 of rows 0..TY-1 (next 4*TY threads)
  int er = 0, ec = 0;
  bool has_e = false;
  if (t < 4 * TX) {
    er = ty < 2 ? ty - 2 : TY + ty - 2;
    ec = tx;
    has_e = true;
  } else if (t < 4 * TX + 4 * TY) {
    const int k = t - 4 * TX, q = k / TY;
    ec = q < 2 ? q - 2 : TX + q - 2;
    er = k % TY;
    has_e = true;
  }
  const int ei = i0 + ec, ej = j0 + er;
  const bool e_in = has_e && ei >= imin && ei <= imax && ej >= jmin && ej <= jmax;
  // ghost cells or cells beyond the mesh (boundary tiles only) first, so that
  // no value of the interior path is live across this register-hungry code.
  // Every non-corner ghost cell of the mesh lies in some tile's footprint, so
  // the ghost part of the data checks and of the CFL reduction (ghosts
  // included, timedisc_base.f90:869-872) and the ghost cells of the new state
  // are produced here; a cell seen by two tiles contributes identical values.
  const bool own_g = !out, e_g = has_e && !e_in;
  if (own_g || e_g) {
    double ginv = -DBL_MAX;
    int gf = 0;
#pragma unroll
    for (int p = 0; p < 2; p++) {
      if (!(p ? e_g : own_g)) continue;
      const int ii = p ? ei : i, jj = p ? ej : j;
      const bool xin = ii >= imin && ii <= imax, yin = jj >= jmin && jj <= jmax;
      double u[V];
      if ((xin || yin) && ii >= 0 && jj >= 0 && ii < g.NI && jj < g.NJ) {
        double pp[V];
        in.ghost_final(ii, jj, u, pp, gf);
        if (a.want_cfl) {
          const double cs = (V == 4) ? sound_speed(g.gamma, pp[0], pp[3]) : ring(g, R_CS, ii);
          ginv = fmax2(ginv, cell_invdt(g, ii, pp[1], pp[2], cs));
        }
        if (a.yout) {
          const unsigned ix = (unsigned)jj * (unsigned)g.P + (unsigned)ii;
#pragma unroll
          for (int l = 0; l < V; l++) (a.yout + l * g.FS)[ix] = u[l];
        }
      } else {
#pragma unroll
        for (int l = 0; l < V; l++) u[l] = 0.0;  // corner or beyond the mesh: never needed
      }
#pragma unroll
      for (int l = 0; l < V; l++) UC(l, p ? er : ty, p ? ec : tx) = u[l];
    }
    gcfl[t] = ginv;
    gflg[t] = gf;
  }

  double ur[V], pr[V];
  {
    double so[V], se[V], uo[V], ue[V], re[V], pe[V];
    in.interior2(out, i, j, so, e_in, ei, ej, se);
    if (out) {
      in.transform(i, so, ur, pr, uo);
#pragma unroll
      for (int l = 0; l < V; l++) UC(l, ty, tx) = uo[l];
    }
    if (e_in) {
      in.transform(ei, se, re, pe, ue);
#pragma unroll
      for (int l = 0; l < V; l++) UC(l, er, ec) = ue[l];
    }
  }
  // sources, data checks and CFL of the own interior cell
  double geo[V], ext[V];
  double invdt = -DBL_MAX;
  int flags = 0;
#pragma unroll
  for (int l = 0; l < V; l++) geo[l] = ext[l] = 0.0;
  if (out) {
    // geometrical sources (physics_eulerisotherm.f90:689-832, elementals 3069-3088;
    // physics_euler.f90:1706-1846); zero on cartesian meshes
    if (g.curvilinear) {
      const double cxyx = ring(g, R_CXYX, i), cyxy = ring(g, R_CYXY, i);
      const double czxz = ring(g, R_CZXZ, i), czyz = ring(g, R_CZYZ, i);
      const double vx = pr[1], vy = pr[2], vz = 0.0, mx = ur[1], my = ur[2], mz = 0.0;
      double P;
      if (V == 4) {
        P = pr[3];
      } else {
        const double cs = ring(g, R_CS, i);
        P = pr[0] * (cs * cs);
      }
      geo[1] = my * (cyxy * vy - cxyx * vx) + mz * (czxz * vz - 0.0 * vx) + (cyxy + czxz) * P;
      geo[2] = mz * (czyz * vz - 0.0 * vy) + mx * (cxyx * vx - cyxy * vy) + (cxyx + czyz) * P;
    }
    // FARGO source (physics_eulerisotherm.f90:1531-1586)
    if (g.fargo)
      geo[2] = geo[2] - ur[1] * 0.5 * (ring(g, R_W, i + 1) - ring(g, R_W, i - 1)) / ring(g, R_DLX, i);
    // point-mass gravity through the source list (sources_generic.f90:220-261,
    // sources_gravity.f90:164-199, physics_eulerisotherm.f90:1201-1226,
    // physics_euler.f90:1849-1885)
    if (g.gravity) {
      const double ax = ring(g, R_ACCX, i), ay = ring(g, R_ACCY, i);
      ext[1] = 0.0 + ur[0] * ax;
      ext[2] = 0.0 + ur[0] * ay;
      if (V == 4) ext[3] = 0.0 + (ur[1] * ax + ur[2] * ay);
    }
    // CheckData (timedisc_base.f90:1666-1742) and CFL (timedisc_base.f90:818-909)
    if (a.want_check) {
      if (pr[0] < g.rhomin) flags |= F_RHOMIN;
      if (V == 4 && pr[3] < g.pmin) flags |= F_PMIN;
#pragma unroll
      for (int l = 0; l < V; l++)
        if (bad(pr[l]) || bad(ur[l])) flags |= F_INVALID;
    }
    if (a.want_cfl) {
      const double vy = a.sub_bg ? pr[2] - ring(g, R_W, i) : pr[2];
      const double cs = (V == 4) ? sound_speed(g.gamma, pr[0], pr[3]) : ring(g, R_CS, i);
      invdt = cell_invdt(g, i, pr[1], vy, cs);
    }
  }
  __syncthreads();

  // ---- phase 2: limited slopes and face states (reconstruction_linear.f90)
  const int ir = i < g.NI - 1 ? i : g.NI - 2;  // ring index, kept inside the mesh arrays
  const int jr = j < g.NJ ? j : g.NJ - 1;
  double uc[V], fe[V], fn[V];
  {
    const double dw = ring(g, R_RDXW, ir), de = ring(g, R_RDXE, ir);
    const double ds = g.col[C_RDYS * g.NJ + jr], dn = g.col[C_RDYN * g.NJ + jr];
#pragma unroll
    for (int l = 0; l < V; l++) {
      uc[l] = UC(l, ty, tx);
      const double sx = g.invdx * slope<LIM>(g, UC(l, ty, tx - 1), uc[l], UC(l, ty, tx + 1));
      FW(l, ty, tx) = uc[l] + sx * dw;
      fe[l] = uc[l] + sx * de;
      const double sy = g.invdy * slope<LIM>(g, UC(l, ty - 1, tx), uc[l], UC(l, ty + 1, tx));
      FS(l, ty, tx) = uc[l] + sy * ds;
      fn[l] = uc[l] + sy * dn;
    }
  }
  // extra faces: N of row -1 and S of row TY (t < 2*TX); E of column -1 and
  // W of column TX (next 2*TY threads)
  if (t < 2 * TX) {
    const int xk = t / TX;  // 0: row -1 (N face), 1: row TY (S face)
    const int xr = xk ? TY : -1, xc = tx;
    const int jj = j0 + xr, jc = jj < g.NJ ? jj : g.NJ - 1;
    const double off = xk ? g.col[C_RDYS * g.NJ + jc] : g.col[C_RDYN * g.NJ + jc];
#pragma unroll
    for (int l = 0; l < V; l++) {
      const double c0 = UC(l, xr, xc);
      const double f = c0 + g.invdy * slope<LIM>(g, UC(l, xr - 1, xc), c0, UC(l, xr + 1, xc)) * off;
      if (xk)
        FS(l, TY, xc) = f;
      else
        FN1(l, xc) = f;
    }
  } else if (t < 2 * TX + 2 * TY) {
    const int k = t - 2 * TX;
    const int xk = k / TY;  // 0: column -1 (E face), 1: column TX (W face)
    const int xr = k % TY, xc = xk ? TX : -1;
    const int ii = i0 + xc, ic = ii < g.NI ? ii : g.NI - 1;
    const double off = xk ? ring(g, R_RDXW, ic) : ring(g, R_RDXE, ic);
#pragma unroll
    for (int l = 0; l < V; l++) {
      const double c0 = UC(l, xr, xc);
      const double f = c0 + g.invdx * slope<LIM>(g, UC(l, xr, xc - 1), c0, UC(l, xr, xc + 1)) * off;
      if (xk)
        FW(l, xr, TX) = f;
      else
        FE1(l, xr) = f;
    }
  }
  __syncthreads();

  // ---- phase 3: KT interface fluxes (fluxes_kt.f90:126-176)
  double xf[V], yf[V];
  {
    double r2[V], n2[V];
#pragma unroll
    for (int l = 0; l < V; l++) {
      r2[l] = FW(l, ty, tx + 1);
      n2[l] = FS(l, ty + 1, tx);
    }
    kt_flux<V>(g, 0, fe, r2, ring(g, R_CSE, ir), ring(g, R_CSW, ir + 1), ring(g, R_DAXW, ir + 1), xf);
    kt_flux<V>(g, 1, fn, n2, ring(g, R_CSN, ir), ring(g, R_CSS, ir), ring(g, R_DAYS, ir), yf);
#pragma unroll
    for (int l = 0; l < V; l++) {
      FLX(l, ty, tx + 1) = xf[l];
      FLY(l, ty + 1, tx) = yf[l];
    }
  }
  // extra fluxes: south interface of row 0 (t < TX), west interface of
  // column 0 (threads 2*TX .. 2*TX+TY-1)
  if (t < TX) {
    double s1[V], n1[V], ef[V];
#pragma unroll
    for (int l = 0; l < V; l++) {
      n1[l] = FN1(l, tx);
      s1[l] = FS(l, 0, tx);
    }
    kt_flux<V>(g, 1, n1, s1, ring(g, R_CSN, ir), ring(g, R_CSS, ir), ring(g, R_DAYS, ir), ef);
#pragma unroll
    for (int l = 0; l < V; l++) FLY(l, 0, tx) = ef[l];
  } else if (t >= 2 * TX && t < 2 * TX + TY) {
    const int fr = t - 2 * TX;
    double e1[V], w1[V], ef[V];
#pragma unroll
    for (int l = 0; l < V; l++) {
      e1[l] = FE1(l, fr);
      w1[l] = FW(l, fr, 0);
    }
    kt_flux<V>(g, 0, e1, w1, ring(g, R_CSE, i0 - 1), ring(g, R_CSW, i0), ring(g, R_DAXW, i0), ef);
#pragma unroll
    for (int l = 0; l < V; l++) FLX(l, fr, 0) = ef[l];
  }
  __syncthreads();

  // ---- phase 4: rhs assembly (timedisc_base.f90:1254-1271)
  double err = -DBL_MAX;
  if (out) {
    const double dydzdV = ring(g, R_DYDZDV, i), dzdxdV = ring(g, R_DZDXDV, i);
    const double dxdydV = ring(g, R_DXDYDV, i);
    const unsigned ix = (unsigned)j * (unsigned)g.P + (unsigned)i;
    double rhs[V];
#pragma unroll
    for (int l = 0; l < V; l++) {
      const double xw = FLX(l, ty, tx), ys = FLY(l, ty, tx);
      rhs[l] = dydzdV * (xf[l] - xw) + dzdxdV * (yf[l] - ys) + dxdydV * (0.0 - 0.0) - geo[l] - ext[l];
      if (a.yout) (a.yout + l * g.FS)[ix] = uc[l];
    }
    if (a.ynew) {
      // final combination with this stage's rhs as k_s, and the error of the
      // embedded solution (same operation order as the reference)
      double y[V], kq[MAXK][V];
#pragma unroll
      for (int l = 0; l < V; l++) y[l] = (a.Y + l * g.FS)[ix];
#pragma unroll
      for (int q = 0; q < MAXK; q++)
        if (q < in.nk) {
#pragma unroll
          for (int l = 0; l < V; l++) kq[q][l] = (a.k[q] + l * g.FS)[ix];
        }
#pragma unroll
      for (int l = 0; l < V; l++) {
        double yh = y[l], yl = y[l];
#pragma unroll
        for (int q = 0; q < MAXK; q++)
          if (q < in.nk) {
            yl = yl - (a.bl[q] * in.dt) * kq[q][l];
            yh = yh - (a.bh[q] * in.dt) * kq[q][l];
          }
        yl = yl - (a.bl[in.nk] * in.dt) * rhs[l];
        yh = yh - (a.bh[in.nk] * in.dt) * rhs[l];
        (a.ynew + l * g.FS)[ix] = yh;
        err = fmax2(err, fabs(yh - yl) / (a.tol_rel * fabs(yh) + a.tol_abs[l]));
      }
    } else {
#pragma unroll
      for (int l = 0; l < V; l++) (a.kout + l * g.FS)[ix] = rhs[l];
    }
  }
  if (a.ynew) {
    const double m = blk_max(err, redd);
    if (t == 0) {
      const int b = blockIdx.x + blockIdx.y * gridDim.x;
      a.part_c[2 * b] = m;
      a.part_c[2 * b + 1] = 0.0;
    }
  }

  if (!(a.want_cfl || a.want_check)) return;
  if (own_g || e_g) {  // ghost cells evaluated in phase 1
    invdt = fmax2(invdt, gcfl[t]);
    flags |= gflg[t];
  }
  // ---- the corner ghost cells (in no footprint) of the corner tiles:
  // data checks, CFL and the new state (SetCornerEdges, boundary_generic.f90:434-517)
  if (t < 16) {
    const bool west = blockIdx.x == 0, east = blockIdx.x == gridDim.x - 1;
    const bool south = blockIdx.y == 0, north = blockIdx.y == gridDim.y - 1;
    const int c = t / 4, ga = 1 + (t & 1), gb = 1 + ((t >> 1) & 1);
    const int ce = c & 1, cn = c >> 1;
    if ((ce ? east : west) && (cn ? north : south)) {
      const int gi = ce ? imax + ga : imin - ga, gj = cn ? jmax + gb : jmin - gb;
      double u[V], p[V];
      in.ghost_final(gi, gj, u, p, flags);
      if (a.want_cfl) {
        const double cs = (V == 4) ? sound_speed(g.gamma, p[0], p[3]) : ring(g, R_CS, gi);
        invdt = fmax2(invdt, cell_invdt(g, gi, p[1], p[2], cs));
      }
      if (a.yout) {
        const unsigned ix = (unsigned)gj * (unsigned)g.P + (unsigned)gi;
#pragma unroll
        for (int l = 0; l < V; l++) (a.yout + l * g.FS)[ix] = u[l];
      }
    }
  }
  const double m = blk_max(invdt, redd);
  const int f = blk_or(flags, redi);
  if (t == 0) {
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    a.part[2 * b] = m;
    a.part[2 * b + 1] = (double)f;
  }
}

// ------------------------------------------------------------ FARGO remap
// FargoAdvectionY (timedisc_base.f90:1935-2069) of one interior cell from
// the combined state: per ring the shift d = w*dt/dly split into its nearest
// integer and a fraction, minmod-limited linear transport by the fraction,
// then CSHIFT(., -shift) over JMIN:JMAX.  Rows outside JMIN:JMAX come from
// the comoving-frame CenterBoundary (prim round trip, no frame change).
template <int V>
__device__ __forceinline__ void ycell(const Geom &g, const double *Y, int i, int jj, double *u) {
  const int jmin = GH, jmax = GH + g.jnum - 1;
  int js = jj;
  bool ghost = false;
  if (jj < jmin) {
    ghost = true;
    js = g.bc[2] == PERIODIC ? jj + g.jnum : jmin;
  } else if (jj > jmax) {
    ghost = true;
    js = g.bc[3] == PERIODIC ? jj - g.jnum : jmax;
  }
  const long ix = cidx(g, i, js);
#pragma unroll
  for (int l = 0; l < V; l++) u[l] = Y[l * g.FS + ix];
  if (ghost) {
    double p[V];
    c2p<V>(u, p, g.gamma);
    p2c<V>(p, u, g.gamma);
  }
}

template <int V>
__global__ __launch_bounds__(256) void k_remap(Geom g, RemapArgs a) {
  if (a.ctl && a.ctl->halt) return;
  double dts = a.dt_step;
  if (a.ctl) {  // this%dt as AcceptSolution will set it
    const double t = a.ctl->time, dt = a.ctl->dt;
    dts = (t + dt) - t;
  }
  const int i = GH + blockIdx.x * 64 + threadIdx.x, j = GH + blockIdx.y * 4 + threadIdx.y;
  if (i >= GH + g.inum || j >= GH + g.jnum) return;
  const int jmin = GH, J = g.jnum;
  double d = ring(g, R_W, i) * dts / ring(g, R_DLY, i);
  const double shift = round(d);
  d = d - shift;
  // source row: jmin + ((j - jmin - shift) mod J); fmod is exact
  int jj = (j - jmin) - (int)fmod(shift, (double)J);
  if (jj < 0) jj += J;
  if (jj >= J) jj -= J;
  const int js = jmin + jj;
  double v[5][V];
#pragma unroll
  for (int q = 0; q < 5; q++) ycell<V>(g, a.Y, i, js - 2 + q, v[q]);
  const long ix = cidx(g, i, j);
#pragma unroll
  for (int l = 0; l < V; l++) {
    double dq[4], dql[3];
#pragma unroll
    for (int q = 0; q < 4; q++) dq[q] = v[q + 1][l] - v[q][l];
#pragma unroll
    for (int q = 0; q < 3; q++)
      dql[q] = (fsign(1.0, dq[q]) * fsign(1.0, dq[q + 1]) > 0.) ? fsign(fmin2(fabs(dq[q]), fabs(dq[q + 1])), dq[q])
                                                              : 0.;
    double f0, f1;  // flux(js-1), flux(js)
    if (d > 0.) {
      f0 = v[1][l] + .5 * dql[0] * (1. - d);
      f1 = v[2][l] + .5 * dql[1] * (1. - d);
    } else {
      f0 = v[2][l] - .5 * dql[1] * (1. + d);
      f1 = v[3][l] - .5 * dql[2] * (1. + d);
    }
    a.out[l * g.FS + ix] = v[2][l] - d * (f1 - f0);
  }
}

// Step control of a device-controlled batch (one block): embedded error of
// the step from the partials of the last stage, accept test, then CFL and data checks
// from the end-of-step k_rhs partials and the next time step.
__global__ __launch_bounds__(1024) void k_control(Ctl *ctl, const double *part_c, const double *part_r, int nb) {
  __shared__ double redd[16];
  __shared__ int redi[16];
  if (ctl->halt) return;
  double e = -DBL_MAX, m = -DBL_MAX;
  int f = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    e = fmax2(e, part_c[2 * b]);
    m = fmax2(m, part_r[2 * b]);
    f |= (int)part_r[2 * b + 1];
  }
  e = blk_max(e, redd);
  m = blk_max(m, redd);
  f = blk_or(f, redi);
  if (threadIdx.x == 0) {
    control_accept(ctl, e);
    if (!ctl->halt) control_next(ctl, m, f);
  }
}

// FARGO frame change on all cells (physics_eulerisotherm.f90:1263-1413):
// sign +1 adds rho*w to the azimuthal momentum, -1 subtracts it
template <int V>
__global__ __launch_bounds__(256) void k_frame(Geom g, const double *src, double *dst, int sign) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.NI || j >= g.NJ) return;
  const long ix = cidx(g, i, j);
  const double w = ring(g, R_W, i);
#pragma unroll
  for (int l = 0; l < V; l++) {
    double v = src[l * g.FS + ix];
    if (l == 2) v = sign > 0 ? v + src[ix] * w : v - src[ix] * w;
    dst[l * g.FS + ix] = v;
  }
}

// max over part[2b] and OR over part[2b+1]; out[0] = max, out[1] = flags
__global__ __launch_bounds__(256) void k_reduce(const double *part, int n, double *out) {
  __shared__ double redd[4];
  __shared__ int redi[4];
  double m = -DBL_MAX;
  int f = 0;
  for (int b = threadIdx.x; b < n; b += 256) {
    m = fmax2(m, part[2 * b]);
    f |= (int)part[2 * b + 1];
  }
  m = blk_max(m, redd);
  f = blk_or(f, redi);
  if (threadIdx.x == 0) {
    out[0] = m;
    out[1] = (double)f;
  }
}

// -------------------------------------------------------------- launchers
int rhs_blocks(const Geom &g) { return ((g.inum + g.tile_x - 1) / g.tile_x) * ((g.jnum + TY_RHS - 1) / TY_RHS); }

template <int V, int LIM, int TX>
static void rhs_lim(const Geom &g, const StageArgs &a, hipStream_t s) {
  dim3 grid((g.inum + TX - 1) / TX, (g.jnum + TY_RHS - 1) / TY_RHS), block(TX, TY_RHS);
  hipLaunchKernelGGL((k_rhs<V, LIM, TX>), grid, block, 0, s, g, a);
}
template <int V, int TX>
static void rhs_v(const Geom &g, const StageArgs &a, hipStream_t s) {
  // compile-time limiters for the common choices, run-time switch otherwise
  if (g.limiter == VANLEER)
    rhs_lim<V, VANLEER, TX>(g, a, s);
  else if (g.limiter == MINMOD)
    rhs_lim<V, MINMOD, TX>(g, a, s);
  else
    rhs_lim<V, 0, TX>(g, a, s);
}
void launch_rhs(const Geom &g, const StageArgs &a, int V, hipStream_t s) {
  if (V == 3) {
    if (g.tile_x == 32) rhs_v<3, 32>(g, a, s);
    else rhs_v<3, 64>(g, a, s);
  } else {
    if (g.tile_x == 32) rhs_v<4, 32>(g, a, s);
    else rhs_v<4, 64>(g, a, s);
  }
}
void launch_remap(const Geom &g, const RemapArgs &a, int V, hipStream_t s) {
  dim3 grid((g.inum + 63) / 64, (g.jnum + 3) / 4), block(64, 4);
  if (V == 3)
    hipLaunchKernelGGL(k_remap<3>, grid, block, 0, s, g, a);
  else
    hipLaunchKernelGGL(k_remap<4>, grid, block, 0, s, g, a);
}
void launch_control(Ctl *ctl, const double *part_c, const double *part_r, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_control, dim3(1), dim3(1024), 0, s, ctl, part_c, part_r, nb);
}
void launch_frame(const Geom &g, const double *src, double *dst, int sign, int V, hipStream_t s) {
  dim3 grid((g.NI + 63) / 64, (g.NJ + 3) / 4), block(64, 4);
  if (V == 3)
    hipLaunchKernelGGL(k_frame<3>, grid, block, 0, s, g, src, dst, sign);
  else
    hipLaunchKernelGGL(k_frame<4>, grid, block, 0, s, g, src, dst, sign);
}
void launch_reduce(const double *part, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, part, n, out);
}

}  // namespace fv
