// fv_kernels.hip -- CDNA4 (gfx950) kernels of fosite's unsplit finite-volume
// step, FP64.
//
// Every floating-point expression keeps the reference's operation order and
// the file is compiled with -ffp-contract=off, so IEEE +,-,*,/ and sqrt give
// the reference's bits.  Citations are file:line in the fosite source tree.
//
// Kernels per right-hand-side evaluation:
//   k_ghost  boundary conditions + corner fill + ghost prim->cons for the
//            stage state (boundary_generic.f90:287-517), written to the ghost
//            cells of a work array;
//   k_rhs    fused: stage state on the fly (timedisc_rkfehlberg.f90:277-302),
//            FARGO frame round trip, linear reconstruction, face conversions,
//            wave speeds, physical and KT fluxes, geometric/gravity/FARGO
//            sources, rhs assembly, CFL and data checks.  64 x BY threads per
//            tile, LDS-staged halo, one cell per thread.
// Per step: k_combine (final stage combination + embedded error norm),
// k_remap (FARGO linear transport + integer shift), small reductions.
#include <float.h>
#include <math.h>

#include "fv_device.h"

namespace fv {

#define TINYD DBL_MIN

__device__ __forceinline__ double fmin2(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double fmax2(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }
__device__ __forceinline__ bool bad(double x) { return (x != x) || (x > DBL_MAX); }

__device__ __forceinline__ double ring(const Geom &g, int id, int i) { return g.ring[id * g.NI + i]; }
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

// ------------------------------------------------------------- stage state
// S = Y - c0*k0 - c1*k1 - ... in the reference's accumulation order
template <int V>
__device__ __forceinline__ void stage_state(const Geom &g, const StageArgs &a, long ix, double *s) {
#pragma unroll
  for (int l = 0; l < V; l++) {
    double v = a.Y[l * g.FS + ix];
    for (int q = 0; q < a.nk; q++) v = v - a.c[q] * a.k[q][l * g.FS + ix];
    s[l] = v;
  }
}

// ComputeRHS state transform on one interior cell (timedisc_base.f90:1098-1208):
// ur = cvar after the frame change and CenterBoundary/PMIN (real frame),
// pr = pvar at the same point, uc = cvar handed to the fluxes.
template <int V>
__device__ __forceinline__ void transform(const Geom &g, const StageArgs &a, int i, const double *s,
                                          double *ur, double *pr, double *uc) {
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
template <int V>
__device__ __forceinline__ void prim_interior(const Geom &g, const StageArgs &a, int i, int j, double *p) {
  double s[V];
  stage_state<V>(g, a, cidx(g, i, j), s);
  if (a.add_bg) s[2] = s[2] + s[0] * ring(g, R_W, i);
  c2p<V>(s, p, g.gamma);
}

// ------------------------------------------------------------- boundaries
// absorbing boundary for isothermal physics along x: boundary_absorbing.f90:114-150,
// physics_eulerisotherm.f90:1764-1849, 2027-2085, 2895-2969
template <int V>
__device__ void absorbing_x(const Geom &g, const StageArgs &a, int east, int j, double *p) {
  const int imin = GH, imax = GH + g.inum - 1;
  const int i1 = east ? imax : imin, i2 = east ? imax - 1 : imin + 1;
  const int iL = i1 < i2 ? i1 : i2, iR = i1 < i2 ? i2 : i1;
  const int delta = east ? 1 : -1;
  double pL[V], pR[V];
  prim_interior<V>(g, a, iL, j, pL);
  prim_interior<V>(g, a, iR, j, pR);
  const double *p1 = east ? pR : pL;  // values at i1
  double cs = ring(g, R_CS, i1), u1 = p1[1];
  double lam0 = u1 - cs, lam1 = u1, lam2 = u1 + cs;
  double csf = ring(g, R_CSW, iR);
  double dlnrho = log(pR[0] / pL[0]);
  double ducs = (pR[1] - pL[1]) / csf;
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
  double cs2 = ring(g, east ? R_CSE : R_CSW, i1);
  p[0] = p1[0] * exp(delta * 0.5 * (x2 + x0));
  p[1] = u1 + delta * 0.5 * cs2 * (x2 - x0);
  p[2] = p1[2] + delta * x1;
}

// pvar of the x-ghost cell (side, layer gl in 1..2) at interior row j
template <int V>
__device__ void ghost_we(const Geom &g, const StageArgs &a, int east, int gl, int j, double *p) {
  const int imin = GH, imax = GH + g.inum - 1, bc = g.bc[east];
  if (bc == NO_GRADIENTS)
    prim_interior<V>(g, a, east ? imax : imin, j, p);
  else if (bc == PERIODIC)
    prim_interior<V>(g, a, east ? imin + gl - 1 : imax - gl + 1, j, p);
  else
    absorbing_x<V>(g, a, east, j, p);
}
// pvar of the y-ghost cell (side, layer gl) at interior column i
template <int V>
__device__ void ghost_sn(const Geom &g, const StageArgs &a, int north, int gl, int i, double *p) {
  const int jmin = GH, jmax = GH + g.jnum - 1, bc = g.bc[2 + north];
  if (bc == NO_GRADIENTS)
    prim_interior<V>(g, a, i, north ? jmax : jmin, p);
  else
    prim_interior<V>(g, a, i, north ? jmin + gl - 1 : jmax - gl + 1, p);
}

__device__ __forceinline__ double blk_max(double v, double *red) {
  // wave64 reduction then LDS across waves
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = fmax2(v, w);
  }
  const int t = threadIdx.x + threadIdx.y * blockDim.x, nw = (blockDim.x * blockDim.y + 63) / 64;
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double r = -DBL_MAX;
  for (int k = 0; k < nw; k++) r = fmax2(r, red[k]);
  return r;
}
__device__ __forceinline__ int blk_or(int v, int *red) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  const int t = threadIdx.x + threadIdx.y * blockDim.x, nw = (blockDim.x * blockDim.y + 63) / 64;
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  int r = 0;
  for (int k = 0; k < nw; k++) r |= red[k];
  return r;
}

// CFL contribution of one cell (timedisc_base.f90:869-872 with
// physics_eulerisotherm.f90:530-552); p = pvar after the frame change back
__device__ __forceinline__ double cell_invdt(const Geom &g, int i, double u, double v, double cs) {
  double mnx = fmin2(0., u - cs), mxx = fmax2(0., u + cs);
  double mny = fmin2(0., v - cs), mxy = fmax2(0., v + cs);
  return fmax2(mxx, -mnx) / ring(g, R_DLX, i) + fmax2(mxy, -mny) / ring(g, R_DLY, i);
}

// One thread per ghost cell: W/E slabs (interior rows), S/N slabs (interior
// columns), 16 corner cells (SetCornerEdges, boundary_generic.f90:434-517).
template <int V>
__global__ __launch_bounds__(256) void k_ghost(Geom g, StageArgs a) {
  __shared__ double redd[8];
  __shared__ int redi[8];
  const int imin = GH, imax = GH + g.inum - 1, jmin = GH, jmax = GH + g.jnum - 1;
  const int nwe = 2 * 2 * g.jnum, nsn = 2 * 2 * g.inum;
  int t = blockIdx.x * 256 + threadIdx.x;
  int i = -1, j = -1;
  double p[V];
  bool active = true;
  if (t < nwe) {  // x-ghosts: t -> (side, layer, j)
    int east = t / (2 * g.jnum), r = t % (2 * g.jnum), gl = 1 + r / g.jnum;
    j = jmin + r % g.jnum;
    i = east ? imax + gl : imin - gl;
    ghost_we<V>(g, a, east, gl, j, p);
  } else if ((t -= nwe) < nsn) {
    int north = t / (2 * g.inum), r = t % (2 * g.inum), gl = 1 + r / g.inum;
    i = imin + r % g.inum;
    j = north ? jmax + gl : jmin - gl;
    ghost_sn<V>(g, a, north, gl, i, p);
  } else if ((t -= nsn) < 16) {
    int corner = t / 4, ga = 1 + (t & 1), gb = 1 + ((t >> 1) & 1);
    int east = corner & 1, north = corner >> 1;
    int ir = east ? imax : imin, jr = north ? jmax : jmin;
    i = east ? imax + ga : imin - ga;
    j = north ? jmax + gb : jmin - gb;
    double px[V], py[V];
    if (ga == 2 && gb == 1) {
      ghost_we<V>(g, a, east, 2, jr, p);  // copy from the x-ghost at the boundary row
    } else if (ga == 1 && gb == 2) {
      ghost_sn<V>(g, a, north, 2, ir, p);  // copy from the y-ghost at the boundary column
    } else {
      ghost_we<V>(g, a, east, ga, jr, px);
      ghost_sn<V>(g, a, north, gb, ir, py);
#pragma unroll
      for (int l = 0; l < V; l++) p[l] = 0.5 * (px[l] + py[l]);
    }
  } else {
    active = false;
  }
  double invdt = -DBL_MAX;
  int flags = 0;
  if (active) {
    double u[V];
    if (V == 4 && a.pmin_clamp) p[3] = fmax2(p[3], g.pmin);
    p2c<V>(p, u, g.gamma);
    if (a.want_check) {
#pragma unroll
      for (int l = 0; l < V; l++)
        if (bad(u[l]) || bad(p[l])) flags |= F_INVALID;
    }
    if (a.sub_bg) {
      double w = ring(g, R_W, i);
      p[2] = p[2] - w;
      u[2] = u[2] - u[0] * w;
    }
    if (a.want_cfl) {
      double cs = (V == 4) ? sound_speed(g.gamma, p[0], p[3]) : ring(g, R_CS, i);
      invdt = cell_invdt(g, i, p[1], p[2], cs);
    }
    const long ix = cidx(g, i, j);
#pragma unroll
    for (int l = 0; l < V; l++) a.gdst[l * g.FS + ix] = u[l];
  }
  if (a.want_cfl || a.want_check) {
    double m = blk_max(invdt, redd);
    int f = blk_or(flags, redi);
    if (threadIdx.x == 0) {
      a.part[2 * (a.part_off + blockIdx.x)] = m;
      a.part[2 * (a.part_off + blockIdx.x) + 1] = (double)f;
    }
  }
}

// ----------------------------------------------------------- reconstruction
// limiters, reconstruction_linear.f90:645-733
__device__ __forceinline__ double limiter(const Geom &g, double a1, double a2) {
  switch (g.limiter) {
    case MINMOD:
      return (fsign(1.0, a1) * fsign(1.0, a2) > 0) ? fsign(fmin2(fabs(a1), fabs(a2)), a1) : 0.;
    case SWEBY:
    case SUPERBEE: {
      double q = g.limiter == SUPERBEE ? 2.0 : g.theta;
      return (fsign(1.0, a1) * fsign(1.0, a2) > 0)
                 ? fsign(fmax2(fmin2(q * fabs(a1), fabs(a2)), fmin2(fabs(a1), q * fabs(a2))), a1)
                 : 0.;
    }
    case OSPRE:
      return 1.5 * a1 * a2 * (a1 + a2) / (a1 * (a1 + 0.5 * a2) + a2 * (a2 + 0.5 * a1) + TINYD);
    case VANLEER: {
      double b1 = fabs(a1), b2 = fabs(a2);
      return (a1 * b2 + a2 * b1) / (b1 + b2 + TINYD);
    }
    default:
      return 0.5 * (a1 + a2);
  }
}
__device__ __forceinline__ double slope(const Geom &g, double um, double u0, double up) {
  if (g.limiter == MONOCENT) {  // minmod3 of theta-scaled one-sided and central differences
    double a1 = g.theta * (u0 - um), a2 = g.theta * (up - u0), a3 = 0.5 * (up - um);
    if ((fsign(1.0, a1) * fsign(1.0, a2) > 0) && (fsign(1.0, a2) * fsign(1.0, a3) > 0))
      return fsign(fmin2(fmin2(fabs(a1), fabs(a2)), fabs(a3)), a1);
    return 0.;
  }
  return limiter(g, u0 - um, up - u0);
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
  double vL = pl[un], vR = pr[un];
  double amin = fmin2(fmin2(0.0, fmin2(0., vR - csR)), fmin2(0., vL - csL));
  double amax = fmax2(fmax2(0.0, fmax2(0., vR + csR)), fmax2(0., vL + csL));
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
  double fac = dA / (amax - amin);
#pragma unroll
  for (int l = 0; l < V; l++) F[l] = fac * (amax * fL[l] - amin * fR[l] + amin * amax * (UR[l] - UL[l]));
}

// -------------------------------------------------------------- fused RHS
template <int V, int BY>
__global__ __launch_bounds__(TILE_X *BY) void k_rhs(Geom g, StageArgs a) {
  constexpr int BX = TILE_X, SX = BX + 2, SY = BY + 2;
  constexpr int NUC = V * SY * SX, NF = 2 * V * BY * BX, NL = NUC > NF ? NUC : NF;
  __shared__ double lds[NL];
  __shared__ double redd[BY];
  __shared__ int redi[BY];
  double *uc_s = lds;                  // [V][SY][SX] flux-input state
  double *fa_s = lds;                  // [V][BY][BX] W faces, then x fluxes
  double *fb_s = lds + V * BY * BX;    // [V][BY][BX] S faces, then y fluxes

  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = GH + blockIdx.x * (BX - 2), j0 = GH + blockIdx.y * (BY - 2);
  const int fi = i0 - 1 + tx, fj = j0 - 1 + ty;  // this thread's cell
  const int imax = GH + g.inum - 1, jmax = GH + g.jnum - 1;
  const bool out = tx >= 1 && tx <= BX - 2 && ty >= 1 && ty <= BY - 2 && fi <= imax && fj <= jmax;

  // ---- phase 1: flux-input state of the tile and its halo into LDS
  double ur[V], pr[V], uc[V];
  auto load = [&](int i, int j, int sx, int sy, bool own) {
    double u[V];
    if (i < 0 || j < 0 || i >= g.NI || j >= g.NJ) {
#pragma unroll
      for (int l = 0; l < V; l++) u[l] = 0.0;
    } else if (i < GH || j < GH || i > imax || j > jmax) {
      const long ix = cidx(g, i, j);
#pragma unroll
      for (int l = 0; l < V; l++) u[l] = a.gsrc[l * g.FS + ix];
    } else {
      double s[V], r[V], p[V];
      stage_state<V>(g, a, cidx(g, i, j), s);
      transform<V>(g, a, i, s, r, p, u);
      if (own) {
#pragma unroll
        for (int l = 0; l < V; l++) { ur[l] = r[l]; pr[l] = p[l]; }
      }
    }
#pragma unroll
    for (int l = 0; l < V; l++) uc_s[(l * SY + sy) * SX + sx] = u[l];
    if (own) {
#pragma unroll
      for (int l = 0; l < V; l++) uc[l] = u[l];
    }
  };
  load(fi, fj, tx + 1, ty + 1, true);
  if (tx == 0 && ty >= 1 && ty <= BY - 2) load(fi - 1, fj, 0, ty + 1, false);
  if (tx == BX - 1 && ty >= 1 && ty <= BY - 2) load(fi + 1, fj, SX - 1, ty + 1, false);
  if (ty == 0 && tx >= 1 && tx <= BX - 2) load(fi, fj - 1, tx + 1, 0, false);
  if (ty == BY - 1 && tx >= 1 && tx <= BX - 2) load(fi, fj + 1, tx + 1, SY - 1, false);

  // sources of the own cell, from the real-frame state (ur, pr)
  double geo[V], ext[V];
  double invdt = -DBL_MAX;
  int flags = 0;
#pragma unroll
  for (int l = 0; l < V; l++) geo[l] = ext[l] = 0.0;
  if (out) {
    // geometrical sources (physics_eulerisotherm.f90:689-832, elementals 3069-3088;
    // physics_euler.f90:1706-1846); zero on cartesian meshes
    if (g.curvilinear) {
      const double cxyx = ring(g, R_CXYX, fi), cyxy = ring(g, R_CYXY, fi);
      const double czxz = ring(g, R_CZXZ, fi), czyz = ring(g, R_CZYZ, fi);
      const double vx = pr[1], vy = pr[2], vz = 0.0, mx = ur[1], my = ur[2], mz = 0.0;
      double P;
      if (V == 4) {
        P = pr[3];
      } else {
        const double cs = ring(g, R_CS, fi);
        P = pr[0] * (cs * cs);
      }
      geo[1] = my * (cyxy * vy - cxyx * vx) + mz * (czxz * vz - 0.0 * vx) + (cyxy + czxz) * P;
      geo[2] = mz * (czyz * vz - 0.0 * vy) + mx * (cxyx * vx - cyxy * vy) + (cxyx + czyz) * P;
    }
    // FARGO source (physics_eulerisotherm.f90:1531-1586)
    if (g.fargo)
      geo[2] = geo[2] - ur[1] * 0.5 * (ring(g, R_W, fi + 1) - ring(g, R_W, fi - 1)) / ring(g, R_DLX, fi);
    // point-mass gravity summed into the source list (sources_generic.f90:220-261,
    // sources_gravity.f90:164-199, physics_eulerisotherm.f90:1201-1226,
    // physics_euler.f90:1849-1885)
    if (g.gravity) {
      const double ax = ring(g, R_ACCX, fi), ay = ring(g, R_ACCY, fi);
      ext[1] = 0.0 + ur[0] * ax;
      ext[2] = 0.0 + ur[0] * ay;
      if (V == 4) ext[3] = 0.0 + (ur[1] * ax + ur[2] * ay);
    }
    // data checks (timedisc_base.f90:1666-1742) and CFL (timedisc_base.f90:818-909)
    if (a.want_check) {
      if (pr[0] < g.rhomin) flags |= F_RHOMIN;
      if (V == 4 && pr[3] < g.pmin) flags |= F_PMIN;
#pragma unroll
      for (int l = 0; l < V; l++)
        if (bad(pr[l]) || bad(ur[l])) flags |= F_INVALID;
    }
    if (a.want_cfl) {
      double vy = a.sub_bg ? pr[2] - ring(g, R_W, fi) : pr[2];
      double cs = (V == 4) ? sound_speed(g.gamma, pr[0], pr[3]) : ring(g, R_CS, fi);
      invdt = cell_invdt(g, fi, pr[1], vy, cs);
    }
  }
  __syncthreads();

  // ---- phase 2: limited slopes and face states
  const int sx = tx + 1, sy = ty + 1;
  const bool xface = ty >= 1 && ty <= BY - 2;       // rows that need x fluxes
  const bool yface = tx >= 1 && tx <= BX - 2;       // columns that need y fluxes
  double fw[V], fe[V], fs[V], fn[V];
  if (xface) {
    const double dw = ring(g, R_RDXW, fi < g.NI ? fi : g.NI - 1), de = ring(g, R_RDXE, fi < g.NI ? fi : g.NI - 1);
#pragma unroll
    for (int l = 0; l < V; l++) {
      const double *r = uc_s + (l * SY + sy) * SX;
      double sl = g.invdx * slope(g, r[sx - 1], r[sx], r[sx + 1]);
      fw[l] = uc[l] + sl * dw;
      fe[l] = uc[l] + sl * de;
    }
  }
  if (yface) {
    const int jc = fj < g.NJ ? fj : g.NJ - 1;
    const double ds = g.col[C_RDYS * g.NJ + jc], dn = g.col[C_RDYN * g.NJ + jc];
#pragma unroll
    for (int l = 0; l < V; l++) {
      const double *c = uc_s + l * SY * SX + sx;
      double sl = g.invdy * slope(g, c[(sy - 1) * SX], c[sy * SX], c[(sy + 1) * SX]);
      fs[l] = uc[l] + sl * ds;
      fn[l] = uc[l] + sl * dn;
    }
  }
  __syncthreads();
#pragma unroll
  for (int l = 0; l < V; l++) {
    if (xface) fa_s[(l * BY + ty) * BX + tx] = fw[l];
    if (yface) fb_s[(l * BY + ty) * BX + tx] = fs[l];
  }
  __syncthreads();

  // ---- phase 3: interface fluxes (east interface of x-cells, north of y-cells)
  double xf[V], yf[V];
  const int fic = fi < g.NI - 1 ? fi : g.NI - 2;
  if (xface && tx <= BX - 2) {
    double ur2[V];
#pragma unroll
    for (int l = 0; l < V; l++) ur2[l] = fa_s[(l * BY + ty) * BX + tx + 1];
    kt_flux<V>(g, 0, fe, ur2, ring(g, R_CSE, fic), ring(g, R_CSW, fic + 1), ring(g, R_DAXW, fic + 1), xf);
  }
  if (yface && ty <= BY - 2) {
    double un2[V];
#pragma unroll
    for (int l = 0; l < V; l++) un2[l] = fb_s[(l * BY + ty + 1) * BX + tx];
    kt_flux<V>(g, 1, fn, un2, ring(g, R_CSN, fic), ring(g, R_CSS, fic), ring(g, R_DAYS, fic), yf);
  }
  __syncthreads();
#pragma unroll
  for (int l = 0; l < V; l++) {
    if (xface && tx <= BX - 2) fa_s[(l * BY + ty) * BX + tx] = xf[l];
    if (yface && ty <= BY - 2) fb_s[(l * BY + ty) * BX + tx] = yf[l];
  }
  __syncthreads();

  // ---- phase 4: rhs assembly (timedisc_base.f90:1254-1271)
  if (out) {
    const double dydzdV = ring(g, R_DYDZDV, fi), dzdxdV = ring(g, R_DZDXDV, fi);
    const double dxdydV = ring(g, R_DXDYDV, fi);
    const long ix = cidx(g, fi, fj);
#pragma unroll
    for (int l = 0; l < V; l++) {
      double xw = fa_s[(l * BY + ty) * BX + tx - 1], ys = fb_s[(l * BY + ty - 1) * BX + tx];
      double r = dydzdV * (xf[l] - xw) + dzdxdV * (yf[l] - ys) + dxdydV * (0.0 - 0.0) - geo[l] - ext[l];
      a.kout[l * g.FS + ix] = r;
      if (a.yout) a.yout[l * g.FS + ix] = uc[l];
    }
  }
  if (a.want_cfl || a.want_check) {
    double m = blk_max(invdt, redd);
    int f = blk_or(flags, redi);
    if (tx == 0 && ty == 0) {
      const int b = a.part_off + blockIdx.x + blockIdx.y * gridDim.x;
      a.part[2 * b] = m;
      a.part[2 * b + 1] = (double)f;
    }
  }
}

// ------------------------------------------------------ final combination
template <int V>
__global__ __launch_bounds__(256) void k_combine(Geom g, CombineArgs a) {
  __shared__ double redd[4];
  const int i = GH + blockIdx.x * 64 + threadIdx.x, j = GH + blockIdx.y * 4 + threadIdx.y;
  double e = -DBL_MAX;
  if (i < GH + g.inum && j < GH + g.jnum) {
    const long ix = cidx(g, i, j);
#pragma unroll
    for (int l = 0; l < V; l++) {
      double y = a.Y[l * g.FS + ix], yh = y, yl = y;
      for (int q = 0; q < a.nk; q++) {
        double k = a.k[q][l * g.FS + ix];
        yl = yl - a.cl[q] * k;
        yh = yh - a.ch[q] * k;
      }
      a.ynew[l * g.FS + ix] = yh;
      double c = fabs(yh - yl) / (a.tol_rel * fabs(yh) + a.tol_abs[l]);
      e = fmax2(e, c);
    }
  }
  double m = blk_max(e, redd);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    a.part[2 * b] = m;
    a.part[2 * b + 1] = 0.0;
  }
}

// ------------------------------------------------------------ FARGO remap
// timedisc_base.f90:1935-2069: minmod-limited linear transport by the
// fractional shift, then the integer circular shift CSHIFT(., -shift).
// One thread per interior output cell; x-ghost rings are rebuilt by the next
// boundary update, so only interior rings are transported.
template <int V>
__global__ __launch_bounds__(256) void k_remap(Geom g, const double *u, double *z, double dt) {
  const int i = GH + blockIdx.x * 64 + threadIdx.x, j = GH + blockIdx.y * 4 + threadIdx.y;
  if (i >= GH + g.inum || j >= GH + g.jnum) return;
  const int jmin = GH, J = g.jnum;
  double d = ring(g, R_W, i) * dt / ring(g, R_DLY, i);
  double shift = round(d);
  d = d - shift;
  long sh = (long)shift;
  // source row of the integer shift: result(j) = a(j - shift) cyclically
  long jj = (long)(j - jmin) - sh;
  jj %= J;
  if (jj < 0) jj += J;
  const int js = jmin + (int)jj;
#pragma unroll
  for (int l = 0; l < V; l++) {
    const double *c = u + l * g.FS + i;
    double v[5];  // rows js-2 .. js+2
#pragma unroll
    for (int q = 0; q < 5; q++) v[q] = c[(long)(js - 2 + q) * g.P];
    double dq[4];
#pragma unroll
    for (int q = 0; q < 4; q++) dq[q] = v[q + 1] - v[q];  // dq(js-2+q)
    double dql[3];                                     // dql(js-1+q)
#pragma unroll
    for (int q = 0; q < 3; q++)
      dql[q] = (fsign(1.0, dq[q]) * fsign(1.0, dq[q + 1]) > 0.) ? fsign(fmin2(fabs(dq[q]), fabs(dq[q + 1])), dq[q]) : 0.;
    double f0, f1;  // flux(js-1), flux(js)
    if (d > 0.) {
      f0 = v[1] + .5 * dql[0] * (1. - d);
      f1 = v[2] + .5 * dql[1] * (1. - d);
    } else {
      f0 = v[2] - .5 * dql[1] * (1. + d);
      f1 = v[3] - .5 * dql[2] * (1. + d);
    }
    z[l * g.FS + cidx(g, i, j)] = v[2] - d * (f1 - f0);
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
constexpr int RHS_BY = 8;

int rhs_blocks(const Geom &g) {
  return ((g.inum + TILE_X - 3) / (TILE_X - 2)) * ((g.jnum + RHS_BY - 3) / (RHS_BY - 2));
}
int ghost_blocks(const Geom &g) { return (4 * g.jnum + 4 * g.inum + 16 + 255) / 256; }
int combine_blocks(const Geom &g) { return ((g.inum + 63) / 64) * ((g.jnum + 3) / 4); }

void launch_ghost(const Geom &g, const StageArgs &a, int V, hipStream_t s) {
  dim3 grid(ghost_blocks(g)), block(256);
  if (V == 3)
    hipLaunchKernelGGL(k_ghost<3>, grid, block, 0, s, g, a);
  else
    hipLaunchKernelGGL(k_ghost<4>, grid, block, 0, s, g, a);
}
void launch_rhs(const Geom &g, const StageArgs &a, int V, hipStream_t s) {
  dim3 grid((g.inum + TILE_X - 3) / (TILE_X - 2), (g.jnum + RHS_BY - 3) / (RHS_BY - 2));
  dim3 block(TILE_X, RHS_BY);
  if (V == 3)
    hipLaunchKernelGGL((k_rhs<3, RHS_BY>), grid, block, 0, s, g, a);
  else
    hipLaunchKernelGGL((k_rhs<4, RHS_BY>), grid, block, 0, s, g, a);
}
void launch_combine(const Geom &g, const CombineArgs &a, int V, hipStream_t s) {
  dim3 grid((g.inum + 63) / 64, (g.jnum + 3) / 4), block(64, 4);
  if (V == 3)
    hipLaunchKernelGGL(k_combine<3>, grid, block, 0, s, g, a);
  else
    hipLaunchKernelGGL(k_combine<4>, grid, block, 0, s, g, a);
}
void launch_remap(const Geom &g, const double *u, double *z, double dt, int V, hipStream_t s) {
  dim3 grid((g.inum + 63) / 64, (g.jnum + 3) / 4), block(64, 4);
  if (V == 3)
    hipLaunchKernelGGL(k_remap<3>, grid, block, 0, s, g, u, z, dt);
  else
    hipLaunchKernelGGL(k_remap<4>, grid, block, 0, s, g, u, z, dt);
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
