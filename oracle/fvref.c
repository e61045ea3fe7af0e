/*
 * fvref.c -- CPU oracle: plain-C restatement of fosite's unsplit
 * finite-volume step (reference: tillense/fosite v0.8.3, Fortran 2008).
 *
 * TEST INFRASTRUCTURE ONLY (checker, never the product).  Every routine
 * mirrors the reference's full-array arithmetic in the reference's operation
 * order, so that -- built without FMA contraction -- it reproduces the
 * compiled reference bit for bit.  Citations are file:line in the reference.
 *
 * Scope: 2-D meshes (knum = 1), midpoint mesh, cartesian and logcylindrical
 * geometry, Euler and isothermal Euler physics, KT fluxes with linear
 * reconstruction (all non-PP limiters) of conservative or primitive
 * variables, embedded RK (RKF 3/5, Cash-Karp 5, Dormand-Prince 5, SSPRK 3/5)
 * with error control and reject path, no-gradient, periodic and absorbing
 * boundaries, point-mass gravity with accretion, the rotating frame about
 * the origin (centrifugal and Coriolis terms), the viscosity source
 * (MOLECULAR, ALPHA and BETA models, with its time-step limit), FARGO along y for both
 * physics and geometries with the per-ring mean (type 1) or a user-supplied
 * (type 2) background velocity, CheckData breaks,
 * output times (datafile count) and a start time.
 *
 * Every function is pinned by the golden fixtures of the compiled reference
 * (tests/golden/fv_golden.npz, tests/test_oracle_golden.py) or by the
 * invariants tested there.
 */
#include "fvref.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define G 2 /* ghost layers, mesh_base.f90:462 */
#define TINY_ DBL_MIN
#define EPS_ DBL_EPSILON
#define HUGE_ DBL_MAX

enum { PRIMITIVE = 1, CONSERVATIVE = 2 };
enum { EULER_ISOTHERM = 16, EULER = 17 };
enum { CARTESIAN = 1, LOGCYLINDRICAL = 31 };
enum { NO_GRADIENTS = 1, PERIODIC = 2, ABSORBING = 13 };
enum { MINMOD = 1, MONOCENT = 2, SWEBY = 3, SUPERBEE = 4, OSPRE = 5, VANLEER = 7, NOLIMIT = 8 };
enum { RK_FEHLBERG = 2, CASH_KARP = 3, DORMAND_PRINCE = 4, SSPRK = 5 };
enum { W_ = 0, E_ = 1, S_ = 2, N_ = 3 };
/* timedisc_base.f90:193-203 */
#define CHECK_NOTHING 0
#define CHECK_ALL (~0)
#define CHECK_CSOUND 1
#define CHECK_PMIN 2
#define CHECK_RHOMIN 4
#define CHECK_INVALID 8
#define DTCAUSE_ERRADJ (-1)
#define DTCAUSE_SMALLERR (-2)
#define DTCAUSE_FILEIO (-4)

typedef struct { double *d; int flag; } sv; /* state vector + fargo_transformation_applied */

struct fvr {
  fvr_cfg c;
  int NI, NJ, N, V, m; /* cells per row/col/total, fields, RK stages */
  double dx, dy, dz, invdx, invdy;
  /* mesh arrays (all cells) */
  double *dAx[2], *dAy[2], *dydzdV, *dzdxdV, *dxdydV, *dlx, *dly;
  double *cxyx, *cyxy, *czxz, *czyz, *rdx[4], *rad, *radf[4], *posx, *posy, *bcx, *bcy;
  double *caccel[2], *twoomega, *raccel[2]; /* rotating frame (sources_rotframe.f90:110-175, 226-234) */
  /* viscosity (sources_viscosity.f90, physics_eulerisotherm.f90:982-1198):
   * dynamic and bulk viscosity of the last update and its time, e_phi/r in
   * curvilinear components, stresses, work arrays */
  double *dynvis, *bulkvis, vis_time, *ephir[2], *btxx, *btxy, *btyy, *vtmp, *vtmp1, *vsterm;
  double *accel[2];
  double pm_mass, pm_accrate, pm_massloss, pm_mdot; /* gravity_pointmass.f90:79-85 */
  /* physics */
  double *bccsound, *fcsound; /* fcsound [face][cell] */
  double *w;                  /* fargo background velocity per x-ring */
  double *wfac;               /* FARGO type 1: geometric factor per x-ring */
  /* timedisc */
  sv pvar, cvar, ptmp, ctmp, cold;
  double *geo_src, *src, *sterm, **coeff, *cerr, *xflux, *yflux;
  /* boundary fluxes (fluxes_base.f90:142-143): bx[(l*2+side)*NJ + j],
   * by[(l*2+side)*NI + i]; bxo/byo are bxfold/byfold */
  double *bx, *by, *bxo, *byo;
  double *a, *b_high, *b_low, *cc; /* Butcher tableau, a row-major m x m */
  /* fluxes */
  double *prim, *cons, *pfluxes, *minwav, *maxwav, *tmp, *tmp1, *slopes;
  double time, dt, dtold, dtmin, maxerrold, pmin, rhomin;
  int dtcause, n_adj, iter, brk;
  double t_out; int out_count; /* Datafile output time and count */
};

#define IDX(s, i, j) ((j) * (s)->NI + (i))

/* Loops over cells whose iterations are independent run on OpenMP threads
 * on meshes large enough to pay for it.  Every value is computed by the same
 * expression as in a serial run, so the results do not depend on the number
 * of threads; reductions (sums in index order, the min/max scans) stay serial
 * or are per ring. */
#ifdef _OPENMP
#define PAR_FOR _Pragma("omp parallel for schedule(static) if (s->N >= 32768)")
#else
#define PAR_FOR
#endif

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* Fortran intrinsics with the reference's semantics */
static inline double fmin2(double a, double b) { return (b < a) ? b : a; }
static inline double fmax2(double a, double b) { return (b > a) ? b : a; }
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

/* ------------------------------------------------------------------ mesh */

/* geometry_logcylindrical.f90:310-319 (h = gp*exp(logr)); cartesian: 1 */
static double hscale(const fvr *s, double x) {
  return (s->c.geometry == LOGCYLINDRICAL) ? s->c.gparam * exp(x) : 1.0;
}

/* mesh_base.f90:558-672, mesh_midpoint.f90:97-145, geometry_logcylindrical.f90:310-342 */
static void mesh_build(fvr *s) {
  int NI = s->NI, NJ = s->NJ;
  double mdx, mdy;
  s->dx = (s->c.xmax - s->c.xmin) / s->c.inum;
  s->dy = (s->c.ymax - s->c.ymin) / s->c.jnum;
  s->dz = (0.0 - 0.0) / 1;
  mdx = s->dx;
  mdy = s->dy;
  if (s->dz < 2 * EPS_) s->dz = 1.0; /* knum = 1 -> mesh_dz = 0 */
  s->invdx = 1. / s->dx;
  s->invdy = 1. / s->dy;
  for (int j0 = 0; j0 < NJ; j0++) {
    int fj = j0 - 1;
    double yc = s->c.ymin + ((double)(2 * fj - 1) * 0.5) * mdy;
    double ys = yc + (-(0.5 * mdy)), yn = yc + 0.5 * mdy;
    for (int i0 = 0; i0 < NI; i0++) {
      int fi = i0 - 1, c = IDX(s, i0, j0);
      double xc = s->c.xmin + ((double)(2 * fi - 1) * 0.5) * mdx;
      double xw = xc + (-(0.5 * mdx)), xe = xc + 0.5 * mdx, xsn = xc + 0.0;
      double hxc = hscale(s, xc), hyc = hxc, hzc = 1.0;
      double hyW = hscale(s, xw), hyE = hscale(s, xe), hzW = 1.0, hzE = 1.0;
      double hxS = hscale(s, xsn), hxN = hxS, hzS = 1.0, hzN = 1.0;
      double sqrtg, vol;
      if (s->c.geometry == CARTESIAN) { hyW = hyE = hxS = hxN = 1.0; }
      s->bcx[c] = xc;
      s->bcy[c] = yc;
      s->dAx[0][c] = hyW * hzW;
      s->dAx[1][c] = hyE * hzE;
      s->dAy[0][c] = hzS * hxS;
      s->dAy[1][c] = hzN * hxN;
      sqrtg = hxc * (hyc * hzc);
      vol = sqrtg * s->dx * s->dy * s->dz;
      s->dxdydV[c] = s->dx * s->dy / (vol + TINY_);
      s->dydzdV[c] = s->dy * s->dz / (vol + TINY_);
      s->dzdxdV[c] = s->dz * s->dx / (vol + TINY_);
      /* top/bottom faces coincide with the centre (mesh_dz = 0) */
      s->cyxy[c] = 0.5 * (hzE + hzW) * (hyE - hyW) * s->dydzdV[c];
      s->cxyx[c] = 0.5 * (hzN + hzS) * (hxN - hxS) * s->dzdxdV[c];
      s->czxz[c] = 0.5 * (hyE + hyW) * (hzE - hzW) * s->dydzdV[c];
      s->czyz[c] = 0.5 * (hxN + hxS) * (hzN - hzS) * s->dzdxdV[c];
      s->dlx[c] = hxc * s->dx;
      s->dly[c] = hyc * s->dy;
      /* reconstruction_linear.f90:156-177: faces - bcenter (bcenter = center) */
      s->rdx[W_][c] = xw - xc;
      s->rdx[E_][c] = xe - xc;
      s->rdx[S_][c] = ys - yc;
      s->rdx[N_][c] = yn - yc;
      if (s->c.geometry == LOGCYLINDRICAL) {
        double r = s->c.gparam * exp(xc), rw = s->c.gparam * exp(xw), re = s->c.gparam * exp(xe);
        double rs = s->c.gparam * exp(xsn);
        s->rad[c] = sqrt(r * r + 0.0 * 0.0);
        s->radf[W_][c] = sqrt(rw * rw + 0.0 * 0.0);
        s->radf[E_][c] = sqrt(re * re + 0.0 * 0.0);
        s->radf[S_][c] = sqrt(rs * rs + 0.0 * 0.0);
        s->radf[N_][c] = s->radf[S_][c];
        s->posx[c] = s->rad[c];
        s->posy[c] = 0.0;
      } else {
        s->rad[c] = sqrt(xc * xc + yc * yc + 0.0 * 0.0);
        s->radf[W_][c] = sqrt(xw * xw + yc * yc);
        s->radf[E_][c] = sqrt(xe * xe + yc * yc);
        s->radf[S_][c] = sqrt(xc * xc + ys * ys);
        s->radf[N_][c] = sqrt(xc * xc + yn * yn);
        s->posx[c] = xc;
        s->posy[c] = yc;
      }
    }
  }
  /* gravity_pointmass.f90:263-282 (point mass at the origin, NEWTON, eps = 0),
   * then sources_gravity.f90:226 (accel = 0 + accel_pointmass) */
  if (s->c.gravity) {
    for (int c = 0; c < s->N; c++) {
      double r = s->rad[c];
      double om2 = 1. / fmax2(10 * TINY_, ((r * ((r - 0.0) * (r - 0.0))) + 0.0 * 0.0 * 0.0) /
                                            (s->c.GN * s->c.mass));
      s->accel[0][c] = 0.0 + (-(om2 * s->posx[c]));
      s->accel[1][c] = 0.0 + (-(om2 * s->posy[c]));
    }
  }
  /* sources_rotframe.f90:110-175 with the rotation centre at the origin:
   * Omez = Omega*e_z on the interior cells, 0 on ghost cells, in curvilinear
   * components (e_z is a basis vector of both geometries); vphi = Omez x r,
   * caccel = vphi x Omez (marray_base.f90:855-897, axb1 = a2*b3 - a3*b2);
   * r = posvec at the bary centres, z component 0 on these 2-D meshes;
   * twoOmega = 2*Omez_z (VECTOR_X|VECTOR_Y) */
  if (s->c.rotframe) {
    for (int j = 0; j < s->NJ; j++)
      for (int i = 0; i < s->NI; i++) {
        const int c = IDX(s, i, j);
        const int inner = i >= G && i < G + s->c.inum && j >= G && j < G + s->c.jnum;
        const double o1 = 0.0, o2 = 0.0, o3 = inner ? s->c.omega : 0.0;
        const double r1 = s->posx[c], r2 = s->posy[c], r3 = 0.0;
        const double v1 = o2 * r3 - o3 * r2, v2 = o3 * r1 - o1 * r3, v3 = o1 * r2 - o2 * r1;
        s->caccel[0][c] = v2 * o3 - v3 * o2;
        s->caccel[1][c] = v3 * o1 - v1 * o3;
        s->twoomega[c] = 2 * o3;
      }
  }
  /* sources_viscosity.f90:169-178 (dynvis = dynconst, bulkvis = bulkconst,
   * the update time -1) and 196-223 (ALPHA): e_phi/r = (-y/r^2, x/r^2, 0)
   * of the Cartesian bary centres (geometry_logcylindrical.f90:344-356),
   * in curvilinear components (geometry_cylindrical.f90:485-494; the
   * Cartesian geometry keeps them) */
  if (s->c.vismodel) {
    s->vis_time = -1.0;
    for (int c = 0; c < s->N; c++) {
      s->dynvis[c] = s->c.dynconst;
      s->bulkvis[c] = s->c.bulkconst;
    }
  }
  if (s->c.vismodel == 2 || s->c.vismodel == 3) {
    for (int c = 0; c < s->N; c++) {
      const double phi = s->bcy[c];
      double x = s->bcx[c], y = s->bcy[c];
      if (s->c.geometry == LOGCYLINDRICAL) {
        const double r = s->c.gparam * exp(s->bcx[c]);
        x = r * cos(phi);
        y = r * sin(phi);
      }
      const double r2 = s->rad[c] * s->rad[c];
      const double ex = -(y / r2), ey = x / r2;
      if (s->c.geometry == LOGCYLINDRICAL) {
        s->ephir[0][c] = ex * cos(phi) + ey * sin(phi);
        s->ephir[1][c] = -(ex * sin(phi)) + ey * cos(phi);
      } else {
        s->ephir[0][c] = ex;
        s->ephir[1][c] = ey;
      }
    }
  }
}

/* --------------------------------------------------------------- physics */

/* physics_eulerisotherm.f90:2835-2846, physics_euler.f90:3056-3068 */
static void cons2prim_pt(const fvr *s, double *p, const double *c, size_t n, size_t k) {
  double rho = c[0 * n + k], inv = 1. / rho;
  double mu = c[1 * n + k], mv = c[2 * n + k];
  p[0 * n + k] = rho;
  p[1 * n + k] = mu * inv;
  p[2 * n + k] = mv * inv;
  if (s->V == 4) p[3 * n + k] = (s->c.gamma - 1.) * (c[3 * n + k] - 0.5 * inv * (mu * mu + mv * mv));
}

/* physics_eulerisotherm.f90:2872-2881, physics_euler.f90:3094-3104 */
static void prim2cons_pt(const fvr *s, double *c, const double *p, size_t n, size_t k) {
  double rho = p[0 * n + k], u = p[1 * n + k], v = p[2 * n + k];
  c[0 * n + k] = rho;
  c[1 * n + k] = rho * u;
  c[2 * n + k] = rho * v;
  if (s->V == 4) c[3 * n + k] = p[3 * n + k] / (s->c.gamma - 1.) + 0.5 * rho * (u * u + v * v);
}

/* physics_euler.f90:2682-2689 */
static inline double sound_speed(double gamma, double rho, double p) {
  return sqrt(fmax2(2.0 * TINY_, gamma * p / rho));
}

/* Convert2Primitive/Conservative on an i,j box of a cell-centred vector */
static void c2p_box(const fvr *s, double *p, const double *c, int i1, int i2, int j1, int j2) {
  PAR_FOR
  for (int j = j1; j <= j2; j++)
    for (int i = i1; i <= i2; i++) cons2prim_pt(s, p, c, s->N, IDX(s, i, j));
}
static void p2c_box(const fvr *s, double *c, const double *p, int i1, int i2, int j1, int j2) {
  PAR_FOR
  for (int j = j1; j <= j2; j++)
    for (int i = i1; i <= i2; i++) prim2cons_pt(s, c, p, s->N, IDX(s, i, j));
}

/* Euler UpdateSoundSpeed on cell centres, physics_euler.f90:2555-2574 */
static void update_bccsound(fvr *s, const double *p) {
  for (int c = 0; c < s->N; c++) s->bccsound[c] = sound_speed(s->c.gamma, p[c], p[3 * s->N + c]);
}

/* FARGO frame change, physics_eulerisotherm.f90:1263-1303 / 1373-1413;
 * Euler physics_euler.f90:2003-2047 / 2159-2203: the energy first, from the
 * old momentum */
static void add_bg_velocity(fvr *s, sv *p, sv *c) {
  if (!c->flag) return;
  if (!p->flag) return; /* "should not happen" branch, kept as in the reference */
  const int N = s->N;
  PAR_FOR
  for (int j = 0; j < s->NJ; j++)
    for (int i = 0; i < s->NI; i++) {
      int k = IDX(s, i, j);
      const double w = s->w[i];
      if (s->V == 4) c->d[3 * N + k] = c->d[3 * N + k] + w * (c->d[2 * N + k] + 0.5 * c->d[k] * w);
      p->d[2 * N + k] = p->d[2 * N + k] + w;
      c->d[2 * N + k] = c->d[2 * N + k] + c->d[k] * w;
    }
  c->flag = 0;
  p->flag = 0;
}
static void sub_bg_velocity(fvr *s, sv *p, sv *c) {
  if (c->flag) return;
  if (p->flag) return;
  const int N = s->N;
  PAR_FOR
  for (int j = 0; j < s->NJ; j++)
    for (int i = 0; i < s->NI; i++) {
      int k = IDX(s, i, j);
      const double w = s->w[i];
      if (s->V == 4) c->d[3 * N + k] = c->d[3 * N + k] - w * (c->d[2 * N + k] - 0.5 * c->d[k] * w);
      p->d[2 * N + k] = p->d[2 * N + k] - w;
      c->d[2 * N + k] = c->d[2 * N + k] - c->d[k] * w;
    }
  c->flag = 1;
  p->flag = 1;
}

/* -------------------------------------------------------------- boundary */

/* absorbing boundary, isothermal: boundary_absorbing.f90:114-150,
 * physics_eulerisotherm.f90:1764-1849 (char. system), 2027-2085 (boundary
 * data), 2895-2969 (elementals).  side 0 = west, 1 = east. */
static void absorbing_x_iso(fvr *s, double *p, int side) {
  int N = s->N, imin = G, imax = G + s->c.inum - 1;
  int i1 = side ? imax : imin, i2n = side ? imax - 1 : imin + 1; /* char system */
  int iL = i1 < i2n ? i1 : i2n, iR = i1 < i2n ? i2n : i1;
  int ig = side ? imax + 1 : imin - 1, delta = ig - i1, fidx = side ? E_ : W_;
  for (int j = G; j < G + s->c.jnum; j++) {
    int k1 = IDX(s, i1, j), kL = IDX(s, iL, j), kR = IDX(s, iR, j), kg = IDX(s, ig, j);
    double cs = s->bccsound[k1], u1 = p[1 * N + k1];
    double lam[3] = {u1 - cs, u1, u1 + cs};
    double csf = s->fcsound[W_ * N + kR];
    double dlnrho = log(p[kR] / p[kL]);
    double ducs = (p[1 * N + kR] - p[1 * N + kL]) / csf;
    double xv[3] = {dlnrho - ducs, p[2 * N + kR] - p[2 * N + kL], dlnrho + ducs};
    for (int l = 0; l < 3; l++)
      if (side ? (lam[l] <= 0.0) : (lam[l] >= 0.0)) xv[l] = 0.0;
    {
      double cs2 = s->fcsound[fidx * N + k1];
      double rho1 = p[k1], v1 = p[2 * N + k1];
      p[kg] = rho1 * exp(delta * 0.5 * (xv[2] + xv[0]));
      p[1 * N + kg] = u1 + delta * 0.5 * cs2 * (xv[2] - xv[0]);
      p[2 * N + kg] = v1 + delta * xv[1];
    }
    for (int gi = 2; gi <= G; gi++) {
      int kk = IDX(s, side ? imax + gi : imin - gi, j);
      for (int l = 0; l < s->V; l++) p[l * N + kk] = p[l * N + kg];
    }
  }
}

/* absorbing boundary, Euler: boundary_absorbing.f90:114-150 with
 * physics_euler.f90:743-846 (CalculateCharSystemX: eigenvalues at i1 from
 * SetEigenValues1d/2d, characteristic variables from the cells iL, iR by
 * SetCharVars1d/2d) and 1072-1129 (CalculateBoundaryDataX with
 * SetBoundaryData1d/2d).  side 0 = west, 1 = east. */
static void absorbing_x_euler(fvr *s, double *p, int side) {
  int N = s->N, imin = G, imax = G + s->c.inum - 1;
  int i1 = side ? imax : imin, i2n = side ? imax - 1 : imin + 1; /* char system */
  int iL = i1 < i2n ? i1 : i2n, iR = i1 < i2n ? i2n : i1;
  int ig = side ? imax + 1 : imin - 1, delta = ig - i1;
  const double gam = s->c.gamma;
  for (int j = G; j < G + s->c.jnum; j++) {
    int k1 = IDX(s, i1, j), kL = IDX(s, iL, j), kR = IDX(s, iR, j), kg = IDX(s, ig, j);
    double rho1 = p[k1], u1 = p[1 * N + k1], v1 = p[2 * N + k1], P1 = p[3 * N + k1];
    double cs = sound_speed(gam, rho1, P1);
    double lam[4] = {u1 - cs, u1, u1, u1 + cs};
    double dlnP = log(p[3 * N + kR] / p[3 * N + kL]);
    double gammadu_cs = 2 * gam * (p[1 * N + kR] - p[1 * N + kL]) / (lam[3] - lam[0]);
    double xv[4] = {dlnP - gammadu_cs, dlnP - gam * log(p[kR] / p[kL]), p[2 * N + kR] - p[2 * N + kL],
                    dlnP + gammadu_cs};
    for (int l = 0; l < 4; l++)
      if (side ? (lam[l] <= 0.0) : (lam[l] >= 0.0)) xv[l] = 0.0;
    {
      double dlp = 0.5 * (xv[3] + xv[0]);
      double rho2 = rho1 * exp((double)delta / gam * (dlp - xv[1]));
      double P2 = P1 * exp(delta * dlp);
      double cs2 = sound_speed(gam, rho1 + rho2, P1 + P2);
      p[kg] = rho2;
      p[1 * N + kg] = u1 + 0.5 * delta * cs2 / gam * (xv[3] - xv[0]);
      p[2 * N + kg] = v1 + delta * xv[2];
      p[3 * N + kg] = P2;
    }
    for (int gi = 2; gi <= G; gi++) {
      int kk = IDX(s, side ? imax + gi : imin - gi, j);
      for (int l = 0; l < s->V; l++) p[l * N + kk] = p[l * N + kg];
    }
  }
}

/* boundary_nogradients.f90:76-129, boundary_periodic.f90:73-116 */
static void set_boundary(fvr *s, double *p, int dir) {
  int N = s->N, V = s->V, imin = G, imax = G + s->c.inum - 1, jmin = G, jmax = G + s->c.jnum - 1;
  int bc = s->c.bc[dir];
  if (bc == NO_GRADIENTS) {
    for (int g = 1; g <= G; g++) {
      if (dir == W_ || dir == E_) {
        int is = dir == W_ ? imin : imax, ig = dir == W_ ? imin - g : imax + g;
        for (int j = jmin; j <= jmax; j++)
          for (int l = 0; l < V; l++) p[l * N + IDX(s, ig, j)] = p[l * N + IDX(s, is, j)];
      } else {
        int js = dir == S_ ? jmin : jmax, jg = dir == S_ ? jmin - g : jmax + g;
        for (int i = imin; i <= imax; i++)
          for (int l = 0; l < V; l++) p[l * N + IDX(s, i, jg)] = p[l * N + IDX(s, i, js)];
      }
    }
  } else if (bc == PERIODIC) {
    for (int g = 1; g <= G; g++) {
      if (dir == W_ || dir == E_) {
        int is = dir == W_ ? imax - g + 1 : imin + g - 1, ig = dir == W_ ? imin - g : imax + g;
        for (int j = 0; j < s->NJ; j++)
          for (int l = 0; l < V; l++) p[l * N + IDX(s, ig, j)] = p[l * N + IDX(s, is, j)];
      } else {
        int js = dir == S_ ? jmax - g + 1 : jmin + g - 1, jg = dir == S_ ? jmin - g : jmax + g;
        for (int i = 0; i < s->NI; i++)
          for (int l = 0; l < V; l++) p[l * N + IDX(s, i, jg)] = p[l * N + IDX(s, i, js)];
      }
    }
  } else if (bc == ABSORBING) {
    /* the x-direction is what the configurations need */
    if (s->V == 4)
      absorbing_x_euler(s, p, dir == E_);
    else
      absorbing_x_iso(s, p, dir == E_);
  }
}

/* boundary_generic.f90:434-517 (2-D corners) */
static void set_corners(fvr *s, double *p) {
  int N = s->N, V = s->V, imin = G, imax = G + s->c.inum - 1, jmin = G, jmax = G + s->c.jnum - 1;
#define CP(ia, ja, ib, jb) \
  for (int l = 0; l < V; l++) p[l * N + IDX(s, ia, ja)] = p[l * N + IDX(s, ib, jb)]
#define AV(ia, ja, i1, j1, i2, j2)                                                             \
  for (int l = 0; l < V; l++)                                                                  \
  p[l * N + IDX(s, ia, ja)] = 0.5 * (p[l * N + IDX(s, i1, j1)] + p[l * N + IDX(s, i2, j2)])
  for (int j = 1; j <= G; j++) { /* south-west */
    for (int i = j + 1; i <= G; i++) {
      CP(imin - i, jmin - j, imin - i, jmin - j + 1);
      CP(imin - j, jmin - i, imin - j + 1, jmin - i);
    }
    AV(imin - j, jmin - j, imin - j, jmin, imin, jmin - j);
  }
  for (int j = 1; j <= G; j++) { /* south-east */
    for (int i = j + 1; i <= G; i++) {
      CP(imax + i, jmin - j, imax + i, jmin - j + 1);
      CP(imax + j, jmin - i, imax + j - 1, jmin - i);
    }
    AV(imax + j, jmin - j, imax + j, jmin, imax, jmin - j);
  }
  for (int j = 1; j <= G; j++) { /* north-west */
    for (int i = j + 1; i <= G; i++) {
      CP(imin - i, jmax + j, imin - i, jmax + j - 1);
      CP(imin - j, jmax + i, imin - j + 1, jmax + i);
    }
    AV(imin - j, jmax + j, imin - j, jmax, imin, jmax + j);
  }
  for (int j = 1; j <= G; j++) { /* north-east */
    for (int i = j + 1; i <= G; i++) {
      CP(imax + i, jmax + j, imax + i, jmax + j - 1);
      CP(imax + j, jmax + i, imax + j - 1, jmax + i);
    }
    AV(imax + j, jmax + j, imax + j, jmax, imax, jmax + j);
  }
#undef CP
#undef AV
}

/* boundary_generic.f90:287-418 (serial, 2-D) */
static void center_boundary(fvr *s, sv *p, sv *c) {
  int imin = G, imax = G + s->c.inum - 1, jmin = G, jmax = G + s->c.jnum - 1;
  c2p_box(s, p->d, c->d, imin, imax, jmin, jmax);
  p->flag = c->flag;
  set_boundary(s, p->d, W_);
  set_boundary(s, p->d, E_);
  set_boundary(s, p->d, S_);
  set_boundary(s, p->d, N_);
  set_corners(s, p->d);
  p2c_box(s, c->d, p->d, 0, imin - 1, 0, s->NJ - 1);
  p2c_box(s, c->d, p->d, imax + 1, s->NI - 1, 0, s->NJ - 1);
  p2c_box(s, c->d, p->d, imin, imax, 0, jmin - 1);
  p2c_box(s, c->d, p->d, imin, imax, jmax + 1, s->NJ - 1);
}

/* -------------------------------------------------------------- sources */

/* physics_eulerisotherm.f90:689-832, physics_euler.f90:1706-1846 (x-y plane),
 * elementals physics_eulerisotherm.f90:3069-3088 */
static void geometrical_sources(fvr *s, const double *p, const double *c, double *st) {
  int N = s->N, V = s->V;
  if (s->c.geometry == CARTESIAN) return;
  PAR_FOR
  for (int k = 0; k < N; k++) {
    double vx = p[1 * N + k], vy = p[2 * N + k], vz = 0.0, mx = c[1 * N + k], my = c[2 * N + k],
           mz = 0.0;
    double P = (V == 4) ? p[3 * N + k] : p[k] * (s->bccsound[k] * s->bccsound[k]);
    st[k] = 0.0;
    if (V == 4) st[3 * N + k] = 0.0;
    /* GetGeometricalSourceX(cxyx,cxzx,cyxy,czxz,vx,vy,vz,P,my,mz), cxzx = 0 */
    st[1 * N + k] = my * (s->cyxy[k] * vy - s->cxyx[k] * vx) + mz * (s->czxz[k] * vz - 0.0 * vx) +
                    (s->cyxy[k] + s->czxz[k]) * P;
    /* GetGeometricalSourceY(cxyx,cyxy,cyzy,czyz,vx,vy,vz,P,mx,mz), cyzy = 0 */
    st[2 * N + k] = mz * (s->czyz[k] * vz - 0.0 * vy) + mx * (s->cxyx[k] * vx - s->cyxy[k] * vy) +
                    (s->cxyx[k] + s->czyz[k]) * P;
  }
  PAR_FOR
  for (int j = 0; j < s->NJ; j++)
    for (int i = 0; i < s->NI; i++)
      if (i < G || i >= G + s->c.inum || j < G || j >= G + s->c.jnum)
        for (int l = 0; l < V; l++) st[l * N + IDX(s, i, j)] = 0.0;
}

/* Fluxes%GetBoundaryFlux (fluxes_base.f90:220-300, serial build) of field l:
 * SUM over the boundary's interior range in index order */
static double boundary_flux_sum(const fvr *s, int direction, int l) {
  const int xd = direction <= 2, side = (direction - 1) & 1;
  const double *a = xd ? s->bx + (size_t)(2 * l + side) * s->NJ : s->by + (size_t)(2 * l + side) * s->NI;
  const int hi = G + (xd ? s->c.jnum : s->c.inum) - 1;
  double acc = 0.0;
  for (int k = G; k <= hi; k++) acc = acc + a[k];
  return acc;
}

/* UpdateGravity_single with accretion (gravity_pointmass.f90:383-428; the
 * switch-on phase is off): the mass that crossed the outflow boundary since
 * the last evaluation is added to the point mass and the acceleration is
 * rescaled by new/old mass.  s->accel holds sources_gravity's sum
 * 0 + accel_pointmass (sources_gravity.f90:213-226); scaling it gives the same
 * values as scaling the point mass's own array and adding it to zero. */
static void update_gravity(fvr *s, double dt) {
  if (!s->c.gravity || !s->c.outbound) return;
  const double bflux = boundary_flux_sum(s, s->c.outbound, 0);
  const double oldmass = s->pm_mass;
  double dmass = bflux - s->pm_mdot;
  if (s->c.acclimit >= 0.) {
    const double dmasslim = fmin2(dmass, s->c.acclimit * dt * s->pm_mass);
    s->pm_massloss = s->pm_massloss + dmass - dmasslim;
    dmass = dmasslim;
  }
  s->pm_mass = s->pm_mass + dmass;
  if (dt > 0.0) s->pm_accrate = dmass / dt;
  const double massfac = s->pm_mass / oldmass;
  PAR_FOR
  for (int k = 0; k < s->N; k++) {
    s->accel[0][k] = s->accel[0][k] * massfac;
    s->accel[1][k] = s->accel[1][k] * massfac;
  }
  s->pm_mdot = bflux;
}

/* sources_generic.f90:220-261: sterm = 0, then every source of the list in
 * order (gravity first, sources_generic.f90:133-199), each through
 * Physics%ExternalSources (physics_eulerisotherm.f90:1201-1226,
 * physics_euler.f90:1849-1885):
 *  - point-mass gravity, sources_gravity.f90:164-199;
 *  - the rotating frame, sources_rotframe.f90:212-247 (2-D acceleration:
 *    caccel + twoOmega*v_y, caccel - twoOmega*v_x; with disable_centaccel
 *    twoOmega*v_y, -twoOmega*v_x) with the real-frame velocities p;
 *  - the viscosity, sources_viscosity.f90:481-500 (configured before the
 *    rotating frame; the list holds the configuration in reverse order);
 * dt < 0: no UpdateGravity (test aid); t: ComputeRHS's time argument */
static void source_term(const fvr *s, const double *c, const double *ax, const double *ay, int N, int V) {
  PAR_FOR
  for (int k = 0; k < N; k++) {
    s->sterm[k] = 0.0;
    s->sterm[1 * N + k] = c[k] * ax[k];
    s->sterm[2 * N + k] = c[k] * ay[k];
    if (V == 4) s->sterm[3 * N + k] = c[1 * N + k] * ax[k] + c[2 * N + k] * ay[k];
  }
}

/* sources_viscosity.f90:356-460 (UpdateViscosity): the dynamic viscosity
 * is recomputed only when the time differs from the last update's, or is 0;
 * ALPHA: eta = alpha*rho*cs^2/|e_phi/r . v + Omega| on every cell, bulk
 * viscosity -2/3 eta; MOLECULAR keeps its constants */
static void update_viscosity(fvr *s, double time, const double *p) {
  const int N = s->N;
  if (!(time != s->vis_time || time == 0.0)) return;
  if (s->c.vismodel == 2) {
    for (int k = 0; k < N; k++) {
      double om = s->ephir[0][k] * p[1 * N + k];
      om = om + s->ephir[1][k] * p[2 * N + k];
      om = om + s->c.omega;
      const double cs = s->bccsound[k];
      s->dynvis[k] = s->c.dynconst * p[k] * (cs * cs) / fabs(om);
      s->bulkvis[k] = -(2. / 3.) * s->dynvis[k];
    }
  } else if (s->c.vismodel == 3) { /* BETA: beta*rho*r^2*|Omega| (sources_viscosity.f90:388-394, 457-466) */
    for (int k = 0; k < N; k++) {
      double om = s->ephir[0][k] * p[1 * N + k];
      om = om + s->ephir[1][k] * p[2 * N + k];
      om = om + s->c.omega;
      s->dynvis[k] = s->c.dynconst * p[k] * (s->rad[k] * s->rad[k]) * fabs(om);
      s->bulkvis[k] = -(2. / 3.) * s->dynvis[k];
    }
  }
  s->vis_time = time;
}

/* mesh_midpoint.f90:586-630 (Divergence3D) of a 2-D x-y vector field on
 * face-averaged components (VectorDivergence2D_1, 156-232) */
static double div2d(const fvr *s, int k, const double *vx, const double *vy) {
  const int NI = s->NI;
  const double xw = 0.5 * (vx[k - 1] + vx[k]), xe = 0.5 * (vx[k + 1] + vx[k]);
  const double ys = 0.5 * (vy[k - NI] + vy[k]), yn = 0.5 * (vy[k + NI] + vy[k]);
  return s->dydzdV[k] * (s->dAx[1][k] * xe - s->dAx[0][k] * xw) + s->dzdxdV[k] * (s->dAy[1][k] * yn - s->dAy[0][k] * ys) +
         0.0 * (0.0 * 0.0 - 0.0 * 0.0) + 0.0 * 0.0 + 0.0 * 0.0 - 0.0 * 0.0 - 0.0 * 0.0;
}

/* physics_eulerisotherm.f90:1017-1072 (CalcStresses, 2-D x-y) and 982-1016
 * (ViscositySources: TensorDivergence2D_1, mesh_midpoint.f90:279-335, with
 * btyx = btxy), physics_euler.f90:1887-1918 (energy: divergence of T.v);
 * result in s->vsterm */
static void viscosity_sources(fvr *s, const double *p) {
  const int N = s->N, NI = s->NI, NJ = s->NJ, V = s->V;
  const double *vx = p + N, *vy = p + 2 * N;
  for (int j = 1; j < NJ - 1; j++)
    for (int i = 1; i < NI - 1; i++) s->vtmp[IDX(s, i, j)] = div2d(s, IDX(s, i, j), vx, vy);
  for (int k = 0; k < N; k++) s->vtmp[k] = s->bulkvis[k] * s->vtmp[k];
  for (int j = G - 1; j <= G + s->c.jnum; j++)
    for (int i = G - 1; i <= G + s->c.inum; i++) {
      const int k = IDX(s, i, j);
      const double eta = s->dynvis[k];
      s->btxx[k] = eta * ((vx[k + 1] - vx[k - 1]) / s->dlx[k] + 2.0 * s->cxyx[k] * vy[k]) + s->vtmp[k];
      s->btyy[k] = eta * ((vy[k + NI] - vy[k - NI]) / s->dly[k] + 2.0 * s->cyxy[k] * vx[k]) + s->vtmp[k];
      s->btxy[k] = eta * (0.5 * ((vy[k + 1] - vy[k - 1]) / s->dlx[k] + (vx[k + NI] - vx[k - NI]) / s->dly[k]) -
                          s->cxyx[k] * vx[k] - s->cyxy[k] * vy[k]);
    }
  double *st = s->vsterm;
  for (int k = 0; k < N; k++) st[k] = 0.0;
  for (int j = 1; j < NJ - 1; j++)
    for (int i = 1; i < NI - 1; i++) {
      const int k = IDX(s, i, j);
      const double *txx = s->btxx, *txy = s->btxy, *tyy = s->btyy;
      st[N + k] = s->dydzdV[k] * (s->dAx[1][k] * (0.5 * (txx[k + 1] + txx[k])) -
                                  s->dAx[0][k] * (0.5 * (txx[k - 1] + txx[k]))) +
                  s->dzdxdV[k] * (s->dAy[1][k] * (0.5 * (txy[k + NI] + txy[k])) -
                                  s->dAy[0][k] * (0.5 * (txy[k - NI] + txy[k]))) +
                  0.0 * (0.0 * 0.0 - 0.0 * 0.0) + s->cxyx[k] * txy[k] + 0.0 * 0.0 - s->cyxy[k] * tyy[k] - 0.0 * 0.0;
      st[2 * N + k] = s->dydzdV[k] * (s->dAx[1][k] * (0.5 * (txy[k + 1] + txy[k])) -
                                      s->dAx[0][k] * (0.5 * (txy[k - 1] + txy[k]))) +
                      s->dzdxdV[k] * (s->dAy[1][k] * (0.5 * (tyy[k + NI] + tyy[k])) -
                                      s->dAy[0][k] * (0.5 * (tyy[k - NI] + tyy[k]))) +
                      0.0 * (0.0 * 0.0 - 0.0 * 0.0) + (-s->cxyx[k]) * txx[k] + 0.0 * 0.0 - (-s->cyxy[k]) * txy[k] -
                      0.0 * 0.0;
    }
  if (V == 4) {
    for (int k = 0; k < N; k++) {
      s->vtmp[k] = vx[k] * s->btxx[k] + vy[k] * s->btxy[k];
      s->vtmp1[k] = vx[k] * s->btxy[k] + vy[k] * s->btyy[k];
    }
    for (int j = 1; j < NJ - 1; j++)
      for (int i = 1; i < NI - 1; i++) st[3 * N + IDX(s, i, j)] = div2d(s, IDX(s, i, j), s->vtmp, s->vtmp1);
  }
}

static void external_sources(fvr *s, const double *p, const double *c, double *st, double dt, double time) {
  int N = s->N, V = s->V;
  if (!s->c.gravity && !s->c.rotframe && !s->c.vismodel) return;
  PAR_FOR
  for (int k = 0; k < V * N; k++) st[k] = 0.0;
  if (s->c.gravity) {
    if (dt >= 0.0) update_gravity(s, dt);
    source_term(s, c, s->accel[0], s->accel[1], N, V);
    PAR_FOR
    for (int k = 0; k < V * N; k++) st[k] = st[k] + s->sterm[k];
  }
  if (s->c.rotframe) {
    double *ax = s->raccel[0], *ay = s->raccel[1];
    PAR_FOR
    for (int k = 0; k < N; k++) {
      const double tw = s->twoomega[k], vx = p[1 * N + k], vy = p[2 * N + k];
      if (s->c.rotframe == 2) {
        ax[k] = tw * vy;
        ay[k] = -(tw * vx);
      } else {
        ax[k] = s->caccel[0][k] + tw * vy;
        ay[k] = s->caccel[1][k] - tw * vx;
      }
    }
    source_term(s, c, ax, ay, N, V);
    PAR_FOR
    for (int k = 0; k < V * N; k++) st[k] = st[k] + s->sterm[k];
  }
  if (s->c.vismodel) {  /* sources_viscosity.f90:481-500 */
    update_viscosity(s, time, p);
    viscosity_sources(s, p);
    for (int k = 0; k < V * N; k++) st[k] = st[k] + s->vsterm[k];
  }
  PAR_FOR
  for (int j = 0; j < s->NJ; j++)
    for (int i = 0; i < s->NI; i++)
      if (i < G || i >= G + s->c.inum || j < G || j >= G + s->c.jnum)
        for (int l = 0; l < V; l++) st[l * N + IDX(s, i, j)] = 0.0;
}

/* physics_eulerisotherm.f90:1531-1586 (2-D x-y momentum); Euler
 * physics_euler.f90:2368-2400: the same momentum source, then the energy
 * source -v_y*(that + GetGeometricalSourceY with vz = P = mz = 0), the
 * elemental of physics_eulerisotherm.f90:3080-3088; p is in the real frame */
static void add_fargo_sources(fvr *s, const double *p, const double *c, double *st) {
  int N = s->N;
  PAR_FOR
  for (int j = G; j < G + s->c.jnum; j++)
    for (int i = G; i < G + s->c.inum; i++) {
      int k = IDX(s, i, j);
      double tmp = c[1 * N + k] * 0.5 * (s->w[i + 1] - s->w[i - 1]) / s->dlx[k];
      st[2 * N + k] = st[2 * N + k] - tmp;
      if (s->V == 4) {
        const double vx = p[1 * N + k], vy = p[2 * N + k], mx = c[1 * N + k];
        tmp = tmp + (0.0 * (s->czyz[k] * 0.0 - 0.0 * vy) + mx * (s->cxyx[k] * vx - s->cyxy[k] * vy) +
                     (s->cxyx[k] + s->czyz[k]) * 0.0);
        st[3 * N + k] = st[3 * N + k] - vy * tmp;
      }
    }
}

/* ---------------------------------------------------------------- fluxes */

/* reconstruction_linear.f90:645-733 */
static double minmod2(double a, double b) {
  if (fsign(1.0, a) * fsign(1.0, b) > 0) return fsign(fmin2(fabs(a), fabs(b)), a);
  return 0.;
}
static double minmod3(double a, double b, double c) {
  if ((fsign(1.0, a) * fsign(1.0, b) > 0) && (fsign(1.0, b) * fsign(1.0, c) > 0))
    return fsign(fmin2(fmin2(fabs(a), fabs(b)), fabs(c)), a);
  return 0.;
}
static double sweby(double a, double b, double q) {
  if (fsign(1.0, a) * fsign(1.0, b) > 0)
    return fsign(fmax2(fmin2(q * fabs(a), fabs(b)), fmin2(fabs(a), q * fabs(b))), a);
  return 0.;
}
static double ospre(double a, double b) {
  return 1.5 * a * b * (a + b) / (a * (a + 0.5 * b) + b * (b + 0.5 * a) + TINY_);
}
static double vanleer(double a, double b) {
  double a1 = fabs(a), a2 = fabs(b);
  return (a * a2 + b * a1) / (a1 + a2 + TINY_);
}
static double slope_lim(const fvr *s, double um, double u0, double up) {
  double th = s->c.theta;
  switch (s->c.limiter) {
    case MINMOD: return minmod2(u0 - um, up - u0);
    case MONOCENT: return minmod3(th * (u0 - um), th * (up - u0), 0.5 * (up - um));
    case SWEBY: return sweby(u0 - um, up - u0, th);
    case SUPERBEE: return sweby(u0 - um, up - u0, 2.0);
    case OSPRE: return ospre(u0 - um, up - u0);
    case VANLEER: return vanleer(u0 - um, up - u0);
    default: return 0.5 * ((u0 - um) + (up - u0));
  }
}

/* reconstruction_linear.f90:215-240 (states) and 243-630 (slopes) */
static void calculate_states(fvr *s, const double *rv, double *rs) {
  int N = s->N, NI = s->NI, V = s->V;
  for (int l = 0; l < V; l++) {
    const double *u = rv + (size_t)l * N;
    double *sx = s->slopes + ((size_t)l * 2 + 0) * N, *sy = s->slopes + ((size_t)l * 2 + 1) * N;
    PAR_FOR
    for (int j = 0; j < s->NJ; j++)
      for (int i = 1; i < NI - 1; i++) {
        int k = IDX(s, i, j);
        sx[k] = s->invdx * slope_lim(s, u[k - 1], u[k], u[k + 1]);
      }
    PAR_FOR
    for (int k = NI; k < N - NI; k++) sy[k] = s->invdy * slope_lim(s, u[k - NI], u[k], u[k + NI]);
  }
  for (int l = 0; l < V; l++)
    for (int n = 0; n < 4; n++)
      PAR_FOR
      for (int k = 0; k < N; k++)
        rs[((size_t)l * 4 + n) * N + k] =
            rv[(size_t)l * N + k] + s->slopes[((size_t)l * 2 + n / 2) * N + k] * s->rdx[n][k];
}

/* physics_eulerisotherm.f90:555-686 with SetWaveSpeeds 2742-2750 */
static void wave_speeds_faces(fvr *s) {
  int N = s->N, NI = s->NI;
  for (int d = 0; d < 2; d++) {
    const double *vW = s->prim + ((size_t)(1 + d) * 4 + 2 * d) * N;
    const double *vE = s->prim + ((size_t)(1 + d) * 4 + 2 * d + 1) * N;
    const double *cW = s->fcsound + (size_t)(2 * d) * N, *cE = s->fcsound + (size_t)(2 * d + 1) * N;
    double *mn = s->minwav + (size_t)d * N, *mx = s->maxwav + (size_t)d * N;
    PAR_FOR
    for (int k = 0; k < N; k++) {
      s->tmp[k] = fmin2(0., vW[k] - cW[k]);
      s->tmp1[k] = fmax2(0., vW[k] + cW[k]);
      mn[k] = fmin2(0., vE[k] - cE[k]);
      mx[k] = fmax2(0., vE[k] + cE[k]);
    }
    if (d == 0) {
      PAR_FOR
      for (int j = 0; j < s->NJ; j++)
        for (int i = G - 1; i <= G + s->c.inum - 1; i++) {
          int k = IDX(s, i, j);
          mn[k] = fmin2(fmin2(0.0, s->tmp[k + 1]), mn[k]);
          mx[k] = fmax2(fmax2(0.0, s->tmp1[k + 1]), mx[k]);
        }
    } else {
      PAR_FOR
      for (int j = G - 1; j <= G + s->c.jnum - 1; j++)
        for (int i = 0; i < NI; i++) {
          int k = IDX(s, i, j);
          mn[k] = fmin2(fmin2(0.0, s->tmp[k + NI]), mn[k]);
          mx[k] = fmax2(fmax2(0.0, s->tmp1[k + NI]), mx[k]);
        }
    }
  }
}

/* physical fluxes on the two faces of direction d:
 * iso physics_eulerisotherm.f90:835-930 / 2787-2817,
 * Euler physics_euler.f90:492-600 / 2999-3030 */
static void physical_fluxes(fvr *s, int d) {
  int N = s->N, V = s->V;
  int un = 1 + d, ut = 2 - d; /* normal / tangential momentum index */
  for (int f = 2 * d; f <= 2 * d + 1; f++)
    PAR_FOR
    for (int k = 0; k < N; k++) {
#define PF(l) s->prim[((size_t)(l) * 4 + f) * N + k]
#define CF(l) s->cons[((size_t)(l) * 4 + f) * N + k]
#define FF(l) s->pfluxes[((size_t)(l) * 4 + f) * N + k]
      double rho = PF(0), u = PF(un), mu = CF(un), mv = CF(ut);
      FF(0) = rho * u;
      if (V == 3) {
        double cs = s->fcsound[(size_t)f * N + k];
        FF(un) = mu * u + rho * cs * cs;
      } else {
        double P = PF(3), E = CF(3);
        FF(un) = mu * u + P;
        FF(3) = (E + P) * u;
      }
      FF(ut) = mv * u;
#undef PF
#undef CF
#undef FF
    }
}

/* fluxes_kt.f90:109-208 */
static void calculate_fluxes(fvr *s, const double *pv, const double *cv) {
  int N = s->N, NI = s->NI, V = s->V;
  /* fluxes_base.f90:368-397 */
  if (s->c.recvars == PRIMITIVE) {
    calculate_states(s, pv, s->prim);
    for (int f = 0; f < 4; f++)
      PAR_FOR
      for (int k = 0; k < N; k++) prim2cons_pt(s, s->cons + (size_t)f * N, s->prim + (size_t)f * N, 4 * (size_t)N, k);
  } else {
    calculate_states(s, cv, s->cons);
    for (int f = 0; f < 4; f++)
      PAR_FOR
      for (int k = 0; k < N; k++) cons2prim_pt(s, s->prim + (size_t)f * N, s->cons + (size_t)f * N, 4 * (size_t)N, k);
  }
  if (V == 4)
    for (int f = 0; f < 4; f++)
      PAR_FOR
      for (int k = 0; k < N; k++)
        s->fcsound[(size_t)f * N + k] =
            sound_speed(s->c.gamma, s->prim[(size_t)f * N + k], s->prim[((size_t)3 * 4 + f) * N + k]);
  wave_speeds_faces(s);
  physical_fluxes(s, 0);
  for (int l = 0; l < V; l++)
    PAR_FOR
    for (int j = 0; j < s->NJ; j++)
      for (int i = G - 1; i <= G + s->c.inum - 1; i++) {
        int k = IDX(s, i, j);
        double amax = s->maxwav[k], amin = s->minwav[k];
        const double *pf = s->pfluxes + (size_t)l * 4 * N, *cs = s->cons + (size_t)l * 4 * N;
        s->xflux[(size_t)l * N + k] =
            s->dAx[0][k + 1] / (amax - amin) *
            (amax * pf[1 * (size_t)N + k] - amin * pf[0 * (size_t)N + k + 1] +
             amin * amax * (cs[0 * (size_t)N + k + 1] - cs[1 * (size_t)N + k]));
      }
  physical_fluxes(s, 1);
  for (int l = 0; l < V; l++)
    PAR_FOR
    for (int j = G - 1; j <= G + s->c.jnum - 1; j++)
      for (int i = 0; i < NI; i++) {
        int k = IDX(s, i, j);
        double amax = s->maxwav[N + k], amin = s->minwav[N + k];
        const double *pf = s->pfluxes + (size_t)l * 4 * N, *cs = s->cons + (size_t)l * 4 * N;
        s->yflux[(size_t)l * N + k] =
            s->dAy[0][k + NI] / (amax - amin) *
            (amax * pf[3 * (size_t)N + k] - amin * pf[2 * (size_t)N + k + NI] +
             amin * amax * (cs[2 * (size_t)N + k + NI] - cs[3 * (size_t)N + k]));
      }
}

/* ------------------------------------------------------------- timedisc */

/* timedisc_base.f90:1666-1742 */
static void check_data(fvr *s, const double *p, const double *c, int bm) {
  int N = s->N, V = s->V;
  if (bm & CHECK_CSOUND) {
    double mn = HUGE_;
    for (int k = 0; k < N; k++) mn = fmin2(mn, s->bccsound[k]);
    if (mn <= 0.) s->brk = 1;
  }
  if ((bm & CHECK_PMIN) && V == 4) {
    double mn = HUGE_;
    for (int j = G; j < G + s->c.jnum; j++)
      for (int i = G; i < G + s->c.inum; i++) mn = fmin2(mn, p[3 * N + IDX(s, i, j)]);
    if (mn < s->pmin) s->brk = 1;
  }
  if (bm & CHECK_RHOMIN) {
    double mn = HUGE_;
    for (int j = G; j < G + s->c.jnum; j++)
      for (int i = G; i < G + s->c.inum; i++) mn = fmin2(mn, p[IDX(s, i, j)]);
    if (mn < s->rhomin) s->brk = 1;
  }
  if (bm & CHECK_INVALID) {
    for (int k = 0; k < V * N; k++)
      if (c[k] != c[k] || p[k] != p[k] || c[k] > HUGE_ || p[k] > HUGE_) s->brk = 1;
  }
}

/* timedisc_base.f90:1073-1271 (rhstype 0, 2-D) */
/* dt: ComputeRHS's own dt argument (timedisc_base.f90:1073), read by the
 * point-mass accretion update */
static void compute_rhs(fvr *s, sv *p, sv *c, int checkbm, double *rhs, double dt, double t) {
  int N = s->N, V = s->V, imin = G, imax = G + s->c.inum - 1, jmin = G, jmax = G + s->c.jnum - 1;
  int fargo = s->c.fargo == 1 || s->c.fargo == 2;
  if (fargo) add_bg_velocity(s, p, c);
  center_boundary(s, p, c);
  if ((checkbm & CHECK_PMIN) && V == 4) {
    PAR_FOR
    for (int k = 0; k < N; k++) p->d[3 * N + k] = fmax2(p->d[3 * N + k], s->pmin);
    PAR_FOR
    for (int k = 0; k < N; k++) prim2cons_pt(s, c->d, p->d, N, k);
    c->flag = p->flag;
  }
  if (V == 4) update_bccsound(s, s->pvar.d);
  if (checkbm != CHECK_NOTHING) check_data(s, p->d, c->d, checkbm);
  geometrical_sources(s, p->d, c->d, s->geo_src);
  external_sources(s, p->d, c->d, s->src, dt, t);
  if (fargo) {
    if (s->c.geometry == CARTESIAN) memset(s->geo_src, 0, sizeof(double) * V * N);
    add_fargo_sources(s, p->d, c->d, s->geo_src);
    sub_bg_velocity(s, p, c);
  }
  calculate_fluxes(s, p->d, c->d);
  for (int l = 0; l < V; l++) {
    double *r = rhs + (size_t)l * N, *xf = s->xflux + (size_t)l * N, *yf = s->yflux + (size_t)l * N;
    for (int j = jmin; j <= jmax; j++) {
      r[IDX(s, imin - 1, j)] = s->dy * s->dz * xf[IDX(s, imin - 1, j)];
      r[IDX(s, imax + 1, j)] = -s->dy * s->dz * xf[IDX(s, imax, j)];
    }
    for (int i = imin; i <= imax; i++) {
      r[IDX(s, i, jmin - 1)] = s->dz * s->dx * yf[IDX(s, i, jmin - 1)];
      r[IDX(s, i, jmax + 1)] = -s->dz * s->dx * yf[IDX(s, i, jmax)];
    }
  }
  for (int l = 0; l < V; l++) {
    double *r = rhs + (size_t)l * N, *xf = s->xflux + (size_t)l * N, *yf = s->yflux + (size_t)l * N;
    const double *gs = s->geo_src + (size_t)l * N, *sr = s->src + (size_t)l * N;
    PAR_FOR
    for (int j = jmin; j <= jmax; j++)
      for (int i = imin; i <= imax; i++) {
        int k = IDX(s, i, j);
        r[k] = s->dydzdV[k] * (xf[k] - xf[k - 1]) + s->dzdxdV[k] * (yf[k] - yf[k - s->NI]) +
               s->dxdydV[k] * (0.0 - 0.0) - gs[k] - sr[k];
      }
  }
}

/* timedisc_base.f90:818-909 (2-D) with CalcWaveSpeeds_center,
 * physics_eulerisotherm.f90:530-552 */
static double calc_timestep(fvr *s) {
  int N = s->N;
  double invdt = -HUGE_;
  const double *p = s->pvar.d;
  for (int k = 0; k < N; k++) {
    double cs = s->bccsound[k], u = p[1 * N + k], v = p[2 * N + k];
    double mnx = fmin2(0., u - cs), mxx = fmax2(0., u + cs);
    double mny = fmin2(0., v - cs), mxy = fmax2(0., v + cs);
    s->minwav[k] = mnx; s->maxwav[k] = mxx; s->minwav[N + k] = mny; s->maxwav[N + k] = mxy;
    double x = fmax2(mxx, -mnx) / s->dlx[k] + fmax2(mxy, -mny) / s->dly[k];
    if (x > invdt) invdt = x;
  }
  s->dtcause = 0;
  double dt = s->c.cfl / invdt;
  /* the source list's time steps (sources_generic.f90:264-293): gravity and
   * the rotating frame give HUGE (sources_c_accel.f90:125); the viscosity
   * its advection-diffusion limit over every cell (sources_viscosity.f90:
   * 507-577, the 2-D x-y branch) with the kinematic viscosity of
   * Timedisc%pvar, after an UpdateViscosity at the current time */
  if (s->c.vismodel) {
    update_viscosity(s, s->time, p);
    double iv = -HUGE_;
    for (int k = 0; k < N; k++) {
      const double kin = s->dynvis[k] / p[k];
      const double x = (fmax2(s->maxwav[k], -s->minwav[k]) + 2 * kin / s->dlx[k]) / s->dlx[k] +
                       (fmax2(s->maxwav[N + k], -s->minwav[N + k]) + 2 * kin / s->dly[k]) / s->dly[k];
      if (x > iv) iv = x;
    }
    const double dtv = iv > TINY_ ? s->c.cvis / iv : HUGE_;
    if (dtv < dt) s->dtcause = 3; /* VISCOSITY (sources_generic.f90:74) */
    dt = fmin2(dt, dtv);
  }
  return dt;
}

/* timedisc_rkfehlberg.f90:277-302 */
static void compute_cvar(fvr *s, int m, double dt) {
  size_t n = (size_t)s->V * s->N;
  const double *a = s->a + (size_t)m * s->m;
  PAR_FOR
  for (size_t k = 0; k < n; k++) s->ctmp.d[k] = s->cvar.d[k] - a[0] * dt * s->coeff[0][k];
  for (int mm = 1; mm < m; mm++)
    PAR_FOR
    for (size_t k = 0; k < n; k++) s->ctmp.d[k] = s->ctmp.d[k] - a[mm] * dt * s->coeff[mm][k];
}

/* timedisc_base.f90:1012-1053 */
static double compute_error(fvr *s) {
  int N = s->N;
  double maxerr = 0.0;
  for (int l = 0; l < s->V; l++) {
    double e = -HUGE_;
    PAR_FOR
    for (int k = 0; k < N; k++) {
      double yh = s->cvar.d[(size_t)l * N + k], yl = s->ctmp.d[(size_t)l * N + k];
      s->cerr[(size_t)l * N + k] = fabs(yh - yl) / (s->c.tol_rel * fabs(yh) + s->c.tol_abs[l]);
    }
    for (int j = G; j < G + s->c.jnum; j++)
      for (int i = G; i < G + s->c.inum; i++) e = fmax2(e, s->cerr[(size_t)l * N + IDX(s, i, j)]);
    maxerr = (l == 0) ? e : fmax2(maxerr, e);
  }
  return maxerr;
}

/* timedisc_base.f90:759-814 */
static double adjust_timestep(fvr *s, double maxerr, double dtold) {
  double dttmp, dtnew;
  if (s->c.tol_rel >= 1.0) return HUGE_;
  if (maxerr > 0.)
    dttmp = 0.9 * dtold * exp(-log(maxerr) / s->c.order);
  else
    dttmp = 4. * dtold;
  if (s->maxerrold > 0.)
    dtnew = dttmp * exp(-s->c.beta * log(s->maxerrold / maxerr));
  else
    dtnew = dttmp;
  if (maxerr < 1.0) {
    dtnew = fmin2(dtnew, 4. * dtold);
    if (dtnew < dtold) s->dtcause = DTCAUSE_SMALLERR;
  } else {
    if (dtnew > dtold) dtnew = dttmp;
    dtnew = fmax2(dtnew, 0.25 * dtold);
  }
  s->maxerrold = maxerr;
  return dtnew;
}

/* boundary-flux accumulation of one SolveODE (timedisc_rkfehlberg.f90:216-259):
 * the boundary rows of every stage's rhs (compute_rhs stores the boundary
 * fluxes there, timedisc_base.f90:1224-1252), weighted by -dt*b_high(m) */
static void accumulate_bflux(fvr *s, double dt) {
  const int N = s->N, NI = s->NI, NJ = s->NJ;
  const int imin = G, imax = G + s->c.inum - 1, jmin = G, jmax = G + s->c.jnum - 1;
  for (int m = 0; m < s->m; m++)
    for (int l = 0; l < s->V; l++) {
      const double *k = s->coeff[m] + (size_t)l * N;
      double *bw = s->bx + (size_t)(2 * l) * NJ, *be = bw + NJ;
      double *bs = s->by + (size_t)(2 * l) * NI, *bn = bs + NI;
      for (int j = jmin; j <= jmax; j++) {
        bw[j] = bw[j] - dt * s->b_high[m] * k[IDX(s, imin - 1, j)];
        be[j] = be[j] - dt * s->b_high[m] * k[IDX(s, imax + 1, j)];
      }
      for (int i = imin; i <= imax; i++) {
        bs[i] = bs[i] - dt * s->b_high[m] * k[IDX(s, i, jmin - 1)];
        bn[i] = bn[i] - dt * s->b_high[m] * k[IDX(s, i, jmax + 1)];
      }
    }
}

/* timedisc_rkfehlberg.f90:170-268 */
static void solve_ode(fvr *s, double time, double *dt, double *err) {
  size_t n = (size_t)s->V * s->N;
  for (int m = 1; m < s->m; m++) {
    compute_cvar(s, m, *dt);
    compute_rhs(s, &s->ptmp, &s->ctmp, CHECK_NOTHING, s->coeff[m], *dt, time + s->cc[m] * *dt);
  }
  memcpy(s->ctmp.d, s->cvar.d, sizeof(double) * n);
  s->ctmp.flag = s->cvar.flag;
  for (int m = 0; m < s->m; m++)
    PAR_FOR
    for (size_t k = 0; k < n; k++) {
      s->ctmp.d[k] = s->ctmp.d[k] - s->b_low[m] * *dt * s->coeff[m][k];
      s->cvar.d[k] = s->cvar.d[k] - s->b_high[m] * *dt * s->coeff[m][k];
    }
  accumulate_bflux(s, *dt);
  if (s->c.tol_rel <= 1.0) {
    *err = compute_error(s);
    *dt = adjust_timestep(s, *err, *dt);
  }
}

static void copy_sv(fvr *s, sv *dst, const sv *src) {
  memcpy(dst->d, src->d, sizeof(double) * s->V * s->N);
  dst->flag = src->flag;
}

/* FARGO type 1 (timedisc_base.f90:2243-2326, direction 2): back to the real
 * frame with the old w, then per x-ring (ghost rings included) the mean of
 * the azimuthal velocity over JMIN:JMAX weighted by wfactor; SUM in index
 * order, as the reference build evaluates it */
static void calc_background_velocity(fvr *s) {
  const int N = s->N, jmin = G, jmax = G + s->c.jnum - 1;
  add_bg_velocity(s, &s->pvar, &s->cvar);
  PAR_FOR
  for (int i = 0; i < s->NI; i++) {
    double wi = 0.0;
    for (int j = jmin; j <= jmax; j++) wi = wi + s->pvar.d[2 * N + IDX(s, i, j)] * s->wfac[i];
    s->w[i] = wi / (double)s->c.jnum;
  }
}

/* timedisc_base.f90:1935-2069 (2-D, direction 2).  Every x-ring is
 * independent: per ring the limited transport by the fractional shift over
 * the whole column (dq, dql and flux indexed by j; dql(0) and the flux below
 * JMIN-1 stay zero as in the full-array form), then CSHIFT(., -shift) over
 * JMIN:JMAX. */
static void fargo_advection_y(fvr *s) {
  const int N = s->N, NI = s->NI, NJ = s->NJ, jmin = G, jmax = G + s->c.jnum - 1, J = s->c.jnum;
  PAR_FOR
  for (int i = 0; i < NI; i++) {
    double *dq = zalloc(NJ), *dql = zalloc(NJ), *flux = zalloc(NJ), *buf = zalloc(J);
    double delxy = s->w[i] * s->dt / s->dly[IDX(s, i, jmin)];
    const long shift = lround(delxy);
    delxy = delxy - (double)shift;
    for (int l = 0; l < s->V; l++) {
      double *u = s->cvar.d + (size_t)l * N;
      dq[0] = u[IDX(s, i, 1)] - u[IDX(s, i, 0)];
      for (int j = 1; j <= NJ - 2; j++) {
        dq[j] = u[IDX(s, i, j + 1)] - u[IDX(s, i, j)];
        if (fsign(1.0, dq[j - 1]) * fsign(1.0, dq[j]) > 0.)
          dql[j] = fsign(fmin2(fabs(dq[j - 1]), fabs(dq[j])), dq[j - 1]);
        else
          dql[j] = 0.;
      }
      for (int j = jmin - 1; j <= jmax; j++) {
        const int k = IDX(s, i, j);
        if (delxy > 0.)
          flux[j] = u[k] + .5 * dql[j] * (1. - delxy);
        else
          flux[j] = u[k + NI] - .5 * dql[j + 1] * (1. + delxy);
        u[k] = u[k] - delxy * (flux[j] - flux[j - 1]);
      }
    }
    /* CSHIFT(a, -shift) over JMIN:JMAX: result(j) = a(j - shift) cyclically */
    const long sh = ((-shift) % J + J) % J;
    for (int l = 0; l < s->V; l++) {
      double *u = s->cvar.d + (size_t)l * N;
      for (int j = 0; j < J; j++) buf[j] = u[IDX(s, i, jmin + (int)((j + sh) % J))];
      for (int j = 0; j < J; j++) u[IDX(s, i, jmin + j)] = buf[j];
    }
    free(dq); free(dql); free(flux); free(buf);
  }
}

/* timedisc_base.f90:961-1009 */
static void reject_solution(fvr *s, double time, double dt) {
  copy_sv(s, &s->cvar, &s->cold);
  compute_rhs(s, &s->pvar, &s->cvar, CHECK_NOTHING, s->coeff[0], dt, time);
  memcpy(s->bx, s->bxo, sizeof(double) * 2 * s->V * s->NJ);
  memcpy(s->by, s->byo, sizeof(double) * 2 * s->V * s->NI);
  s->n_adj++;
  s->dtcause = DTCAUSE_ERRADJ;
  if (dt < s->dtmin) s->dtmin = dt;
}

/* timedisc_base.f90:911-959 */
static void accept_solution(fvr *s, double *time, double dt) {
  *time = *time + dt;
  if (*time < s->time + s->dt) {
    compute_rhs(s, &s->pvar, &s->cvar, CHECK_ALL, s->coeff[0], dt, *time);
    copy_sv(s, &s->cold, &s->cvar);
  }
  memcpy(s->bxo, s->bx, sizeof(double) * 2 * s->V * s->NJ);
  memcpy(s->byo, s->by, sizeof(double) * 2 * s->V * s->NI);
  s->iter++;
}

/* timedisc_base.f90:649-750 */
static void integration_step(fvr *s) {
  double err = 0.0, dtold, dt, time;
  if (s->c.fargo == 1) calc_background_velocity(s); /* timedisc_base.f90:664-665 */
  if (s->c.fargo == 1 || s->c.fargo == 2) sub_bg_velocity(s, &s->pvar, &s->cvar);
  time = s->time;
  dt = s->dt;
  if (dt < s->dtmin && s->dtcause != DTCAUSE_FILEIO) s->dtmin = dt;
  while (!s->brk && time + dt <= s->time + s->dt) {
    dtold = dt;
    solve_ode(s, time, &dt, &err);
    if (err != err)
      s->brk = 1;
    else if (err > HUGE_)
      s->brk = 1;
    else if (err >= 1.0)
      reject_solution(s, time, dt);
    else
      accept_solution(s, &time, dtold);
    if (dt < s->c.dtlimit) s->brk = 1;
  }
  if (!s->brk) {
    s->dt = time - s->time;
    s->time = time;
    s->dtold = dt;
    if (s->c.fargo) {
      center_boundary(s, &s->pvar, &s->cvar);
      fargo_advection_y(s);
      PAR_FOR
      for (int k = 0; k < s->N; k++) cons2prim_pt(s, s->pvar.d, s->cvar.d, s->N, k);
      s->pvar.flag = s->cvar.flag;
      compute_rhs(s, &s->pvar, &s->cvar, CHECK_ALL, s->coeff[0], 0.0, s->time);
    } else {
      compute_rhs(s, &s->pvar, &s->cvar, CHECK_ALL, s->coeff[0], dtold, s->time);
    }
    copy_sv(s, &s->cold, &s->cvar);
  }
}

/* ---------------------------------------------------------- tableaux */

/* timedisc_ssprk.f90:93-126, timedisc_rkfehlberg.f90:305-335,
 * timedisc_dormand_prince.f90:91-114, timedisc_cashkarp.f90:95-117 */
static int set_tableau(fvr *s) {
  int m, o = s->c.order;
  static const double ssp3a[] = {0, 0, 0, 1.0, 0, 0, 0.25, 0.25, 0};
  static const double ssp3h[] = {1. / 6., 1. / 6., 2. / 3.}, ssp3l[] = {0.5, 0.5, 0.}, ssp3c[] = {0., 1., 0.5};
  static const double ssp5a[] = {0., 0., 0., 0., 0., 0.36717, 0., 0., 0., 0., 0.26802, 0.3172, 0., 0., 0.,
                                 0.11606, 0.13735, 0.18816, 0., 0., 0.11212, 0.13269, 0.18178, 0.4198, 0.};
  static const double ssp5h[] = {0.17279, 0.094505, 0.12947, 0.29899, 0.30424};
  static const double ssp5l[] = {0.12293, 0.31981, -0.15316, 0.31887, 0.39155};
  static const double ssp5c[] = {0., 0.36717, 0.58522, 0.44156, 0.8464};
  static const double rkf3a[] = {0.0, 0.0, 0.0, 0.5, 0.0, 0.0, -1.0, 2.0, 0.0};
  static const double rkf3h[] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}, rkf3l[] = {0.0, 1.0, 0.0}, rkf3c[] = {0.0, 0.5, 1.0};
  static const double rkf5a[] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                                 0.25, 0.0, 0.0, 0.0, 0.0, 0.0,
                                 3.0 / 32.0, 9.0 / 32.0, 0.0, 0.0, 0.0, 0.0,
                                 1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0.0, 0.0, 0.0,
                                 439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0.0, 0.0,
                                 -8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0, 0.0};
  static const double rkf5h[] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
  static const double rkf5l[] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -0.2, 0.0};
  static const double rkf5c[] = {0.0, 0.25, 3.0 / 8.0, 12.0 / 13.0, 1.0, 0.5};
  static const double dp5a[] = {0., 0., 0., 0., 0., 0., 0.,
                                .2, 0., 0., 0., 0., 0., 0.,
                                3. / 40., 9. / 40., 0., 0., 0., 0., 0.,
                                44. / 45., -56. / 15., 32. / 9., 0., 0., 0., 0.,
                                19372. / 6561., -25360. / 2187., 64448. / 6561., -212. / 729., 0., 0., 0.,
                                9017. / 3168., -355. / 33., 46732. / 5247., 49. / 176., -5103. / 18656., 0., 0.,
                                35. / 384., 0., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0.};
  static const double dp5h[] = {35. / 384., 0., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0.};
  static const double dp5l[] = {5179. / 57600., 0., 7571. / 16695., 393. / 640., -92097. / 339200., 187. / 2100., 1. / 40.};
  static const double dp5c[] = {0., .2, 3. / 10., 4. / 5., 8. / 9., 1., 1.};
  /* timedisc_cashkarp.f90:101-113 */
  static const double ck5a[] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                                0.2, 0.0, 0.0, 0.0, 0.0, 0.0,
                                0.075, 0.225, 0.0, 0.0, 0.0, 0.0,
                                0.3, -0.9, 1.2, 0.0, 0.0, 0.0,
                                -11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0.0, 0.0,
                                1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0, 0.0};
  static const double ck5h[] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
  static const double ck5l[] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 0.25};
  static const double ck5c[] = {0.0, 0.2, 0.3, 0.6, 1.0, 7.0 / 8.0};
  const double *A, *H, *L, *C;
  if (s->c.method == SSPRK && o == 3) { m = 3; A = ssp3a; H = ssp3h; L = ssp3l; C = ssp3c; }
  else if (s->c.method == SSPRK && o == 5) { m = 5; A = ssp5a; H = ssp5h; L = ssp5l; C = ssp5c; }
  else if (s->c.method == RK_FEHLBERG && o == 3) { m = 3; A = rkf3a; H = rkf3h; L = rkf3l; C = rkf3c; }
  else if (s->c.method == RK_FEHLBERG && o == 5) { m = 6; A = rkf5a; H = rkf5h; L = rkf5l; C = rkf5c; }
  else if (s->c.method == DORMAND_PRINCE && o == 5) { m = 7; A = dp5a; H = dp5h; L = dp5l; C = dp5c; }
  else if (s->c.method == CASH_KARP && o == 5) { m = 6; A = ck5a; H = ck5h; L = ck5l; C = ck5c; }
  else return -1;
  s->m = m;
  s->a = zalloc(m * m); s->b_high = zalloc(m); s->b_low = zalloc(m); s->cc = zalloc(m);
  memcpy(s->a, A, sizeof(double) * m * m);
  memcpy(s->b_high, H, sizeof(double) * m);
  memcpy(s->b_low, L, sizeof(double) * m);
  memcpy(s->cc, C, sizeof(double) * m);
  return 0;
}

/* ------------------------------------------------------------- public */

fvr *fvr_create(const fvr_cfg *cfg) {
  fvr *s;
  if (cfg->physics != EULER && cfg->physics != EULER_ISOTHERM) return NULL;
  if (cfg->geometry != CARTESIAN && cfg->geometry != LOGCYLINDRICAL) return NULL;
  if (cfg->inum < 4 || cfg->jnum < 4) return NULL;
  if (cfg->fargo < 0 || cfg->fargo > 2) return NULL;
  for (int d = 0; d < 4; d++) {
    int b = cfg->bc[d];
    if (b != NO_GRADIENTS && b != PERIODIC && b != ABSORBING) return NULL;
    if (b == ABSORBING && d >= 2) return NULL;
  }
  if (cfg->outbound < 0 || cfg->outbound > 4) return NULL;
  if (cfg->rotframe < 0 || cfg->rotframe > 2) return NULL;
  if (cfg->vismodel < 0 || cfg->vismodel > 3) return NULL;
  s = (fvr *)calloc(1, sizeof(fvr));
  s->c = *cfg;
  s->pm_mass = cfg->mass; /* mdot, massloss, accrate start at 0 (gravity_pointmass.f90:206-209) */
  s->V = cfg->physics == EULER ? 4 : 3;
  s->NI = cfg->inum + 2 * G;
  s->NJ = cfg->jnum + 2 * G;
  s->N = s->NI * s->NJ;
  if (set_tableau(s)) { free(s); return NULL; }
  {
    size_t N = s->N, V = s->V;
    for (int f = 0; f < 2; f++) { s->dAx[f] = zalloc(N); s->dAy[f] = zalloc(N); s->accel[f] = zalloc(N); }
    s->dydzdV = zalloc(N); s->dzdxdV = zalloc(N); s->dxdydV = zalloc(N); s->dlx = zalloc(N); s->dly = zalloc(N);
    s->cxyx = zalloc(N); s->cyxy = zalloc(N); s->czxz = zalloc(N); s->czyz = zalloc(N);
    for (int f = 0; f < 4; f++) { s->rdx[f] = zalloc(N); s->radf[f] = zalloc(N); }
    s->rad = zalloc(N); s->posx = zalloc(N); s->posy = zalloc(N); s->bcx = zalloc(N); s->bcy = zalloc(N);
    s->caccel[0] = zalloc(N); s->caccel[1] = zalloc(N); s->twoomega = zalloc(N);
    s->raccel[0] = zalloc(N); s->raccel[1] = zalloc(N);
    s->dynvis = zalloc(N); s->bulkvis = zalloc(N); s->ephir[0] = zalloc(N); s->ephir[1] = zalloc(N);
    s->btxx = zalloc(N); s->btxy = zalloc(N); s->btyy = zalloc(N); s->vtmp = zalloc(N); s->vtmp1 = zalloc(N);
    s->vsterm = zalloc(V * N);
    s->bccsound = zalloc(N); s->fcsound = zalloc(4 * N); s->w = zalloc(s->NI); s->wfac = zalloc(s->NI);
    s->pvar.d = zalloc(V * N); s->cvar.d = zalloc(V * N); s->ptmp.d = zalloc(V * N);
    s->ctmp.d = zalloc(V * N); s->cold.d = zalloc(V * N);
    s->geo_src = zalloc(V * N); s->src = zalloc(V * N); s->sterm = zalloc(V * N); s->cerr = zalloc(V * N);
    s->xflux = zalloc(V * N); s->yflux = zalloc(V * N);
    s->bx = zalloc(2 * V * s->NJ); s->by = zalloc(2 * V * s->NI);
    s->bxo = zalloc(2 * V * s->NJ); s->byo = zalloc(2 * V * s->NI);
    s->coeff = (double **)calloc(s->m, sizeof(double *));
    for (int k = 0; k < s->m; k++) s->coeff[k] = zalloc(V * N);
    s->prim = zalloc(4 * V * N); s->cons = zalloc(4 * V * N); s->pfluxes = zalloc(4 * V * N);
    s->minwav = zalloc(2 * N); s->maxwav = zalloc(2 * N); s->tmp = zalloc(N); s->tmp1 = zalloc(N);
    s->slopes = zalloc(2 * V * N);
  }
  mesh_build(s);
  /* FARGO type 1 wfactor (timedisc_base.f90:428-447, direction 2): the
   * smallest hy over the W, E, bottom and top faces of the ring over hy at
   * the bary-centre; the bottom/top faces sit at the centre's x (mesh_dz = 0) */
  for (int i = 0; i < s->NI; i++) {
    const double mdx = s->dx, xc = s->c.xmin + ((double)(2 * (i - 1) - 1) * 0.5) * mdx;
    const double xw = xc + (-(0.5 * mdx)), xe = xc + 0.5 * mdx, xbt = xc + 0.0;
    double hw = hscale(s, xw), he = hscale(s, xe), hbt = hscale(s, xbt), hc = hscale(s, xc);
    if (s->c.geometry == CARTESIAN) hw = he = hbt = hc = 1.0;
    s->wfac[i] = fmin2(fmin2(hw, he), fmin2(hbt, hbt)) / hc;
  }
  /* timedisc_base.f90:224-359 defaults */
  s->time = cfg->starttime;
  s->dt = cfg->stoptime;
  s->dtold = s->dt;
  s->dtmin = s->dt;
  if (s->c.dtlimit <= 0.0) s->c.dtlimit = EPS_ * cfg->stoptime;
  s->pmin = cfg->pmin > 0.0 ? cfg->pmin : TINY_;
  s->rhomin = cfg->rhomin > 0.0 ? cfg->rhomin : TINY_;
  s->maxerrold = 0.;
  s->iter = -1;
  /* Datafile (fileio_base.f90:474-487, 591-600): the output clock starts at
   * the start time and advances by |stoptime - starttime|/count; one IncTime
   * happens before the first Step, in InitFileIO (start time > 0) or after
   * the initial dataset of FirstStep (start time 0) */
  s->out_count = cfg->count > 0 ? cfg->count : 1;
  s->t_out = s->time + fabs(cfg->stoptime - s->time) / s->out_count;
  return s;
}

void fvr_destroy(fvr *s) {
  if (!s) return;
  free(s->dAx[0]); free(s->dAx[1]); free(s->dAy[0]); free(s->dAy[1]); free(s->accel[0]); free(s->accel[1]);
  free(s->dydzdV); free(s->dzdxdV); free(s->dxdydV); free(s->dlx); free(s->dly);
  free(s->cxyx); free(s->cyxy); free(s->czxz); free(s->czyz);
  for (int f = 0; f < 4; f++) { free(s->rdx[f]); free(s->radf[f]); }
  free(s->rad); free(s->posx); free(s->posy); free(s->bcx); free(s->bcy);
  free(s->caccel[0]); free(s->caccel[1]); free(s->twoomega); free(s->raccel[0]); free(s->raccel[1]);
  free(s->dynvis); free(s->bulkvis); free(s->ephir[0]); free(s->ephir[1]);
  free(s->btxx); free(s->btxy); free(s->btyy); free(s->vtmp); free(s->vtmp1); free(s->vsterm);
  free(s->bccsound); free(s->fcsound); free(s->w); free(s->wfac);
  free(s->pvar.d); free(s->cvar.d); free(s->ptmp.d); free(s->ctmp.d); free(s->cold.d);
  free(s->geo_src); free(s->src); free(s->sterm); free(s->cerr); free(s->xflux); free(s->yflux);
  free(s->bx); free(s->by); free(s->bxo); free(s->byo);
  for (int k = 0; k < s->m; k++) free(s->coeff[k]);
  free(s->coeff);
  free(s->prim); free(s->cons); free(s->pfluxes); free(s->minwav); free(s->maxwav);
  free(s->tmp); free(s->tmp1); free(s->slopes);
  free(s->a); free(s->b_high); free(s->b_low); free(s->cc);
  free(s);
}

int fvr_ncells(const fvr *s) { return s->N; }
int fvr_vnum(const fvr *s) { return s->V; }

int fvr_mesh(const fvr *s, int id, double *out) {
  const double *src[30] = {s->dAx[0], s->dAx[1], s->dAy[0], s->dAy[1], s->dydzdV, s->dzdxdV, s->dxdydV,
                           s->dlx, s->dly, s->cxyx, s->cyxy, s->czxz, s->czyz, s->rdx[0], s->rdx[1],
                           s->rdx[2], s->rdx[3], s->rad, s->radf[0], s->radf[1], s->radf[2], s->radf[3],
                           s->posx, s->posy, s->bcx, s->bcy, s->accel[0], s->accel[1],
                           s->caccel[0], s->caccel[1]};
  if (id < 0 || id >= 30) return -1;
  memcpy(out, src[id], sizeof(double) * s->N);
  return 0;
}

void fvr_set_sound_speeds(fvr *s, const double *bc, const double *fc) {
  memcpy(s->bccsound, bc, sizeof(double) * s->N);
  memcpy(s->fcsound, fc, sizeof(double) * 4 * s->N);
}

void fvr_set_background_velocity(fvr *s, const double *w) { memcpy(s->w, w, sizeof(double) * s->NI); }

void fvr_set_pvar(fvr *s, const double *pvar) {
  memcpy(s->pvar.d, pvar, sizeof(double) * s->V * s->N);
  for (int k = 0; k < s->N; k++) prim2cons_pt(s, s->cvar.d, s->pvar.d, s->N, k);
  s->cvar.flag = s->pvar.flag;
}

/* fosite.f90:258-348 */
int fvr_first_step(fvr *s) {
  s->iter = 0;
  if (s->c.fargo == 1) { /* fosite.f90:305-311 */
    center_boundary(s, &s->pvar, &s->cvar);
    calc_background_velocity(s);
  }
  compute_rhs(s, &s->pvar, &s->cvar, CHECK_ALL, s->coeff[0], 0.0, s->time);
  s->dt = calc_timestep(s);
  copy_sv(s, &s->cold, &s->cvar);
  /* initial WriteDataset: fileio_base.f90:523-533 moves the state to the
   * real frame (the next IntegrationStep moves it back) */
  if (s->time == 0.0 && s->c.fargo) add_bg_velocity(s, &s->pvar, &s->cvar);
  return s->brk;
}

/* fosite.f90:367-464 (outputs at the Datafile times) */
int fvr_step(fvr *s) {
  if (s->iter == -1) fvr_first_step(s);
  if (fabs(s->c.stoptime - s->time) <= 1.0E-05 * s->c.stoptime) return 1;
  /* fileio_base.f90:604-620 */
  if ((s->time + s->dt) / s->t_out > 1.0) {
    s->dt = s->t_out - s->time;
    s->dtcause = DTCAUSE_FILEIO;
  } else if ((s->time + 1.5 * s->dt) / s->t_out > 1.0) {
    s->dt = 0.5 * (s->t_out - s->time);
    s->dtcause = DTCAUSE_FILEIO;
  }
  integration_step(s);
  if (fabs(s->t_out - s->time) <= 1.0E-5 * s->t_out || s->brk) {
    if (s->c.fargo) add_bg_velocity(s, &s->pvar, &s->cvar); /* WriteDataset */
    if (s->brk) return 1;
    s->t_out = s->t_out + fabs(s->c.stoptime - s->c.starttime) / s->out_count; /* IncTime */
    s->dtmin = s->c.stoptime;
    s->n_adj = 0;
  }
  s->dt = calc_timestep(s);
  return 0;
}

void fvr_get(const fvr *s, double *cvar, double *pvar, fvr_scalars *sc) {
  if (cvar) memcpy(cvar, s->cvar.d, sizeof(double) * s->V * s->N);
  if (pvar) memcpy(pvar, s->pvar.d, sizeof(double) * s->V * s->N);
  if (sc) {
    sc->iter = s->iter; sc->n_adj = s->n_adj; sc->brk = s->brk; sc->dtcause = s->dtcause;
    sc->time = s->time; sc->dt = s->dt; sc->dtold = s->dtold; sc->maxerrold = s->maxerrold;
  }
}

void fvr_get_background_velocity(const fvr *s, double *w) { memcpy(w, s->w, sizeof(double) * s->NI); }

void fvr_get_bflux(const fvr *s, double *bx, double *by) {
  if (bx) memcpy(bx, s->bx, sizeof(double) * 2 * s->V * s->NJ);
  if (by) memcpy(by, s->by, sizeof(double) * 2 * s->V * s->NI);
}

void fvr_get_pointmass(const fvr *s, double *out) {
  out[0] = s->pm_mass;
  out[1] = s->pm_accrate;
  out[2] = s->pm_massloss;
  out[3] = s->pm_mdot;
}

void fvr_compute_rhs(fvr *s, double *rhs_out) {
  size_t n = (size_t)s->V * s->N;
  sv p = {(double *)malloc(sizeof(double) * n), s->pvar.flag};
  sv c = {(double *)malloc(sizeof(double) * n), s->cvar.flag};
  memcpy(p.d, s->pvar.d, sizeof(double) * n);
  memcpy(c.d, s->cvar.d, sizeof(double) * n);
  memset(rhs_out, 0, sizeof(double) * n);
  compute_rhs(s, &p, &c, CHECK_NOTHING, rhs_out, -1.0, s->time);
  free(p.d);
  free(c.d);
}

/* Test aid with no reference counterpart: rounds every array the FP32 form
 * stores between steps (the state, the saved state and the stored
 * first-stage rhs) to IEEE single precision and recomputes the primitive
 * variables.  Steps run in FP64 as before, so a run that calls this after
 * every step shows the accuracy floor that FP32 storage alone sets. */
void fvr_round_state_f32(fvr *s) {
  const size_t n = (size_t)s->V * s->N;
  for (size_t k = 0; k < n; k++) {
    s->cvar.d[k] = (double)(float)s->cvar.d[k];
    s->cold.d[k] = (double)(float)s->cold.d[k];
    s->coeff[0][k] = (double)(float)s->coeff[0][k];
  }
  for (int k = 0; k < s->N; k++) cons2prim_pt(s, s->pvar.d, s->cvar.d, s->N, k);
}
