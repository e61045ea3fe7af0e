// fv_kernels.h -- CDNA4 (gfx950) kernels of fosite's unsplit finite-volume
// step (templates; instantiated by fv_inst_*.hip, dispatched by fv_kernels.hip).  Every kernel is a template on the real type T of the stored state:
//
//   T = double  the plain drop-in.  Every floating-point expression keeps the
//               reference's operation order and the file is compiled with
//               -ffp-contract=off, so IEEE +,-,*,/ and sqrt give the
//               reference's bits.
//   T = float   the precision form: reconstruction, fluxes, sources and the
//               stage states in FP32 (hardware reciprocal and square root,
//               FP32 copies of the metric arrays); the new state of a step
//               accumulated with compensated (two-sum / two-product)
//               arithmetic; the embedded-error norm in FP64; CFL partials
//               reduced in FP64 by the host path or k_control.
//
// Citations are file:line in the fosite source tree.
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
// k_remap: FARGO azimuthal shift of the combined state (y direction only).
// k_control: one block; in device-controlled batches it reduces the partials
// of the last stage and of the end-of-step k_rhs, runs the accept test and
// the CFL update.  Kernel boundaries order everything: no grid-wide fences.
#pragma once
#include <float.h>
#include <math.h>

#include <cstring>

#include "fv_device.h"

namespace fv {

// ------------------------------------------------------------ arithmetic
template <typename T>
struct Real;
template <>
struct Real<double> {
  static constexpr double tiny = DBL_MIN, huge = DBL_MAX;
};
template <>
struct Real<float> {
  static constexpr float tiny = FLT_MIN, huge = FLT_MAX;
};

template <typename T>
__device__ __forceinline__ T fmin2(T a, T b) { return (b < a) ? b : a; }
template <typename T>
__device__ __forceinline__ T fmax2(T a, T b) { return (b > a) ? b : a; }
template <typename T>
__device__ __forceinline__ T fsign(T a, T b) { return copysign(fabs(a), b); }
template <typename T>
__device__ __forceinline__ bool bad(T x) { return (x != x) || (x > Real<T>::huge); }
// division, square root, log, exp: IEEE in FP64; hardware-rate in FP32
__device__ __forceinline__ double dv(double a, double b) { return a / b; }
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double sq(double a) { return sqrt(a); }
__device__ __forceinline__ float sq(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double lg(double a) { return log(a); }
__device__ __forceinline__ float lg(float a) { return __logf(a); }
__device__ __forceinline__ double ex(double a) { return exp(a); }
__device__ __forceinline__ float ex(float a) { return __expf(a); }

// error-free transformations for the compensated FP32 accumulation
// (Knuth's two-sum; two-product with a fused multiply-add)
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
// hi + lo = y - sum_q c_q k_q, with c_q = (chi_q + clo_q) given in two parts
struct Acc {
  float hi, lo;
  __device__ __forceinline__ explicit Acc(float y) : hi(y), lo(0.f) {}
  __device__ __forceinline__ void sub(float ch, float cl, float k) {
    const float p = ch * k, pe = fmaf(ch, k, -p);  // ch*k = p + pe exactly
    float s, e;
    two_sum(hi, -p, s, e);
    hi = s;
    lo = lo + (e - pe - cl * k);
  }
  __device__ __forceinline__ float value() const { return hi + lo; }
  __device__ __forceinline__ double wide() const { return (double)hi + (double)lo; }
};

// per-ring and per-column metric arrays: uniform base per id, 32-bit lane
// offset; the FP32 form reads FP32 copies
__device__ __forceinline__ double ring(const Geom &g, int id, int i) { return (g.ring + id * g.NI)[(unsigned)i]; }
template <typename T>
__device__ __forceinline__ T rg(const Geom &g, int id, int i) {
  if constexpr (sizeof(T) == 4)
    return (g.ringf + id * g.NI)[(unsigned)i];
  else
    return ring(g, id, i);
}
template <typename T>
__device__ __forceinline__ T cl(const Geom &g, int id, int j) {
  if constexpr (sizeof(T) == 4)
    return (g.colf + id * g.NJ)[(unsigned)j];
  else
    return (g.col + id * g.NJ)[(unsigned)j];
}
__device__ __forceinline__ long cidx(const Geom &g, int i, int j) { return (long)j * g.P + i; }

// ---------------------------------------------------------------- physics
// Cons2Prim2d / Prim2Cons2d (physics_eulerisotherm.f90:2835-2881,
// physics_euler.f90:3056-3104)
template <typename T, int V>
__device__ __forceinline__ void c2p(const T *u, T *p, T gamma) {
  T inv = dv(T(1), u[0]);
  p[0] = u[0];
  p[1] = u[1] * inv;
  p[2] = u[2] * inv;
  if (V == 4) p[3] = (gamma - T(1)) * (u[3] - T(0.5) * inv * (u[1] * u[1] + u[2] * u[2]));
}
template <typename T, int V>
__device__ __forceinline__ void p2c(const T *p, T *u, T gamma) {
  u[0] = p[0];
  u[1] = p[0] * p[1];
  u[2] = p[0] * p[2];
  if (V == 4) u[3] = dv(p[3], gamma - T(1)) + T(0.5) * p[0] * (p[1] * p[1] + p[2] * p[2]);
}
// physics_euler.f90:2682-2689
template <typename T>
__device__ __forceinline__ T sound_speed(T gamma, T rho, T pr) {
  return sq(fmax2(T(2) * Real<T>::tiny, dv(gamma * pr, rho)));
}

// Absorbing boundary in x: characteristic variables from the two edge cells
// (primitive states pL at the lower, pR at the higher index), incoming waves
// set to zero, extrapolation into the ghost cells (boundary_absorbing.f90:
// 114-150).  Isothermal: physics_eulerisotherm.f90:1764-1849, 2027-2085,
// 2895-2969.  Euler: physics_euler.f90:743-846, 1072-1129 with
// SetEigenValues, SetCharVars and SetBoundaryData (1d/2d).  i1 is the edge
// cell, iR the higher-index cell of the pair.
template <typename C, int V>
__device__ __forceinline__ void absorbing_math(const Geom &g, int east, int i1, int iR, C gam, const C *pL,
                                               const C *pR, C *p) {
  const C delta = east ? C(1) : C(-1);
  C p1[V];  // element-wise select keeps both states in registers
#pragma unroll
  for (int l = 0; l < V; l++) p1[l] = east ? pR[l] : pL[l];
  if constexpr (V == 4) {
    const C rho1 = p1[0], u1 = p1[1], v1 = p1[2], P1 = p1[3];
    const C cs = sound_speed(gam, rho1, P1);
    const C lam0 = u1 - cs, lam1 = u1, lam3 = u1 + cs;
    const C dlnP = lg(dv(pR[3], pL[3]));
    const C gammadu_cs = dv(C(2) * gam * (pR[1] - pL[1]), lam3 - lam0);
    C x0 = dlnP - gammadu_cs, x1 = dlnP - gam * lg(dv(pR[0], pL[0])), x2 = pR[2] - pL[2];
    C x3 = dlnP + gammadu_cs;
    if (east) {
      if (lam0 <= C(0)) x0 = C(0);
      if (lam1 <= C(0)) x1 = C(0), x2 = C(0);
      if (lam3 <= C(0)) x3 = C(0);
    } else {
      if (lam0 >= C(0)) x0 = C(0);
      if (lam1 >= C(0)) x1 = C(0), x2 = C(0);
      if (lam3 >= C(0)) x3 = C(0);
    }
    const C dlp = C(0.5) * (x3 + x0);
    const C rho2 = rho1 * ex(dv(delta, gam) * (dlp - x1));
    const C P2 = P1 * ex(delta * dlp);
    const C cs2 = sound_speed(gam, rho1 + rho2, P1 + P2);
    p[0] = rho2;
    p[1] = u1 + dv(C(0.5) * delta * cs2, gam) * (x3 - x0);
    p[2] = v1 + delta * x2;
    p[3] = P2;
  } else {
    const C cs = rg<C>(g, R_CS, i1), u1 = p1[1];
    const C lam0 = u1 - cs, lam1 = u1, lam2 = u1 + cs;
    const C csf = rg<C>(g, R_CSW, iR);
    const C dlnrho = lg(dv(pR[0], pL[0]));
    const C ducs = dv(pR[1] - pL[1], csf);
    C x0 = dlnrho - ducs, x1 = pR[2] - pL[2], x2 = dlnrho + ducs;
    if (east) {
      if (lam0 <= C(0)) x0 = C(0);
      if (lam1 <= C(0)) x1 = C(0);
      if (lam2 <= C(0)) x2 = C(0);
    } else {
      if (lam0 >= C(0)) x0 = C(0);
      if (lam1 >= C(0)) x1 = C(0);
      if (lam2 >= C(0)) x2 = C(0);
    }
    const C cs2 = rg<C>(g, east ? R_CSE : R_CSW, i1);
    p[0] = p1[0] * ex(delta * C(0.5) * (x2 + x0));
    p[1] = u1 + delta * C(0.5) * cs2 * (x2 - x0);
    p[2] = p1[2] + delta * x1;
  }
}

// FARGO frame change of one conservative state (physics_eulerisotherm.f90:
// 1263-1303, 1373-1413): + or - rho*w on the azimuthal momentum; Euler first
// changes the energy by w*(m +- 0.5*rho*w) with the old momentum
// (physics_euler.f90:2003-2047, 2159-2203)
template <typename T, int V>
__device__ __forceinline__ void frame_add(T *u, T w) {
  if (V == 4) u[3] = u[3] + w * (u[2] + T(0.5) * u[0] * w);
  u[2] = u[2] + u[0] * w;
}
template <typename T, int V>
__device__ __forceinline__ void frame_sub(T *u, T w) {
  if (V == 4) u[3] = u[3] - w * (u[2] - T(0.5) * u[0] * w);
  u[2] = u[2] - u[0] * w;
}

// ------------------------------------------------------------ input state
// KM: coefficient slots of the tableau (3 for up to three stages, else MAXK);
// loops over stages are unrolled to KM with run-time guards on nk
template <typename T, int V, int KM>
struct In {
  const Geom &g;
  const StageArgsT<T> &a;
  double dt;
  T gam;
  int nk;
  bool chk;  // CheckData on the ghost cells
  // stage coefficients a(m,q)*dt, formed once (KM = 0: no stage sum)
  static constexpr int KA = KM > 0 ? KM : 1;
  double c[sizeof(T) == 8 ? KA : 1];
  float ch[sizeof(T) == 4 ? KA : 1];

  __forceinline__ __device__ In(const Geom &g_, const StageArgsT<T> &a_, bool chk_) : g(g_), a(a_), chk(chk_) {
    dt = a.ctl ? a.ctl->dt : a.dt;  // device-controlled: this step's dt
    nk = a.nk;
    gam = (T)g.gamma;
#pragma unroll
    for (int q = 0; q < KM; q++) {
      const double cq = a.a[q] * dt;
      if constexpr (sizeof(T) == 8)
        c[q] = cq;
      else
        ch[q] = (float)cq;
    }
  }

  // s = Y - sum_q (a_q*dt) k_q for one cell from loaded values, in the
  // reference's order (FP64) or with FP32 fused multiply-adds.
  __forceinline__ __device__ void stage_sum(const T *y, const T (*kv)[V], T *s) const {
    if constexpr (KM == 0) {
#pragma unroll
      for (int l = 0; l < V; l++) s[l] = y[l];
      return;
    }
    if constexpr (sizeof(T) == 8) {
#pragma unroll
      for (int l = 0; l < V; l++) s[l] = y[l];
#pragma unroll
      for (int q = 0; q < KM; q++)
        if (q < nk) {
#pragma unroll
          for (int l = 0; l < V; l++) s[l] = s[l] - c[q] * kv[q][l];
        }
    } else {
      // the stage state only feeds an FP32 RHS evaluation: plain FP32 sums
#pragma unroll
      for (int l = 0; l < V; l++) {
        float acc = (float)y[l];
#pragma unroll
        for (int q = 0; q < KM; q++)
          if (q < nk) acc = fmaf(-ch[q], (float)kv[q][l], acc);
        s[l] = (T)acc;
      }
    }
  }

  // input state at an interior cell
  __forceinline__ __device__ void interior(int i, int j, T *s) const {
    const long ix = cidx(g, i, j);
    // all loads first (uniform branches only), then the stage sum
    T y[V], kv[KA][V];
#pragma unroll
    for (int l = 0; l < V; l++) y[l] = a.Y[l * g.FS + ix];
#pragma unroll
    for (int q = 0; q < KM; q++)
      if (q < nk) {
#pragma unroll
        for (int l = 0; l < V; l++) kv[q][l] = a.k[q][l * g.FS + ix];
      }
    stage_sum(y, kv, s);
  }

  // stage states of two interior cells (p0, p1 say which are wanted): all
  // loads of both cells are issued before any arithmetic
  __forceinline__ __device__ void interior2(bool p0, int ia, int ja, T *s0, bool p1, int ib, int jb, T *s1) const {
    const unsigned x0 = (unsigned)ja * (unsigned)g.P + (unsigned)ia;
    const unsigned x1 = (unsigned)jb * (unsigned)g.P + (unsigned)ib;
    T y0[V], y1[V], k0[KA][V], k1[KA][V];
    if (p0) {
#pragma unroll
      for (int l = 0; l < V; l++) y0[l] = (a.Y + l * g.FS)[x0];
#pragma unroll
      for (int q = 0; q < KM; q++)
        if (q < nk) {
#pragma unroll
          for (int l = 0; l < V; l++) k0[q][l] = (a.k[q] + l * g.FS)[x0];
        }
    }
    if (p1) {
#pragma unroll
      for (int l = 0; l < V; l++) y1[l] = (a.Y + l * g.FS)[x1];
#pragma unroll
      for (int q = 0; q < KM; q++)
        if (q < nk) {
#pragma unroll
          for (int l = 0; l < V; l++) k1[q][l] = (a.k[q] + l * g.FS)[x1];
        }
    }
    stage_sum(y0, k0, s0);
    stage_sum(y1, k1, s1);
  }

  // ComputeRHS transform (timedisc_base.f90:1098-1208): ur = cvar after the
  // frame change, CenterBoundary and PMIN (real frame), pr = pvar at that
  // point, uc = cvar handed to the fluxes
  __forceinline__ __device__ void transform(int i, const T *s, T *ur, T *pr, T *uc) const {
#pragma unroll
    for (int l = 0; l < V; l++) ur[l] = s[l];
    if (a.add_bg) frame_add<T, V>(ur, rg<T>(g, R_W, i));
    c2p<T, V>(ur, pr, gam);
    if (V == 4 && a.pmin_clamp) {
      pr[3] = fmax2(pr[3], (T)g.pmin);
      p2c<T, V>(pr, ur, gam);
    }
#pragma unroll
    for (int l = 0; l < V; l++) uc[l] = ur[l];
    if constexpr (sizeof(T) == 4) {
      // FP32 (isothermal only): m' + rho*w - rho*w would cancel to a few
      // bits of the residual momentum m'; the stored residual is what the
      // subtraction means
      if (a.sub_bg) uc[2] = a.add_bg ? s[2] : ur[2] - ur[0] * rg<T>(g, R_W, i);
    } else {
      if (a.sub_bg) frame_sub<T, V>(uc, rg<T>(g, R_W, i));
    }
  }

  // PRIMITIVE reconstruction: the pvar handed to the reconstruction, after
  // SubtractBackgroundVelocityY (physics_eulerisotherm.f90:1373-1413) of the
  // real-frame pvar p (as transform returns it); s is the stage state
  __forceinline__ __device__ void prim_comoving(int i, const T *s, const T *p, T *po) const {
#pragma unroll
    for (int l = 0; l < V; l++) po[l] = p[l];
    if constexpr (sizeof(T) == 4) {
      // FP32: (v' + w) - w would keep a few bits of the residual velocity v';
      // the stored residual is what the subtraction means
      if (a.sub_bg) po[2] = a.add_bg ? dv(s[2], s[0]) : p[2] - rg<T>(g, R_W, i);
    } else {
      if (a.sub_bg) po[2] = p[2] - rg<T>(g, R_W, i);
    }
  }

  // pvar of an interior cell right after CenterBoundary's conversion
  __forceinline__ __device__ void prim_interior(int i, int j, T *p) const {
    T s[V];
    interior(i, j, s);
    if (a.add_bg) frame_add<T, V>(s, rg<T>(g, R_W, i));
    c2p<T, V>(s, p, gam);
  }

  // absorbing boundary, x direction (absorbing_math) at row j
  __forceinline__ __device__ void absorbing_x(int east, int j, T *p) const {
    const int imin = GH, imax = GH + g.inum - 1;
    const int i1 = east ? imax : imin, i2 = east ? imax - 1 : imin + 1;
    const int iL = i1 < i2 ? i1 : i2, iR = i1 < i2 ? i2 : i1;
    T pL[V], pR[V];
    prim_interior(iL, j, pL);
    prim_interior(iR, j, pR);
    absorbing_math<T, V>(g, east, i1, iR, gam, pL, pR, p);
  }

  // x-ghost (side, layer 1..2) at interior row j
  __forceinline__ __device__ void ghost_we(int east, int gl, int j, T *p) const {
    const int imin = GH, imax = GH + g.inum - 1, bc = east ? g.bc[1] : g.bc[0];
    if (bc == NO_GRADIENTS)
      prim_interior(east ? imax : imin, j, p);
    else if (bc == PERIODIC)
      prim_interior(east ? imin + gl - 1 : imax - gl + 1, j, p);
    else
      absorbing_x(east, j, p);
  }
  // y-ghost (side, layer) at interior column i
  __forceinline__ __device__ void ghost_sn(int north, int gl, int i, T *p) const {
    const int jmin = GH, jmax = GH + g.jnum - 1, bc = north ? g.bc[3] : g.bc[2];
    if (bc == NO_GRADIENTS)
      prim_interior(i, north ? jmax : jmin, p);
    else
      prim_interior(i, north ? jmin + gl - 1 : jmax - gl + 1, p);
  }

  // pvar of any ghost cell after CenterBoundary: boundary rules in the order
  // W, E, S, N, then SetCornerEdges (boundary_generic.f90:434-517)
  __forceinline__ __device__ void ghost_prim(int i, int j, T *p) const {
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
        T px[V], py[V];
        ghost_we(east, ga, jr, px);
        ghost_sn(north, gb, ir, py);
#pragma unroll
        for (int l = 0; l < V; l++) p[l] = T(0.5) * (px[l] + py[l]);
      }
    }
  }

  // final ghost state: PMIN round trip, prim->cons, data check, frame change
  __forceinline__ __device__ void ghost_final(int i, int j, T *u, T *p, int &flags) const {
    ghost_prim(i, j, p);
    if (V == 4 && a.pmin_clamp) p[3] = fmax2(p[3], (T)g.pmin);
    p2c<T, V>(p, u, gam);
    if (chk) {
#pragma unroll
      for (int l = 0; l < V; l++)
        if (bad(u[l]) || bad(p[l])) flags |= F_INVALID;
    }
    if (a.sub_bg) {
      const T w = rg<T>(g, R_W, i);
      p[2] = p[2] - w;
      frame_sub<T, V>(u, w);
    }
  }
};

// CFL contribution of one cell (timedisc_base.f90:869-872 with
// physics_eulerisotherm.f90:530-552); u, v = pvar after the frame change back
template <typename T>
__device__ __forceinline__ T cell_invdt(const Geom &g, int i, T u, T v, T cs) {
  const T mnx = fmin2(T(0), u - cs), mxx = fmax2(T(0), u + cs);
  const T mny = fmin2(T(0), v - cs), mxy = fmax2(T(0), v + cs);
  return dv(fmax2(mxx, -mnx), rg<T>(g, R_DLX, i)) + dv(fmax2(mxy, -mny), rg<T>(g, R_DLY, i));
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
static __device__ double blk_max(double v, double *red) {
  v = wave_max(v);
  const int t = threadIdx.x + threadIdx.y * blockDim.x, nw = (blockDim.x * blockDim.y + 63) / 64;
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double r = -DBL_MAX;
  for (int k = 0; k < nw; k++) r = fmax2(r, red[k]);
  return r;
}
// block-wide max of v and OR of f with one barrier; the results are valid in
// thread 0 only, and a block may use its scratch (red, redi) once
static __device__ inline void blk_max_or0(double &v, int &f, double *red, int *redi) {
  v = wave_max(v);
  f = wave_or(f);
  const int t = threadIdx.x + threadIdx.y * blockDim.x, nw = (blockDim.x * blockDim.y + 63) / 64;
  if ((t & 63) == 0) {
    red[t >> 6] = v;
    redi[t >> 6] = f;
  }
  __syncthreads();
  if (t == 0) {  // from -DBL_MAX in wave order, as blk_max
    v = -DBL_MAX;
    f = 0;
    for (int k = 0; k < nw; k++) {
      v = fmax2(v, red[k]);
      f |= redi[k];
    }
  }
}
static __device__ int blk_or(int v, int *red) {
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
template <typename T, int LIM>
__device__ __forceinline__ T slope(const Geom &g, T um, T u0, T up) {
  const T a1 = u0 - um, a2 = up - u0;
  const T one = T(1), zero = T(0), tiny = Real<T>::tiny;
  switch (LIM ? LIM : g.limiter) {
    case MINMOD:
      return (fsign(one, a1) * fsign(one, a2) > zero) ? fsign(fmin2(fabs(a1), fabs(a2)), a1) : zero;
    case MONOCENT: {
      const T th = (T)g.theta;
      const T b1 = th * (u0 - um), b2 = th * (up - u0), b3 = T(0.5) * (up - um);
      if ((fsign(one, b1) * fsign(one, b2) > zero) && (fsign(one, b2) * fsign(one, b3) > zero))
        return fsign(fmin2(fmin2(fabs(b1), fabs(b2)), fabs(b3)), b1);
      return zero;
    }
    case SWEBY:
    case SUPERBEE: {
      const T q = (LIM ? LIM : g.limiter) == SUPERBEE ? T(2) : (T)g.theta;
      return (fsign(one, a1) * fsign(one, a2) > zero)
                 ? fsign(fmax2(fmin2(q * fabs(a1), fabs(a2)), fmin2(fabs(a1), q * fabs(a2))), a1)
                 : zero;
    }
    case OSPRE:
      return dv(T(1.5) * a1 * a2 * (a1 + a2), a1 * (a1 + T(0.5) * a2) + a2 * (a2 + T(0.5) * a1) + tiny);
    case VANLEER: {
      const T b1 = fabs(a1), b2 = fabs(a2);
      return dv(a1 * b2 + a2 * b1, b1 + b2 + tiny);
    }
    default:
      return T(0.5) * (a1 + a2);
  }
}

// KT flux across one interface from the two face states (fluxes_kt.f90:126-176,
// physics_eulerisotherm.f90:555-686, 835-930, 2742-2817; physics_euler.f90:492-600)
// d = 0: x-interface (normal momentum index 1), d = 1: y-interface (index 2).
// kt_core takes both representations of the two face states.
template <typename T, int V>
__device__ __forceinline__ void kt_core(const Geom &g, int d, const T *UL, const T *UR, const T *pl, const T *pr, T csL,
                                        T csR, T dA, T *F) {
  const int un = 1 + d, ut = 2 - d;
  const T gam = (T)g.gamma, zero = T(0);
  if (V == 4) {
    csL = sound_speed(gam, pl[0], pl[3]);
    csR = sound_speed(gam, pr[0], pr[3]);
  }
  const T vL = pl[un], vR = pr[un];
  const T amin = fmin2(fmin2(zero, fmin2(zero, vR - csR)), fmin2(zero, vL - csL));
  const T amax = fmax2(fmax2(zero, fmax2(zero, vR + csR)), fmax2(zero, vL + csL));
  T fL[V], fR[V];
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
  const T fac = dv(dA, amax - amin);
#pragma unroll
  for (int l = 0; l < V; l++) F[l] = fac * (amax * fL[l] - amin * fR[l] + amin * amax * (UR[l] - UL[l]));
}
// conservative reconstruction: the face states are conservative, the
// primitive ones follow (Physics%Convert2Primitive, fluxes_base.f90:382-383)
template <typename T, int V>
__device__ __forceinline__ void kt_flux(const Geom &g, int d, const T *UL, const T *UR, T csL, T csR, T dA, T *F) {
  const T gam = (T)g.gamma;
  T pl[V], pr[V];
  c2p<T, V>(UL, pl, gam);
  c2p<T, V>(UR, pr, gam);
  kt_core<T, V>(g, d, UL, UR, pl, pr, csL, csR, dA, F);
}
// PRIMITIVE reconstruction: the face states are primitive, the conservative
// ones follow (Physics%Convert2Conservative, fluxes_base.f90:378-381)
template <typename T, int V>
__device__ __forceinline__ void kt_flux_p(const Geom &g, int d, const T *PL, const T *PR, T csL, T csR, T dA, T *F) {
  const T gam = (T)g.gamma;
  T ul[V], ur[V];
  p2c<T, V>(PL, ul, gam);
  p2c<T, V>(PR, ur, gam);
  kt_core<T, V>(g, d, ul, ur, PL, PR, csL, csR, dA, F);
}

// ------------------------------------------------------------ step control
// PI controller (timedisc_base.f90:759-814), used on the device only to stop
// a batch when the step could fall under dtlimit; the host redoes such steps
static __device__ double pi_dtnew(const Ctl *c, double maxerr, double dtold) {
  double dttmp = maxerr > 0. ? 0.9 * dtold * exp(-log(maxerr) / c->order) : 4. * dtold;
  double dtnew = c->maxerrold > 0. ? dttmp * exp(-c->beta * log(c->maxerrold / maxerr)) : dttmp;
  if (maxerr < 1.0 && 4. * dtold < dtnew) dtnew = 4. * dtold;
  return dtnew;
}
static __device__ void halt(Ctl *c, int reason) {
  c->halt = 1;
  c->reason = reason;
}
// A lower bound of pi_dtnew for an accepted step (0 <= err < 1) without
// exp/log: err^(-1/order) >= 1, and (maxerrold/err)^(-beta) >= min(1,
// err/maxerrold) for 0 <= beta <= 1; other beta take the exact form
static __device__ double pi_dtnew_lower(const Ctl *c, double err, double dtold) {
  if (!(c->beta >= 0.0 && c->beta <= 1.0 && c->order > 0.0)) return pi_dtnew(c, err, dtold);
  const double f = (c->maxerrold > 0. && err < c->maxerrold) ? err / c->maxerrold : 1.0;
  return 0.9 * dtold * f;
}
// after SolveODE: accept when the embedded error is finite and < 1
// (timedisc_base.f90:686-713, AcceptSolution 911-959); halt when the next
// step could fall under dtlimit (the host redoes such a step exactly)
static __device__ void control_accept(Ctl *c, double err) {
  const int s = c->step;
  c->err_hist[s] = err;
  c->dt_hist[s] = c->dt;
  if (!(err < 1.0)) return halt(c, H_ERR);  // NaN, Inf or a rejection
  if (pi_dtnew_lower(c, err, c->dt) < c->dtlimit * (1.0 + 1.0e-6)) return halt(c, H_DTLIMIT);
  c->maxerrold = err;
  const double t1 = c->time + c->dt;
  c->dt_step = t1 - c->time;
  c->time = t1;
}
// after the end-of-step RHS: data checks, output test, next CFL step,
// stop-time test and output-time clipping (fosite.f90:383-462,
// io/fileio_base.f90:604-620)
static __device__ void control_next(Ctl *c, double invdt, int flags, double invdt_v) {
  if (flags) return halt(c, H_FLAGS);
  const double t = c->time;
  if (fabs(c->t_out - t) <= 1.0E-5 * c->t_out) return halt(c, H_OUTPUT);
  double dt = c->cfl / invdt;
  if (c->vis) {  // the viscous limit (sources_viscosity.f90:507-577)
    const double dtv = invdt_v > DBL_MIN ? c->cvis / invdt_v : DBL_MAX;
    dt = dtv < dt ? dtv : dt;
    c->invdt_v = invdt_v;
  }
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

// ------------------------------------------------ boundary-flux accounting
// Element e of the boundary rows (west/east rows of every field first, then
// south/north): its new bxflux(j,side,l) / byflux(i,side,l) after one step
// (timedisc_rkfehlberg.f90:216-259), and in *bi its index in the layout.
__device__ __forceinline__ double bflux_elem(const BfluxArgs &b, double dt, int e, long *bi) {
  const int nx = b.V * 2 * b.jnum;
  int l, side, gx;
  if (e < nx) {  // west / east: ghost column IMIN-1 / IMAX+1 of row j
    l = e / (2 * b.jnum);
    side = (e / b.jnum) & 1;
    const int j = GH + e % b.jnum;
    *bi = (long)(l * 2 + side) * b.NJ + j;
    gx = j * b.P + (side ? GH + b.inum : GH - 1);
  } else {       // south / north: ghost row JMIN-1 / JMAX+1 of column i
    const int f = e - nx;
    l = f / (2 * b.inum);
    side = (f / b.inum) & 1;
    const int i = GH + f % b.inum;
    *bi = 2L * b.V * b.NJ + (long)(l * 2 + side) * b.NI + i;
    gx = (side ? GH + b.jnum : GH - 1) * b.P + i;
  }
  double kq[MAXK];
  const long kx = (long)l * b.FS + gx;
#pragma unroll
  for (int q = 0; q < MAXK; q++)  // all loads first
    if (q < b.nk)
      kq[q] = b.dtype == 1 ? (double)static_cast<const float *>(b.k[q])[kx] : static_cast<const double *>(b.k[q])[kx];
  double acc = b.in[*bi];
  const double ks = b.ks[*bi];
#pragma unroll
  for (int q = 0; q < MAXK; q++)
    if (q < b.nk) acc = acc - (dt * b.bh[q]) * kq[q];
  return acc - (dt * b.bh[b.nk]) * ks;
}
// the whole accumulation: threads t, t+nt, ... of the caller each take one
// boundary row element
__device__ __forceinline__ void bflux_accumulate(const BfluxArgs &b, double dt, int t, int nt) {
  const int n = b.V * 2 * (b.jnum + b.inum);
  for (int e = t; e < n; e += nt) {
    long bi;
    const double v = bflux_elem(b, dt, e, &bi);
    b.out[bi] = v;
  }
}

// ----------------------------------------------------- point-mass accretion
// sum0 + term(0) + ... + term(n-1) in index order: one FP64 chain.  Full
// chunks of PF terms are loaded before their adds (the adds then wait on the
// loads in issue order, not on each load), with no per-term bounds test;
// the remainder follows term by term.  n must be wave-uniform.
template <int PF, class F>
__device__ __forceinline__ double ordered_sum(double sum, int n, F term) {
  int k = 0;
  for (; k + PF <= n; k += PF) {
    double t[PF];
#pragma unroll
    for (int q = 0; q < PF; q++) t[q] = term(k + q);
#pragma unroll
    for (int q = 0; q < PF; q++) sum = sum + t[q];
  }
  for (; k < n; k++) sum = sum + term(k);
  return sum;
}

// SUM(row(0:n-1)) in index order, the density sum of GetBoundaryFlux
// (fluxes_base.f90:220-300), by all 256 threads of a workgroup: they stage
// the row in LDS (2048 values at a time, one memory latency for the meshes
// in use), then wave 0 adds it, every lane on the same broadcast values.
// The result is valid in wave 0.
constexpr int ACC_CHUNK = 2048;
__device__ __forceinline__ double ordered_row_sum(const double *row, int n, int tid) {
  __shared__ double srow[ACC_CHUNK];
  double sum = 0.0;
  for (int c0 = 0; c0 < n; c0 += ACC_CHUNK) {
    const int m = min(ACC_CHUNK, n - c0);
    for (int k = tid; k < m; k += 256) srow[k] = row[c0 + k];
    __syncthreads();
    if (tid < 64) sum = ordered_sum<16>(sum, __builtin_amdgcn_readfirstlane(m), [&](int k) { return srow[k]; });
    __syncthreads();
  }
  return sum;
}

// UpdateGravity_single with accretion (gravity_pointmass.f90:383-428; no
// switch-on phase) given the boundary's density sum: header of gin -> gout,
// acceleration rings rescaled by new/old mass (lanes lane, lane+nl, ...)
__device__ __forceinline__ void accrete_update(const AccreteArgs &p, double sum, int lane, int nl) {
  const double oldmass = p.gin[PM_MASS];
  const double dt = p.dt_ctl ? p.ctl->dt : p.dt;
  double accrate = p.zero_accrate ? 0.0 : p.gin[PM_ACCRATE], massloss = p.gin[PM_MASSLOSS];
  double dmass = sum - p.gin[PM_MDOT];
  if (p.acclimit >= 0.) {
    const double lim = p.acclimit * dt * oldmass;
    const double dmasslim = (lim < dmass) ? lim : dmass;  // MIN(dmass, lim)
    massloss = massloss + dmass - dmasslim;
    dmass = dmasslim;
  }
  const double mass = oldmass + dmass;
  if (dt > 0.0) accrate = dmass / dt;
  const double f = mass / oldmass;
  if (lane == 0) {
    p.gout[PM_MASS] = mass;
    p.gout[PM_ACCRATE] = accrate;
    p.gout[PM_MASSLOSS] = massloss;
    p.gout[PM_MDOT] = sum;
  }
  for (int k = lane; k < 2 * p.NI; k += nl) {
    const double x = p.gin[PM_HDR + k] * f;
    p.gout[PM_HDR + k] = x;
    if (p.gfout) p.gfout[k] = (float)x;
  }
}

// ----------------------------------------------------------- tile orders
// Workgroup b of a launch is dispatched to XCD b mod 8.  A 1-D grid of
// 8*Q blocks (Q = ceil(tiles/8)) whose block b runs tile (b mod 8)*Q + b/8
// gives every XCD a contiguous, row-major run of tiles, so tiles that share
// halo rows share that XCD's L2.  Returns false for the padding blocks.
constexpr int XCDS = 8;
__device__ __forceinline__ bool xcd_tile(int b, int nbx, int nby, int &bx, int &by) {
  const int n = nbx * nby, q = (n + XCDS - 1) / XCDS, t = (b % XCDS) * q + b / XCDS;
  if (t >= n) return false;
  bx = t % nbx;
  by = t / nbx;
  return true;
}
inline int xcd_grid(int nbx, int nby) { return ((nbx * nby + XCDS - 1) / XCDS) * XCDS; }

// -------------------------------------------------------------- fused RHS
// Tile: TX x TY_RHS output cells (TX = 64 or 32), one thread each.  The
// footprint (2-cell halo, corners excluded) is staged in LDS as the state the
// fluxes see; threads with an extra job cover the halo cells and the faces
// and fluxes one cell beyond the tile.  Phases, separated by barriers:
//   1 stage state + transform of the own and one extra footprint cell (ghost
//     cells on boundary tiles), sources, data checks and CFL of the own cell
//   2 limited slopes and face states: own cell; rows -1 and TY; columns -1, TX
//   3 KT interface fluxes: own east and north interfaces; west of column 0,
//     south of row 0
//   4 rhs assembly, or the final RK combination in the last stage; stores
// Faces live in their own LDS region; the fluxes reuse the footprint region,
// which is dead after phase 2, so three barriers separate the four phases.
// Output-aligned tiles keep every lane of the flux phases busy and divide
// power-of-two meshes exactly.
//
// Occupancy: a CU has 160 KB of LDS, allocated in 512-byte granules.  The
// FP64 isothermal stage tiles (64 x 8, V = 3) take 54444 bytes, 107
// granules, so two tiles share a CU (4 waves per SIMD); the launch bounds
// compute that from the layout.  The ghost-cell CFL and flag arrays only
// exist in the end-of-step mode.
constexpr int TY_RHS = TY_TILE;
// launch modes of the RHS kernel
enum RhsMode { M_STAGE = 0, M_LAST = 1, M_END = 2 };
// LDS bytes of one tile (the layout rhs_tile declares)
template <typename T, int V, int TX>
constexpr int rhs_lds_words() {
  constexpr int TY = TY_RHS, SX = TX + 4, SY = TY + 4;
  constexpr int N_UC = V * SY * SX, N_FLUX = V * TY * (TX + 1) + V * (TY + 1) * TX;
  constexpr int N_FACE = V * TY * (TX + 1) + V * TY + V * (TY + 1) * TX + V * TX;
  return (N_UC > N_FLUX ? N_UC : N_FLUX) + N_FACE;
}
// LDS bytes of one tile, rounded to the 512-byte allocation granule
template <typename T, int V, int TX>
constexpr int rhs_lds_bytes() {
  return ((int)sizeof(T) * rhs_lds_words<T, V, TX>() + 12 * TY_RHS + 511) / 512 * 512;
}
// waves per SIMD the launch bounds ask for: the FP64 stage kernels of up to
// three stages fit 80 VGPRs without spilling, so they take as many waves as
// the tiles' LDS lets share a CU (160 KB), at most 6; everything else 4
template <typename T, int V, int TX, int KM, int MODE>
constexpr int rhs_waves_per_simd() {
#ifdef FV_RHS_WAVES_PER_SIMD
  return FV_RHS_WAVES_PER_SIMD;
#else
  if (!(sizeof(T) == 8 && V == 3 && KM <= 3 && MODE != M_END)) return 4;
  constexpr int w = (163840 / rhs_lds_bytes<T, V, TX>()) * (TX * TY_RHS / 64) / 4;
  return w < 4 ? 4 : (w > 6 ? 6 : w);
#endif
}
template <typename T>
inline int rhs_mode(const StageArgsT<T> &a) { return a.ynew ? M_LAST : (a.nk == 0 ? M_END : M_STAGE); }

// one TX x TY tile (bx, by) of a mesh tiled nbx x nby
template <typename T, int V, int LIM, int TX, int KM, int MODE>
__device__ __forceinline__ void rhs_tile(const Geom &g, const StageArgsT<T> &a, int bx, int by, int nbx,
                                         int nby) {
  // what this launch does, fixed at compile time (the run-time flags of the
  // arguments still apply within it): RK stage rhs, last stage with the
  // final combination, or the end-of-step evaluation (no stage sum; CFL,
  // data checks, new state, FARGO type-1 velocities, boundary fluxes)
  constexpr bool kEnd = MODE == M_END, kLast = MODE == M_LAST;
  constexpr int KE = kEnd ? 0 : KM;
  const bool want_cfl = kEnd && a.want_cfl, want_check = kEnd && a.want_check;
  T *const yout = kEnd ? a.yout : nullptr;
  T *const pout = kEnd ? a.pout : nullptr;
  double *const vout = kEnd ? a.vout : nullptr;
  T *const ynew = kLast ? a.ynew : nullptr;
  constexpr int TY = TY_RHS, SX = TX + 4, SY = TY + 4;
  // the threads with an extra footprint cell cover rows -2, -1, TY, TY+1 and
  // columns -2, -1, TX, TX+1
  static_assert(TX * TY >= 4 * TX + 4 * TY, "tile too small for its halo jobs");
  constexpr int N_UC = V * SY * SX;
  constexpr int XW = V * TY * (TX + 1), XE = V * TY, YS = V * (TY + 1) * TX, YN = V * TX;
  constexpr int N_FACE = XW + XE + YS + YN;
  constexpr int FX = V * TY * (TX + 1), FY = V * (TY + 1) * TX;
  constexpr int N_FLUX = FX + FY;
  // region A: footprint state (phases 1-2), then fluxes (phases 3-4);
  // region B: face states (phases 2-3)
  constexpr int NA = N_UC > N_FLUX ? N_UC : N_FLUX;
  __shared__ T lds[rhs_lds_words<T, V, TX>()];
  static_assert(NA + N_FACE == rhs_lds_words<T, V, TX>(), "LDS layout");
  __shared__ double redd[TY];
  __shared__ int redi[TY];
  // CFL term and check flags of the ghost cells a thread evaluated (kept in
  // LDS rather than registers across the interior path)
  __shared__ T gcfl[kEnd ? TX * TY : 1];
  __shared__ int gflg[kEnd ? TX * TY : 1];
  // footprint state, tile-relative row r in [-2, TY+1], column c in [-2, TX+1]
  auto UC = [&](int l, int r, int c) -> T & { return lds[(l * SY + r + 2) * SX + c + 2]; };
  // faces: W of columns 0..TX, E of column -1, S of rows 0..TY, N of row -1
  T *fb = lds + NA;
  auto FW = [&](int l, int r, int c) -> T & { return fb[(l * TY + r) * (TX + 1) + c]; };
  auto FE1 = [&](int l, int r) -> T & { return fb[XW + l * TY + r]; };
  auto FS = [&](int l, int r, int c) -> T & { return fb[XW + XE + (l * (TY + 1) + r) * TX + c]; };
  auto FN1 = [&](int l, int c) -> T & { return fb[XW + XE + YS + l * TX + c]; };
  // fluxes: x interface c-1/2 (c = 0..TX) of row r; y interface r-1/2 (r = 0..TY) of column c
  auto FLX = [&](int l, int r, int c) -> T & { return lds[(l * TY + r) * (TX + 1) + c]; };
  auto FLY = [&](int l, int r, int c) -> T & { return lds[FX + (l * (TY + 1) + r) * TX + c]; };

  if (a.ctl && a.ctl->halt) return;
  // PRIMITIVE reconstruction: the footprint and the face states hold pvar
  // (only the run-time-limiter instantiation carries this path)
  const bool prim_rec = LIM == 0 && g.recon_prim;
  const In<T, V, KE> in(g, a, want_check);
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + TX * ty;
  const int i0 = GH + bx * TX, j0 = GH + by * TY;
  const int imin = GH, jmin = GH, imax = GH + g.inum - 1, jmax = GH + g.jnum - 1;
  const int i = i0 + tx, j = j0 + ty;
  const bool out = i <= imax && j <= jmax;  // own cell is an interior cell
  const T zero = T(0);

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
    T ginv = -Real<T>::huge;
    int gf = 0;
#pragma unroll
    for (int p = 0; p < 2; p++) {
      if (!(p ? e_g : own_g)) continue;
      const int ii = p ? ei : i, jj = p ? ej : j;
      const bool xin = ii >= imin && ii <= imax, yin = jj >= jmin && jj <= jmax;
      T u[V], pp[V];
      if ((xin || yin) && ii >= 0 && jj >= 0 && ii < g.NI && jj < g.NJ) {
        in.ghost_final(ii, jj, u, pp, gf);
        if (want_cfl) {
          const T cs = (V == 4) ? sound_speed(in.gam, pp[0], pp[3]) : rg<T>(g, R_CS, ii);
          ginv = fmax2(ginv, cell_invdt(g, ii, pp[1], pp[2], cs));
        }
        if (yout) {
          const unsigned ix = (unsigned)jj * (unsigned)g.P + (unsigned)ii;
#pragma unroll
          for (int l = 0; l < V; l++) (yout + l * g.FS)[ix] = u[l];
        }
        if (pout) {
          const unsigned ix = (unsigned)jj * (unsigned)g.P + (unsigned)ii;
#pragma unroll
          for (int l = 0; l < V; l++) (pout + l * g.FS)[ix] = pp[l];
        }
        if (vout && yin) {  // x-ghost ring of an interior row: (v - w) + w
          const T w = rg<T>(g, R_W, ii), vc = a.sub_bg ? pp[2] : pp[2] - w;
          vout[(unsigned)jj * (unsigned)g.P + (unsigned)ii] = (double)(vc + w);
        }
      } else {
#pragma unroll
        for (int l = 0; l < V; l++) u[l] = pp[l] = zero;  // corner or beyond the mesh: never needed
      }
#pragma unroll
      for (int l = 0; l < V; l++) UC(l, p ? er : ty, p ? ec : tx) = prim_rec ? pp[l] : u[l];
    }
    if constexpr (kEnd) {
      gcfl[t] = ginv;
      gflg[t] = gf;
    }
  }

  T ur[V], pr[V];
  T vres = zero;  // FP32: azimuthal velocity of the stored state (residual in the FARGO frame)
  T ucons[V];     // PRIMITIVE reconstruction: the own cell's cvar for the new state (the footprint holds pvar)
  {
    T so[V], se[V], uo[V], ue[V], re[V], pe[V];
    in.interior2(out, i, j, so, e_in, ei, ej, se);
    if (out) {
      in.transform(i, so, ur, pr, uo);
      if constexpr (sizeof(T) == 4) vres = dv(so[2], so[0]);
      if (vout) {  // pvar after SubtractBackgroundVelocityY, then + w again
        const T w = rg<T>(g, R_W, i);
        vout[(unsigned)j * (unsigned)g.P + (unsigned)i] = (double)((pr[2] - w) + w);
      }
      if (pout) {  // pvar after SubtractBackgroundVelocityY
        const unsigned ix = (unsigned)j * (unsigned)g.P + (unsigned)i;
#pragma unroll
        for (int l = 0; l < V; l++) (pout + l * g.FS)[ix] = (l == 2 && a.sub_bg) ? pr[2] - rg<T>(g, R_W, i) : pr[l];
      }
      if (prim_rec) {
#pragma unroll
        for (int l = 0; l < V; l++) ucons[l] = uo[l];
        in.prim_comoving(i, so, pr, uo);
      }
#pragma unroll
      for (int l = 0; l < V; l++) UC(l, ty, tx) = uo[l];
    }
    if (e_in) {
      in.transform(ei, se, re, pe, ue);
      if (prim_rec) in.prim_comoving(ei, se, pe, ue);
#pragma unroll
      for (int l = 0; l < V; l++) UC(l, er, ec) = ue[l];
    }
  }
  // sources, data checks and CFL of the own interior cell
  T geo[V], ext[V];
  T invdt = -Real<T>::huge;
  int flags = 0;
#pragma unroll
  for (int l = 0; l < V; l++) geo[l] = ext[l] = zero;
  if (out) {
    // geometrical sources (physics_eulerisotherm.f90:689-832, elementals 3069-3088;
    // physics_euler.f90:1706-1846); zero on cartesian meshes
    if (g.curvilinear) {
      const T cxyx = rg<T>(g, R_CXYX, i), cyxy = rg<T>(g, R_CYXY, i);
      const T czxz = rg<T>(g, R_CZXZ, i), czyz = rg<T>(g, R_CZYZ, i);
      const T vx = pr[1], vy = pr[2], vz = zero, mx = ur[1], my = ur[2], mz = zero;
      T P;
      if (V == 4) {
        P = pr[3];
      } else {
        const T cs = rg<T>(g, R_CS, i);
        P = pr[0] * (cs * cs);
      }
      if constexpr (sizeof(T) == 4) {
        // FP32: the radial momentum source is the small difference of the
        // centrifugal force and gravity (relative size ~(cs/vK)^2 on a disk);
        // both are formed in FP64 from the stored state, with the rotating
        // frame velocity added exactly, and only their sum is rounded
        const double vyd = a.add_bg ? (double)vres + ring(g, R_W, i) : (double)vy;
        const double rho = (double)ur[0], myd = rho * vyd;
        const double cyd = ring(g, R_CYXY, i);
        double s1 = myd * (cyd * vyd - ring(g, R_CXYX, i) * (double)vx) + (cyd + ring(g, R_CZXZ, i)) * (double)P;
        if (g.gravity) s1 = s1 + rho * a.acc[i];
        if (g.rotframe) {  // the rotating frame's radial acceleration joins the balance
          const double ax = g.rotframe == 2 ? g.twoomega * vyd : ring(g, R_CACX, i) + g.twoomega * vyd;
          s1 = s1 + rho * ax;
        }
        geo[1] = (T)s1;
      } else {
        geo[1] = my * (cyxy * vy - cxyx * vx) + mz * (czxz * vz - zero * vx) + (cyxy + czxz) * P;
      }
      geo[2] = mz * (czyz * vz - zero * vy) + mx * (cxyx * vx - cyxy * vy) + (cxyx + czyz) * P;
    }
    // FARGO source (physics_eulerisotherm.f90:1531-1586); Euler also takes
    // -v_y*(that + GetGeometricalSourceY(vz = P = mz = 0)) from the energy
    // (physics_euler.f90:2368-2400, elemental physics_eulerisotherm.f90:
    // 3080-3088), with the real-frame velocities
    if (g.fargo) {
      T tmp = dv(ur[1] * T(0.5) * (rg<T>(g, R_W, i + 1) - rg<T>(g, R_W, i - 1)), rg<T>(g, R_DLX, i));
      geo[2] = geo[2] - tmp;
      if constexpr (V == 4) {
        const T vx = pr[1], vy = pr[2], mx = ur[1];
        T cxyx = zero, cyxy = zero, czyz = zero;
        if (g.curvilinear) cxyx = rg<T>(g, R_CXYX, i), cyxy = rg<T>(g, R_CYXY, i), czyz = rg<T>(g, R_CZYZ, i);
        tmp = tmp + (zero * (czyz * zero - zero * vy) + mx * (cxyx * vx - cyxy * vy) + (cxyx + czyz) * zero);
        geo[3] = geo[3] - vy * tmp;
      }
    }
    // point-mass gravity through the source list (sources_generic.f90:220-261,
    // sources_gravity.f90:164-199, physics_eulerisotherm.f90:1201-1226,
    // physics_euler.f90:1849-1885)
    if (g.gravity) {
      const T *acc = sizeof(T) == 4 ? reinterpret_cast<const T *>(a.accf) : reinterpret_cast<const T *>(a.acc);
      const T ax = acc[(unsigned)i], ay = acc[(unsigned)(g.NI + i)];
      // FP32 curvilinear: already part of the FP64 radial balance above
      if (sizeof(T) == 8 || !g.curvilinear) ext[1] = zero + ur[0] * ax;
      ext[2] = zero + ur[0] * ay;
      if (V == 4) ext[3] = zero + (ur[1] * ax + ur[2] * ay);
    }
    // rotating frame, after gravity in the source list (sources_generic.f90:
    // 133-199, 236-241; sources_rotframe.f90:226-234): centrifugal plus
    // Coriolis acceleration of the real-frame velocities, through the same
    // ExternalSources as gravity
    if (g.rotframe) {
      const T tw = (T)g.twoomega, vx = pr[1], vy = pr[2];
      T ax, ay;
      if (g.rotframe == 2) {
        ax = tw * vy;
        ay = -(tw * vx);
      } else {
        ax = rg<T>(g, R_CACX, i) + tw * vy;
        ay = cl<T>(g, C_CACY, j) - tw * vx;
      }
      // FP32 curvilinear: the radial part is in the FP64 radial balance above
      if (sizeof(T) == 8 || !g.curvilinear) ext[1] = ext[1] + ur[0] * ax;
      ext[2] = ext[2] + ur[0] * ay;
      if (V == 4) ext[3] = ext[3] + (ur[1] * ax + ur[2] * ay);
    }
    // CheckData (timedisc_base.f90:1666-1742) and CFL (timedisc_base.f90:818-909)
    if (want_check) {
      if (pr[0] < (T)g.rhomin) flags |= F_RHOMIN;
      if (V == 4 && pr[3] < (T)g.pmin) flags |= F_PMIN;
#pragma unroll
      for (int l = 0; l < V; l++)
        if (bad(pr[l]) || bad(ur[l])) flags |= F_INVALID;
    }
    if (want_cfl) {
      const T vy = a.sub_bg ? pr[2] - rg<T>(g, R_W, i) : pr[2];
      const T cs = (V == 4) ? sound_speed(in.gam, pr[0], pr[3]) : rg<T>(g, R_CS, i);
      invdt = cell_invdt(g, i, pr[1], vy, cs);
    }
  }
  __syncthreads();

  // ---- phase 2: limited slopes and face states (reconstruction_linear.f90)
  const int ir = i < g.NI - 1 ? i : g.NI - 2;  // ring index, kept inside the mesh arrays
  const int jr = j < g.NJ ? j : g.NJ - 1;
  const T invdx = (T)g.invdx, invdy = (T)g.invdy;
  T uc[V], fe[V], fn[V];
  {
    const T dw = rg<T>(g, R_RDXW, ir), de = rg<T>(g, R_RDXE, ir);
    const T ds = cl<T>(g, C_RDYS, jr), dn = cl<T>(g, C_RDYN, jr);
#pragma unroll
    for (int l = 0; l < V; l++) {
      uc[l] = UC(l, ty, tx);
      const T sx = invdx * slope<T, LIM>(g, UC(l, ty, tx - 1), uc[l], UC(l, ty, tx + 1));
      FW(l, ty, tx) = uc[l] + sx * dw;
      fe[l] = uc[l] + sx * de;
      const T sy = invdy * slope<T, LIM>(g, UC(l, ty - 1, tx), uc[l], UC(l, ty + 1, tx));
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
    const T off = xk ? cl<T>(g, C_RDYS, jc) : cl<T>(g, C_RDYN, jc);
#pragma unroll
    for (int l = 0; l < V; l++) {
      const T c0 = UC(l, xr, xc);
      const T f = c0 + invdy * slope<T, LIM>(g, UC(l, xr - 1, xc), c0, UC(l, xr + 1, xc)) * off;
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
    const T off = xk ? rg<T>(g, R_RDXW, ic) : rg<T>(g, R_RDXE, ic);
#pragma unroll
    for (int l = 0; l < V; l++) {
      const T c0 = UC(l, xr, xc);
      const T f = c0 + invdx * slope<T, LIM>(g, UC(l, xr, xc - 1), c0, UC(l, xr, xc + 1)) * off;
      if (xk)
        FW(l, xr, TX) = f;
      else
        FE1(l, xr) = f;
    }
  }
  __syncthreads();

  // ---- phase 3: KT interface fluxes (fluxes_kt.f90:126-176)
  auto kt = [&](int d, const T *L, const T *R, T csL, T csR, T dA, T *F) {
    if (prim_rec)
      kt_flux_p<T, V>(g, d, L, R, csL, csR, dA, F);
    else
      kt_flux<T, V>(g, d, L, R, csL, csR, dA, F);
  };
  T xf[V], yf[V];
  {
    T r2[V], n2[V];
#pragma unroll
    for (int l = 0; l < V; l++) {
      r2[l] = FW(l, ty, tx + 1);
      n2[l] = FS(l, ty + 1, tx);
    }
    kt(0, fe, r2, rg<T>(g, R_CSE, ir), rg<T>(g, R_CSW, ir + 1), rg<T>(g, R_DAXW, ir + 1), xf);
    kt(1, fn, n2, rg<T>(g, R_CSN, ir), rg<T>(g, R_CSS, ir), rg<T>(g, R_DAYS, ir), yf);
#pragma unroll
    for (int l = 0; l < V; l++) {
      FLX(l, ty, tx + 1) = xf[l];
      FLY(l, ty + 1, tx) = yf[l];
    }
  }
  // extra fluxes: south interface of row 0 (t < TX), west interface of
  // column 0 (threads 2*TX .. 2*TX+TY-1)
  if (t < TX) {
    T s1[V], n1[V], ef[V];
#pragma unroll
    for (int l = 0; l < V; l++) {
      n1[l] = FN1(l, tx);
      s1[l] = FS(l, 0, tx);
    }
    kt(1, n1, s1, rg<T>(g, R_CSN, ir), rg<T>(g, R_CSS, ir), rg<T>(g, R_DAYS, ir), ef);
#pragma unroll
    for (int l = 0; l < V; l++) FLY(l, 0, tx) = ef[l];
  } else if (t >= 2 * TX && t < 2 * TX + TY) {
    const int fr = t - 2 * TX;
    T e1[V], w1[V], ef[V];
#pragma unroll
    for (int l = 0; l < V; l++) {
      e1[l] = FE1(l, fr);
      w1[l] = FW(l, fr, 0);
    }
    kt(0, e1, w1, rg<T>(g, R_CSE, i0 - 1), rg<T>(g, R_CSW, i0), rg<T>(g, R_DAXW, i0), ef);
#pragma unroll
    for (int l = 0; l < V; l++) FLX(l, fr, 0) = ef[l];
  }
  __syncthreads();

  // ---- phase 4: rhs assembly (timedisc_base.f90:1254-1271)
  double err = -DBL_MAX;
  if (out) {
    const T dydzdV = rg<T>(g, R_DYDZDV, i), dzdxdV = rg<T>(g, R_DZDXDV, i);
    const T dxdydV = rg<T>(g, R_DXDYDV, i);
    const unsigned ix = (unsigned)j * (unsigned)g.P + (unsigned)i;
    T rhs[V];
    // the viscous source joins the source list's sum last (sources_generic.f90:
    // 236-241); it has no density component
    if constexpr (sizeof(T) == 8) {
      if (a.vsrc) {
#pragma unroll
        for (int l = 1; l < V; l++) ext[l] = ext[l] + (a.vsrc + (l - 1) * g.FS)[ix];
      }
    }
#pragma unroll
    for (int l = 0; l < V; l++) {
      const T xw = FLX(l, ty, tx), ys = FLY(l, ty, tx);
      rhs[l] = dydzdV * (xf[l] - xw) + dzdxdV * (yf[l] - ys) + dxdydV * (zero - zero) - geo[l] - ext[l];
      if (yout) (yout + l * g.FS)[ix] = prim_rec ? ucons[l] : uc[l];
    }
    // boundary rows: the fluxes through the mesh boundary at the boundary
    // cells' outer interfaces, Mesh%dy*Mesh%dz*xflux (W), -Mesh%dy*Mesh%dz*xflux
    // (E) and likewise in y (timedisc_base.f90:1224-1252), stored at the ghost
    // position next to the cell (the last stage: into bks).  bxs/bys: the x
    // and y side of a boundary cell (0 W/S, 1 E/N, -1 none).
    const int bxs = !a.bflux ? -1 : i == imin ? 0 : i == imax ? 1 : -1;
    const int bys = !a.bflux ? -1 : j == jmin ? 0 : j == jmax ? 1 : -1;
    const unsigned gxx = (unsigned)j * (unsigned)g.P + (unsigned)(bxs == 0 ? imin - 1 : imax + 1);
    const unsigned gxy = (unsigned)(bys == 0 ? jmin - 1 : jmax + 1) * (unsigned)g.P + (unsigned)i;
    const long bix = (long)bxs * g.NJ + j, biy = 2L * V * g.NJ + (long)bys * g.NI + i;
    T kbx[V], kby[V];
    if (bxs >= 0 || bys >= 0) {
      const T ax = (T)g.bdydz, ay = (T)g.bdzdx;
#pragma unroll
      for (int l = 0; l < V; l++) {
        kbx[l] = bxs == 0 ? ax * FLX(l, ty, tx) : -(ax * xf[l]);
        kby[l] = bys == 0 ? ay * FLY(l, ty, tx) : -(ay * yf[l]);
        if (!ynew) {
          if (bxs >= 0) (a.kout + l * g.FS)[gxx] = kbx[l];
          if (bys >= 0) (a.kout + l * g.FS)[gxy] = kby[l];
        }
      }
    }
    if (ynew) {
      // final combination with this stage's rhs as k_s, and the error of the
      // embedded solution (FP64: same operation order as the reference;
      // FP32: compensated sums, error norm in FP64)
      T y[V], kq[KM][V];
#pragma unroll
      for (int l = 0; l < V; l++) y[l] = (a.Y + l * g.FS)[ix];
#pragma unroll
      for (int q = 0; q < KM; q++)
        if (q < in.nk) {
#pragma unroll
          for (int l = 0; l < V; l++) kq[q][l] = (a.k[q] + l * g.FS)[ix];
        }
      // FP32: weights b_h*dt as hi/lo pairs, and (b_h - b_l)*dt for the error
      float bhh[KM], bhl[KM], bde[KM];
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < KM; q++) {
          const double bh = a.bh[q] * in.dt;
          bhh[q] = (float)bh;
          bhl[q] = (float)(bh - (double)bhh[q]);
          bde[q] = (float)((a.bh[q] - a.bl[q]) * in.dt);
        }
      }
#pragma unroll
      for (int l = 0; l < V; l++) {
        // k_1 .. k_{s-1} from memory, then this stage's rhs as k_s (every
        // coefficient index is a compile-time constant)
        double yh, dif;
        if constexpr (sizeof(T) == 8) {
          T h = y[l], w = y[l];
#pragma unroll
          for (int q = 0; q < KM; q++)
            if (q <= in.nk) {
              const T kk = q < in.nk ? kq[q][l] : rhs[l];
              w = w - (T)(a.bl[q] * in.dt) * kk;
              h = h - (T)(a.bh[q] * in.dt) * kk;
            }
          (ynew + l * g.FS)[ix] = h;
          yh = (double)h;
          dif = yh - (double)w;
        } else {
          // compensated accumulation of the new state; the difference of the
          // two embedded solutions summed directly (no cancellation)
          Acc h((float)y[l]);
          float e = 0.f;
#pragma unroll
          for (int q = 0; q < KM; q++)
            if (q <= in.nk) {
              const float kk = (float)(q < in.nk ? kq[q][l] : rhs[l]);
              h.sub(bhh[q], bhl[q], kk);
              e = fmaf(bde[q], kk, e);
            }
          (ynew + l * g.FS)[ix] = (T)h.value();
          yh = h.wide();
          dif = (double)e;
        }
        err = fmax2(err, fabs(dif) / (a.tol_rel * fabs(yh) + a.tol_abs[l]));
      }
      // this stage's boundary rows for the boundary-flux accumulation
#pragma unroll
      for (int l = 0; l < V; l++) {
        if (bxs >= 0) a.bks[bix + 2L * l * g.NJ] = (double)kbx[l];
        if (bys >= 0) a.bks[biy + 2L * l * g.NI] = (double)kby[l];
      }
    } else {
#pragma unroll
      for (int l = 0; l < V; l++) (a.kout + l * g.FS)[ix] = rhs[l];
    }
  }
  if (ynew) {
    double m = err;
    int unused = 0;
    blk_max_or0(m, unused, redd, redi);
    if (t == 0) {
      const int b = bx + by * nbx;
      a.part_c[2 * b] = m;
      a.part_c[2 * b + 1] = 0.0;
    }
  }

  if (!(want_cfl || want_check)) return;
  if (own_g || e_g) {  // ghost cells evaluated in phase 1
    if constexpr (kEnd) {
      invdt = fmax2(invdt, gcfl[t]);
      flags |= gflg[t];
    }
  }
  // ---- the corner ghost cells (in no footprint) of the corner tiles:
  // data checks, CFL and the new state (SetCornerEdges, boundary_generic.f90:434-517)
  if (t < 16) {
    const bool west = bx == 0, east = bx == nbx - 1;
    const bool south = by == 0, north = by == nby - 1;
    const int c = t / 4, ga = 1 + (t & 1), gb = 1 + ((t >> 1) & 1);
    const int ce = c & 1, cn = c >> 1;
    if ((ce ? east : west) && (cn ? north : south)) {
      const int gi = ce ? imax + ga : imin - ga, gj = cn ? jmax + gb : jmin - gb;
      T u[V], p[V];
      in.ghost_final(gi, gj, u, p, flags);
      if (want_cfl) {
        const T cs = (V == 4) ? sound_speed(in.gam, p[0], p[3]) : rg<T>(g, R_CS, gi);
        invdt = fmax2(invdt, cell_invdt(g, gi, p[1], p[2], cs));
      }
      if (yout) {
        const unsigned ix = (unsigned)gj * (unsigned)g.P + (unsigned)gi;
#pragma unroll
        for (int l = 0; l < V; l++) (yout + l * g.FS)[ix] = u[l];
      }
      if (pout) {
        const unsigned ix = (unsigned)gj * (unsigned)g.P + (unsigned)gi;
#pragma unroll
        for (int l = 0; l < V; l++) (pout + l * g.FS)[ix] = p[l];
      }
    }
  }
  // CFL partials in FP64 (the maximum of FP32 values is exact)
  double m = invdt == -Real<T>::huge ? -DBL_MAX : (double)invdt;
  int f = flags;
  if (ynew) __syncthreads();  // the reduction scratch is in use above
  blk_max_or0(m, f, redd, redi);
  if (t == 0) {
    const int b = bx + by * nbx;
    a.part[2 * b] = m;
    a.part[2 * b + 1] = (double)f;
  }
  // the step's boundary-flux accumulation (device-controlled steps), after
  // this tile's work so that it adds no register pressure to it
  if (kEnd && a.bfa.out) {
    const int nt = TX * TY;
    bflux_accumulate(a.bfa, in.dt, (bx + by * nbx) * nt + t, nbx * nby * nt);
  }
}

template <typename T, int V, int LIM, int TX, int KM, int MODE>
__global__ __launch_bounds__(TX *TY_RHS, (rhs_waves_per_simd<T, V, TX, KM, MODE>())) void k_rhs(Geom g, StageArgsT<T> a) {
  const int nbx = (g.inum + TX - 1) / TX, nby = (g.jnum + TY_RHS - 1) / TY_RHS;
  int bx, by;
  if (!xcd_tile(blockIdx.x, nbx, nby, bx, by)) return;
  rhs_tile<T, V, LIM, TX, KM, MODE>(g, a, bx, by, nbx, nby);
}

// Many meshes in one launch: block b runs the tile tiles[b] (TileCode) of
// the mesh whose Geom and arguments of this launch are tab[disk].
template <typename T, int V, int LIM, int TX, int KM, int MODE>
__global__ __launch_bounds__(TX *TY_RHS, (rhs_waves_per_simd<T, V, TX, KM, MODE>())) void k_rhs_multi(
    const RhsEntry *__restrict__ tab, const unsigned *__restrict__ tiles) {
  const unsigned e = tiles[blockIdx.x];
  if (e == NO_TILE) return;
  const RhsEntry &E = tab[tile_disk(e)];
  const StageArgsT<T> &a = reinterpret_cast<const StageArgsT<T> &>(E.a);
  const int nbx = (E.g.inum + TX - 1) / TX, nby = (E.g.jnum + TY_RHS - 1) / TY_RHS;
  rhs_tile<T, V, LIM, TX, KM, MODE>(E.g, a, tile_bx(e), tile_by(e), nbx, nby);
}

// ------------------------------------------------------------ FARGO remap
// FargoAdvectionY (timedisc_base.f90:1935-2069) of one interior cell from
// the combined state: per ring the shift d = w*dt/dly split into its nearest
// integer and a fraction, minmod-limited linear transport by the fraction,
// then CSHIFT(., -shift) over JMIN:JMAX.  Rows outside JMIN:JMAX come from
// the comoving-frame CenterBoundary (prim round trip, no frame change).
template <typename T, int V>
__device__ __forceinline__ void ycell(const Geom &g, const T *Y, int i, int jj, T *u) {
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
    T p[V];
    c2p<T, V>(u, p, (T)g.gamma);
    p2c<T, V>(p, u, (T)g.gamma);
  }
}

// one block (bx, by) of the remap: 64 rings x REMAP_ROWS*4 rows.  A thread
// takes REMAP_ROWS consecutive rows of its ring; while their source rows do
// not wrap around the ring, they share one window of REMAP_ROWS+4 source
// cells (each cell's arithmetic is unchanged), otherwise every row loads its
// own five.
template <typename T, int V>
__device__ __forceinline__ void remap_block(const Geom &g, const RemapArgsT<T> &a, int bx, int by) {
  if (a.ctl && a.ctl->halt) return;
  double dts = a.dt_step;
  if (a.ctl) {  // this%dt as AcceptSolution will set it
    const double t = a.ctl->time, dt = a.ctl->dt;
    dts = (t + dt) - t;
  }
  constexpr int RR = REMAP_ROWS;
  const int i = GH + bx * 64 + threadIdx.x, jb = (by * 4 + threadIdx.y) * RR;
  if (i >= GH + g.inum || jb >= g.jnum) return;
  const int jmin = GH, J = g.jnum;
  double dd = ring(g, R_W, i) * dts / ring(g, R_DLY, i);
  const double shift = round(dd);
  const T d = (T)(dd - shift);
  // source row of the first row: jmin + ((jb - shift) mod J); fmod is exact
  int j0 = jb - (int)fmod(shift, (double)J);
  if (j0 < 0) j0 += J;
  if (j0 >= J) j0 -= J;
  const int nr = J - jb < RR ? J - jb : RR;
  const T one = T(1), zero = T(0), half = T(0.5);
  // the limited transport of one row from its five source cells
  auto row = [&](const T (*v)[V], int j) {
    const long ix = cidx(g, i, j);
#pragma unroll
    for (int l = 0; l < V; l++) {
      T dq[4], dql[3];
#pragma unroll
      for (int q = 0; q < 4; q++) dq[q] = v[q + 1][l] - v[q][l];
#pragma unroll
      for (int q = 0; q < 3; q++)
        dql[q] = (fsign(one, dq[q]) * fsign(one, dq[q + 1]) > zero) ? fsign(fmin2(fabs(dq[q]), fabs(dq[q + 1])), dq[q])
                                                                   : zero;
      T f0, f1;  // flux(js-1), flux(js)
      if (d > zero) {
        f0 = v[1][l] + half * dql[0] * (one - d);
        f1 = v[2][l] + half * dql[1] * (one - d);
      } else {
        f0 = v[2][l] - half * dql[1] * (one + d);
        f1 = v[3][l] - half * dql[2] * (one + d);
      }
      a.out[l * g.FS + ix] = v[2][l] - d * (f1 - f0);
    }
  };
  if (j0 + nr <= J) {  // no wrap: one window
    T v[RR + 4][V];
#pragma unroll
    for (int q = 0; q < RR + 4; q++)
      if (q < nr + 4) ycell<T, V>(g, a.Y, i, jmin + j0 - 2 + q, v[q]);
#pragma unroll
    for (int r = 0; r < RR; r++)
      if (r < nr) row(v + r, jmin + jb + r);
  } else {
    for (int r = 0; r < nr; r++) {
      int jj = j0 + r;
      if (jj >= J) jj -= J;
      T v[5][V];
#pragma unroll
      for (int q = 0; q < 5; q++) ycell<T, V>(g, a.Y, i, jmin + jj - 2 + q, v[q]);
      row(v, jmin + jb + r);
    }
  }
}

template <typename T, int V>
__global__ __launch_bounds__(256) void k_remap(Geom g, RemapArgsT<T> a) {
  int bx, by;
  if (!xcd_tile(blockIdx.x, (g.inum + 63) / 64, remap_rows_blocks(g.jnum), bx, by)) return;
  remap_block<T, V>(g, a, bx, by);
}
template <typename T, int V>
__global__ __launch_bounds__(256) void k_remap_multi(const RemapEntry *__restrict__ tab,
                                                     const unsigned *__restrict__ tiles) {
  const unsigned e = tiles[blockIdx.x];
  if (e == NO_TILE) return;
  const RemapEntry &E = tab[tile_disk(e)];
  remap_block<T, V>(E.g, reinterpret_cast<const RemapArgsT<T> &>(E.a), tile_bx(e), tile_by(e));
}

// FARGO frame change on all cells (physics_eulerisotherm.f90:1263-1413):
// sign +1 adds rho*w to the azimuthal momentum, -1 subtracts it.  With w2
// (the w row of other ring arrays) both changes in one pass: + rho*w, then
// - rho*w2, rounded as the two passes round.  ctl: a device-controlled step,
// skipped once the round has halted.
template <typename T, int V>
__global__ __launch_bounds__(256) void k_frame(Geom g, const T *src, T *dst, int sign, const Ctl *ctl, const T *w2) {
  if (ctl && ctl->halt) return;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.NI || j >= g.NJ) return;
  const long ix = cidx(g, i, j);
  const T w = rg<T>(g, R_W, i);
  T u[V];
#pragma unroll
  for (int l = 0; l < V; l++) u[l] = src[l * g.FS + ix];
  if (sign > 0)
    frame_add<T, V>(u, w);
  else
    frame_sub<T, V>(u, w);
  if (w2) frame_sub<T, V>(u, w2[i]);
#pragma unroll
  for (int l = 0; l < V; l++) dst[l * g.FS + ix] = u[l];
}

// ------------------------------------------------------------- viscosity
// dynamic viscosity of one cell from the local angular velocity
// e_phi/r . v + Omega (sources_viscosity.f90:377-394): ALPHA (etafkt_alpha)
// alpha*rho*cs^2/|Omega_l|, BETA (etafkt_beta, 457-466) beta*rho*r^2*|Omega_l|
// with the bary-centre radius
__device__ __forceinline__ double alpha_dynvis(const VisArgs &va, long ix, long FS, double rho, double vx, double vy,
                                               double cs) {
  double om = va.ephir[ix] * vx;
  om = om + va.ephir[FS + ix] * vy;
  om = om + va.omega;
  if (va.model == VIS_BETA) return va.dynconst * rho * va.ephir[2 * FS + ix] * fabs(om);
  return va.dynconst * rho * (cs * cs) / fabs(om);
}
// Divergence of a vector field in 2-D curvilinear coordinates from its face
// averages (mesh_midpoint.f90:156-232, 586-630): arguments are the west,
// own, east x components and the south, own, north y components; the z terms
// of the reference's expression are zero and kept as such (they turn a -0
// into +0)
__device__ __forceinline__ double div_faces(const Geom &g, int i, double xw, double x0, double xe, double ys,
                                            double y0, double yn) {
  const double fw = 0.5 * (xw + x0), fe = 0.5 * (xe + x0), fs = 0.5 * (ys + y0), fn = 0.5 * (yn + y0);
  return ring(g, R_DYDZDV, i) * (ring(g, R_DAXE, i) * fe - ring(g, R_DAXW, i) * fw) +
         ring(g, R_DZDXDV, i) * (ring(g, R_DAYN, i) * fn - ring(g, R_DAYS, i) * fs) + 0.0 * (0.0 * 0.0 - 0.0 * 0.0) +
         0.0 * 0.0 + 0.0 * 0.0 - 0.0 * 0.0 - 0.0 * 0.0;
}

// the divergence part of k_visc for the interior cell (i, j) at (r, c) of
// the inner region
template <int V, int SX, int RX>
__device__ __forceinline__ void visc_divergence(const Geom &g, const VisArgs &va, int i, int j, int r, int c,
                                                double (*Fu)[SX], double (*Fv)[SX], double (*Txx)[RX],
                                                double (*Tyy)[RX], double (*Txy)[RX]) {
  const double dydzdV = ring(g, R_DYDZDV, i), dzdxdV = ring(g, R_DZDXDV, i);
  const double axw = ring(g, R_DAXW, i), axe = ring(g, R_DAXE, i), ays = ring(g, R_DAYS, i), ayn = ring(g, R_DAYN, i);
  const double cxyx = ring(g, R_CXYX, i), cyxy = ring(g, R_CYXY, i);
  const double xx0 = Txx[r][c], xy0 = Txy[r][c], yy0 = Tyy[r][c];
  const double sx = dydzdV * (axe * (0.5 * (Txx[r][c + 1] + xx0)) - axw * (0.5 * (Txx[r][c - 1] + xx0))) +
                    dzdxdV * (ayn * (0.5 * (Txy[r + 1][c] + xy0)) - ays * (0.5 * (Txy[r - 1][c] + xy0))) +
                    0.0 * (0.0 * 0.0 - 0.0 * 0.0) + cxyx * xy0 + 0.0 * 0.0 - cyxy * yy0 - 0.0 * 0.0;
  const double sy = dydzdV * (axe * (0.5 * (Txy[r][c + 1] + xy0)) - axw * (0.5 * (Txy[r][c - 1] + xy0))) +
                    dzdxdV * (ayn * (0.5 * (Tyy[r + 1][c] + yy0)) - ays * (0.5 * (Tyy[r - 1][c] + yy0))) +
                    0.0 * (0.0 * 0.0 - 0.0 * 0.0) + (-cxyx) * xx0 + 0.0 * 0.0 - (-cyxy) * xy0 - 0.0 * 0.0;
  const long ix = cidx(g, i, j);
  va.src[ix] = sx;
  va.src[g.FS + ix] = sy;
  if constexpr (V == 4) {
    // (T.v)_x and (T.v)_y at the cell (dr, dc) of the inner region
    auto tvx = [&](int dr, int dc) {
      return Fu[dr + 1][dc + 1] * Txx[dr][dc] + Fv[dr + 1][dc + 1] * Txy[dr][dc];
    };
    auto tvy = [&](int dr, int dc) {
      return Fu[dr + 1][dc + 1] * Txy[dr][dc] + Fv[dr + 1][dc + 1] * Tyy[dr][dc];
    };
    va.src[2 * g.FS + ix] =
        div_faces(g, i, tvx(r, c - 1), tvx(r, c), tvx(r, c + 1), tvy(r - 1, c), tvy(r, c), tvy(r + 1, c));
  }
}

// viscous 1/dt of one cell (sources_viscosity.f90:507-577, the 2-D x-y
// branch): (max(a+, -a-) + 2 nu/dl)/dl summed over both directions, the wave
// speeds of CalcTimestep, nu = eta/rho
__device__ __forceinline__ double visc_invdt(const Geom &g, int i, double rho, double u, double v, double cs,
                                             double eta) {
  const double mnx = fmin2(0.0, u - cs), mxx = fmax2(0.0, u + cs);
  const double mny = fmin2(0.0, v - cs), mxy = fmax2(0.0, v + cs);
  const double kin = eta / rho, dlx = ring(g, R_DLX, i), dly = ring(g, R_DLY, i);
  return (fmax2(mxx, -mnx) + 2 * kin / dlx) / dlx + (fmax2(mxy, -mny) + 2 * kin / dly) / dly;
}

// Viscous source of one RHS evaluation (sources_viscosity.f90:481-500;
// physics_eulerisotherm.f90:982-1072, physics_euler.f90:1887-1918), FP64:
// ExternalSources sees Timedisc%pvar after CenterBoundary and the PMIN clamp,
// in the real frame.  Tiles of VT_X x VT_Y cells cover the whole
// ghost-inclusive mesh, one thread per cell; LDS holds
//   density and velocities on the tile and a two-cell halo,
//   the dynamic viscosity and the stresses on the tile and a one-cell halo
// (the reference forms the stresses on every cell but the outermost ghost
// layer, and the divergence reads them one cell around each interior cell).
// Three phases: pvar (and, with an ALPHA update, the dynamic viscosity of
// every own cell into va.dyn), stresses, divergence into va.src.
constexpr int VT_X = 64, VT_Y = 8;
#ifndef FV_VISC_WAVES_PER_SIMD
#define FV_VISC_WAVES_PER_SIMD 1
#endif
template <int V, int KM>
__global__ __launch_bounds__(VT_X *VT_Y, FV_VISC_WAVES_PER_SIMD) void k_visc(Geom g, StageArgs a, VisArgs va) {
  constexpr int SX = VT_X + 4, SY = VT_Y + 4, RX = VT_X + 2, RY = VT_Y + 2, NT = VT_X * VT_Y;
  __shared__ double Fr[SY][SX], Fu[SY][SX], Fv[SY][SX];  // rho, vx, vy; [r + 2][c + 2]
  __shared__ double Dy[RY][RX], Txx[RY][RX], Tyy[RY][RX], Txy[RY][RX];  // [r + 1][c + 1]
  __shared__ double Pown[V == 4 ? VT_Y : 1][V == 4 ? VT_X : 1];  // Euler: pressure of the own cells
  __shared__ double red[VT_X * VT_Y / 64];
  if (a.ctl && a.ctl->halt) return;
  const In<double, V, KM> in(g, a, false);
  const int tx = threadIdx.x, ty = threadIdx.y, t = tx + VT_X * ty;
  const int i0 = blockIdx.x * VT_X, j0 = blockIdx.y * VT_Y;
  const int imin = GH, imax = GH + g.inum - 1, jmin = GH, jmax = GH + g.jnum - 1;
  const bool alpha = va.model == VIS_ALPHA || va.model == VIS_BETA;  // viscosity from the local rotation
  const bool euler_cs = V == 4 && va.model == VIS_ALPHA;               // Physics%bccsound of Euler physics
  // UpdateViscosity's time test (sources_viscosity.f90:371): decided by the
  // host, or here from the step's time and dt in device-controlled steps
  bool update = va.update;
  if (a.ctl) {
    const double tc = a.ctl->time + va.c_cur * a.ctl->dt, tp = a.ctl->time + va.c_prev * a.ctl->dt;
    update = tc != tp || tc == 0.0;
  }

  // ---- phase 1: pvar of the footprint; dynamic viscosity of the inner
  // region.  At most two footprint cells per thread: every load of both
  // (state, stage slopes, stored viscosity) is issued before any arithmetic.
  static_assert(SX * SY <= 2 * NT, "two footprint cells per thread");
  int fr[2], fc[2], fi[2], fj[2];
  bool inm[2], inr[2], inreg[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int q = t + h * NT;
    fr[h] = q / SX - 2;
    fc[h] = q % SX - 2;
    fi[h] = i0 + fc[h];
    fj[h] = j0 + fr[h];
    inm[h] = q < SX * SY && fi[h] >= 0 && fi[h] < g.NI && fj[h] >= 0 && fj[h] < g.NJ;
    inr[h] = inm[h] && fi[h] >= imin && fi[h] <= imax && fj[h] >= jmin && fj[h] <= jmax;
    inreg[h] = inm[h] && fr[h] >= -1 && fr[h] <= VT_Y && fc[h] >= -1 && fc[h] <= VT_X;
  }
  double dst[2] = {va.dynconst, va.dynconst};
  if (alpha && !update) {
#pragma unroll
    for (int h = 0; h < 2; h++)
      if (inreg[h]) dst[h] = va.dyn[cidx(g, fi[h], fj[h])];
  }
  double s[2][V];
  in.interior2(inr[0], fi[0], fj[0], s[0], inr[1], fi[1], fj[1], s[1]);
#pragma unroll
  for (int h = 0; h < 2; h++) {
    if (!inm[h]) continue;
    const int r = fr[h], c = fc[h], ii = fi[h], jj = fj[h];
    double p[V];
    if (inr[h]) {
      double ur[V], uc[V];
      in.transform(ii, s[h], ur, p, uc);
    } else {
      in.ghost_prim(ii, jj, p);
      if (V == 4 && a.pmin_clamp) p[3] = fmax2(p[3], g.pmin);
    }
    Fr[r + 2][c + 2] = p[0];
    Fu[r + 2][c + 2] = p[1];
    Fv[r + 2][c + 2] = p[2];
    const bool own = r >= 0 && r < VT_Y && c >= 0 && c < VT_X;
    if constexpr (V == 4) {
      if (va.part && own) Pown[r][c] = p[3];
      if (euler_cs && a.nk == 0 && own) va.cs_out[cidx(g, ii, jj)] = sound_speed(g.gamma, p[0], p[3]);
    }
    if (inreg[h]) {
      double d = dst[h];
      if (alpha && update) {
        const long ix = cidx(g, ii, jj);
        double cs = 0.0;  // (BETA does not use it)
        if constexpr (V == 4) {
          if (euler_cs) cs = a.nk == 0 ? sound_speed(g.gamma, p[0], p[3]) : va.cs_in[ix];
        } else {
          cs = ring(g, R_CS, ii);
        }
        d = alpha_dynvis(va, ix, g.FS, p[0], p[1], p[2], cs);
        if (r >= 0 && r < VT_Y && c >= 0 && c < VT_X) va.dyn[ix] = d;
      }
      Dy[r + 1][c + 1] = d;
    }
  }
  __syncthreads();

  // ---- phase 2: stresses (physics_eulerisotherm.f90:1017-1072, 2-D x-y) on
  // the inner region's cells that are not in the outermost ghost layer
#pragma unroll
  for (int q0 = 0; q0 < RX * RY; q0 += NT) {
    const int q = q0 + t;
    if (q >= RX * RY) break;
    const int r = q / RX - 1, c = q % RX - 1, ii = i0 + c, jj = j0 + r;
    if (ii < 1 || ii > g.NI - 2 || jj < 1 || jj > g.NJ - 2) continue;
    const int fr = r + 2, fc = c + 2;
    const double u0 = Fu[fr][fc], v0 = Fv[fr][fc];
    const double uw = Fu[fr][fc - 1], ue = Fu[fr][fc + 1], us = Fu[fr - 1][fc], un = Fu[fr + 1][fc];
    const double vw = Fv[fr][fc - 1], ve = Fv[fr][fc + 1], vs = Fv[fr - 1][fc], vn = Fv[fr + 1][fc];
    const double eta = Dy[r + 1][c + 1];
    const double bulk = alpha ? -(2. / 3.) * eta : va.bulkconst;
    const double dlx = ring(g, R_DLX, ii), dly = ring(g, R_DLY, ii);
    const double cxyx = ring(g, R_CXYX, ii), cyxy = ring(g, R_CYXY, ii);
    const double dv = bulk * div_faces(g, ii, uw, u0, ue, vs, v0, vn);
    Txx[r + 1][c + 1] = eta * ((ue - uw) / dlx + 2.0 * cxyx * v0) + dv;
    Tyy[r + 1][c + 1] = eta * ((vn - vs) / dly + 2.0 * cyxy * u0) + dv;
    Txy[r + 1][c + 1] = eta * (0.5 * ((ve - vw) / dlx + (un - us) / dly) - cxyx * u0 - cyxy * v0);
  }
  __syncthreads();

  // ---- phase 3: tensor divergence (mesh_midpoint.f90:279-335, btyx = btxy)
  // and, for Euler, the divergence of T.v (physics_euler.f90:1887-1918)
  const int i = i0 + tx, j = j0 + ty;
  if (i >= imin && i <= imax && j >= jmin && j <= jmax)
    visc_divergence<V, SX, RX>(g, va, i, j, ty + 1, tx + 1, Fu, Fv, Txx, Tyy, Txy);
  if (!va.part) return;
  // the viscous limit of this evaluation's pvar on the own cell, every cell of
  // the mesh (as k_visc_cfl on Timedisc%pvar: comoving velocities under FARGO)
  double m = -DBL_MAX;
  if (i < g.NI && j < g.NJ) {
    const double rho = Fr[ty + 2][tx + 2], u = Fu[ty + 2][tx + 2];
    const double v = a.sub_bg ? Fv[ty + 2][tx + 2] - ring(g, R_W, i) : Fv[ty + 2][tx + 2];
    double cs;
    if constexpr (V == 4)
      cs = sound_speed(g.gamma, rho, Pown[ty][tx]);
    else
      cs = ring(g, R_CS, i);
    m = visc_invdt(g, i, rho, u, v, cs, Dy[ty + 1][tx + 1]);
  }
  m = blk_max(m, red);
  if (t == 0) {
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    va.part[2 * b] = m;
    va.part[2 * b + 1] = 0.0;
  }
}

// Viscous time-step limit (sources_viscosity.f90:507-577, the 2-D x-y
// branch) on the stored pvar pv of the last evaluation, all cells:
// (max(a+, -a-) + 2 nu/dl)/dl summed over the directions, nu = eta/rho, with
// the wave speeds CalcTimestep uses (v + w when add_w).  va.update: the
// dynamic viscosity from pv itself (UpdateViscosity at the same time 0).
template <int V>
__global__ __launch_bounds__(256) void k_visc_cfl(Geom g, VisArgs va, const double *__restrict__ pv, int add_w,
                                                  double *part) {
  __shared__ double red[4];
  double m = -DBL_MAX;
  const long n = (long)g.NI * g.NJ;
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long)gridDim.x * 256) {
    const int i = (int)(k % g.NI), j = (int)(k / g.NI);
    const long ix = cidx(g, i, j);
    const double rho = pv[ix], u = pv[g.FS + ix];
    double v = pv[2 * g.FS + ix];
    if (add_w) v = v + ring(g, R_W, i);
    const double cs = (V == 4) ? sound_speed(g.gamma, rho, pv[3 * g.FS + ix]) : ring(g, R_CS, i);
    double d = va.dynconst;
    if (va.model == VIS_ALPHA || va.model == VIS_BETA)
      d = va.update ? alpha_dynvis(va, ix, g.FS, rho, u, v, cs) : va.dyn[ix];
    m = fmax2(m, visc_invdt(g, i, rho, u, v, cs, d));
  }
  m = blk_max(m, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = m;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// typed launchers, explicitly instantiated in fv_inst_*.hip (one translation
// unit per element type and kernel family, compiled in parallel)
template <typename T, int KM>
void rhs_single(const Geom &g, const StageArgsT<T> &a, int V, hipStream_t s);
template <typename T>
void remap_single(const Geom &g, const RemapArgsT<T> &a, int V, hipStream_t s);
template <typename T>
void frame_single(const Geom &g, const void *src, void *dst, int sign, int V, const Ctl *ctl, const void *w2,
                  hipStream_t s);
template <typename T, int KM>
void rhs_multi(const RhsEntry *tab, const unsigned *tiles, int n, int V, int limiter, int tx, int mode, hipStream_t s);
// coefficient slots the kernels are compiled for, by tableau size
constexpr int KM_SMALL = 3;
inline int km_for(int stages) { return stages <= KM_SMALL ? KM_SMALL : MAXK; }
template <typename T>
void remap_multi(const RemapEntry *tab, const unsigned *tiles, int n, int V, hipStream_t s);

}  // namespace fv
