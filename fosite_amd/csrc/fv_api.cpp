// fv_api.cpp -- C ABI and step control of the MI355X finite-volume step.
//
// The control flow is the reference's Timedisc (timedisc/timedisc_base.f90,
// timedisc/timedisc_rkfehlberg.f90) and driver (fosite.f90): the host keeps
// the scalars (time, dt, error history, FARGO frame flags of every state
// vector) and issues device work; all array work runs in HIP kernels on
// device-resident state.  Step control (accept/reject, PI controller,
// output-time clipping) runs on the host in FP64 with the C library's exp/log,
// as the reference does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fv_host.h"

using namespace fv;

namespace {

constexpr int NBUF = 4;
// timedisc_base.f90:193-203
constexpr int CHECK_NOTHING = 0, CHECK_ALL = ~0, CHECK_CSOUND = 1, CHECK_PMIN = 2;
constexpr int DTCAUSE_ERRADJ = -1, DTCAUSE_SMALLERR = -2, DTCAUSE_FILEIO = -4;
enum { RK_FEHLBERG = 2, CASH_KARP = 3, DORMAND_PRINCE = 4, SSPRK = 5 };

struct Tableau {
  int m;
  double a[MAXK][MAXK], bh[MAXK], bl[MAXK];
};

// Butcher tableaux: timedisc_ssprk.f90:93-126, timedisc_rkfehlberg.f90:305-335,
// timedisc_dormand_prince.f90:91-114 (same rational literals, FP64)
bool set_tableau(int method, int order, Tableau &t) {
  std::memset(&t, 0, sizeof(t));
  if (method == SSPRK && order == 3) {
    t.m = 3;
    t.a[1][0] = 1.0;
    t.a[2][0] = 0.25; t.a[2][1] = 0.25;
    double bh[] = {1. / 6., 1. / 6., 2. / 3.}, bl[] = {0.5, 0.5, 0.};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl));
  } else if (method == SSPRK && order == 5) {
    t.m = 5;
    double a[5][5] = {{0., 0., 0., 0., 0.}, {0.36717, 0., 0., 0., 0.}, {0.26802, 0.3172, 0., 0., 0.},
                      {0.11606, 0.13735, 0.18816, 0., 0.}, {0.11212, 0.13269, 0.18178, 0.4198, 0.}};
    for (int i = 0; i < 5; i++) for (int j = 0; j < 5; j++) t.a[i][j] = a[i][j];
    double bh[] = {0.17279, 0.094505, 0.12947, 0.29899, 0.30424};
    double bl[] = {0.12293, 0.31981, -0.15316, 0.31887, 0.39155};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl));
  } else if (method == RK_FEHLBERG && order == 3) {
    t.m = 3;
    t.a[1][0] = 0.5;
    t.a[2][0] = -1.0; t.a[2][1] = 2.0;
    double bh[] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}, bl[] = {0.0, 1.0, 0.0};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl));
  } else if (method == RK_FEHLBERG && order == 5) {
    t.m = 6;
    double a[6][6] = {{0, 0, 0, 0, 0, 0},
                      {0.25, 0, 0, 0, 0, 0},
                      {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0, 0},
                      {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0, 0},
                      {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0, 0},
                      {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0, 0}};
    for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) t.a[i][j] = a[i][j];
    double bh[] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
    double bl[] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -0.2, 0.0};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl));
  } else if (method == DORMAND_PRINCE && order == 5) {
    t.m = 7;
    double a[7][7] = {{0, 0, 0, 0, 0, 0, 0},
                      {.2, 0, 0, 0, 0, 0, 0},
                      {3. / 40., 9. / 40., 0, 0, 0, 0, 0},
                      {44. / 45., -56. / 15., 32. / 9., 0, 0, 0, 0},
                      {19372. / 6561., -25360. / 2187., 64448. / 6561., -212. / 729., 0, 0, 0},
                      {9017. / 3168., -355. / 33., 46732. / 5247., 49. / 176., -5103. / 18656., 0, 0},
                      {35. / 384., 0., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0}};
    for (int i = 0; i < 7; i++) for (int j = 0; j < 7; j++) t.a[i][j] = a[i][j];
    double bh[] = {35. / 384., 0., 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0.};
    double bl[] = {5179. / 57600., 0., 7571. / 16695., 393. / 640., -92097. / 339200., 187. / 2100., 1. / 40.};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl));
  } else {
    return false;
  }
  return true;
}

}  // namespace

struct fv_ctx {
  fv_config c{};
  std::string err;
  int V = 0, NI = 0, NJ = 0, P = 0;
  long FS = 0;
  Tableau tab{};
  Mesh mesh;
  std::vector<double> ring_h, col_h;
  double *d_ring = nullptr, *d_col = nullptr;
  double *buf[NBUF] = {};
  int cvar = 0, cold = 0;
  double *W = nullptr;
  double *k[MAXK] = {};
  double *part = nullptr, *d_red = nullptr, *h_red = nullptr;
  int nb_rhs = 0, nb_ghost = 0, nb_comb = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Geom g{};
  // FARGO frame flags (fargo_transformation_applied) of each state vector
  int fl_pvar = 0, fl_cvar = 0, fl_ptmp = 0, fl_ctmp = 0, fl_cold = 0;
  double time = 0, dt = 0, dtold = 0, dtmin = 0, maxerrold = 0, t_out = 0, dtlimit = 0;
  int iter = -1, n_adj = 0, brk = 0, dtcause = 0;
  bool have_cs = false, have_w = false, have_state = false;
  double cs_min = 0;
  double last_ms = 0;
  double invdt = 0;
  int flags = 0;
};

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);               \
      return -3;                                                               \
    }                                                                          \
  } while (0)

namespace {

int fail(fv_ctx *ctx, const std::string &m, int code = -1) {
  ctx->err = m;
  return code;
}

int pick_free(const fv_ctx *ctx, int a, int b, int c) {
  for (int q = 0; q < NBUF; q++)
    if (q != a && q != b && q != c) return q;
  return -1;
}

// reduce the per-block partials and fetch {max, flags} to the host
int reduce(fv_ctx *ctx, int n, double *mx, int *flags) {
  launch_reduce(ctx->part, n, ctx->d_red, ctx->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ctx->h_red, ctx->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  *mx = ctx->h_red[0];
  *flags = (int)ctx->h_red[1];
  return 0;
}

StageArgs base_args(fv_ctx *ctx) {
  StageArgs a;
  std::memset(&a, 0, sizeof(a));
  a.part = ctx->part;
  return a;
}

// ComputeRHS (timedisc_base.f90:1073-1271) on the main state (pvar, cvar):
// writes coeff(1) and, when the frame/PMIN round trip changes cvar, the new
// cvar into a fresh buffer.  Also yields the CFL inverse time step and the
// CheckData flags (timedisc_base.f90:1666-1742).
int rhs_main(fv_ctx *ctx, int checkbm) {
  const bool fargo = ctx->c.fargo == 2;
  int add = 0, sub = 0;
  if (fargo && ctx->fl_cvar && ctx->fl_pvar) {
    add = 1;
    ctx->fl_cvar = ctx->fl_pvar = 0;
  }
  ctx->fl_pvar = ctx->fl_cvar;  // CenterBoundary
  const int pmin = (ctx->V == 4 && (checkbm & CHECK_PMIN)) ? 1 : 0;
  if (fargo && !ctx->fl_cvar && !ctx->fl_pvar) {
    sub = 1;
    ctx->fl_cvar = ctx->fl_pvar = 1;
  }
  const bool ident = !add && !sub && !pmin;
  const int outb = ident ? ctx->cvar : pick_free(ctx, ctx->cvar, ctx->cold, -1);
  StageArgs a = base_args(ctx);
  a.Y = ctx->buf[ctx->cvar];
  a.nk = 0;
  a.gsrc = a.gdst = ctx->buf[outb];
  a.kout = ctx->k[0];
  a.yout = ident ? nullptr : ctx->buf[outb];
  a.add_bg = add;
  a.sub_bg = sub;
  a.pmin_clamp = pmin;
  a.want_cfl = 1;
  a.want_check = checkbm != CHECK_NOTHING;
  a.part_off = ctx->nb_rhs;
  launch_ghost(ctx->g, a, ctx->V, ctx->st);
  HIPCHK(hipGetLastError());
  a.part_off = 0;
  launch_rhs(ctx->g, a, ctx->V, ctx->st);
  HIPCHK(hipGetLastError());
  ctx->cvar = outb;
  int rc = reduce(ctx, ctx->nb_rhs + ctx->nb_ghost, &ctx->invdt, &ctx->flags);
  if (rc) return rc;
  if (checkbm != CHECK_NOTHING) {
    if ((checkbm & CHECK_CSOUND) && ctx->V == 3 && ctx->cs_min <= 0.) ctx->brk = 1;
    if (ctx->flags) ctx->brk = 1;
  }
  return 0;
}

// ComputeRHS on the stage state (ptmp, ctmp) with ctmp = cvar - sum a*dt*k
int rhs_stage(fv_ctx *ctx, int m, double dt) {
  const bool fargo = ctx->c.fargo == 2;
  int add = 0, sub = 0;
  if (fargo && ctx->fl_ctmp && ctx->fl_ptmp) {
    add = 1;
    ctx->fl_ctmp = ctx->fl_ptmp = 0;
  }
  ctx->fl_ptmp = ctx->fl_ctmp;
  if (fargo && !ctx->fl_ctmp && !ctx->fl_ptmp) {
    sub = 1;
    ctx->fl_ctmp = ctx->fl_ptmp = 1;
  }
  StageArgs a = base_args(ctx);
  a.Y = ctx->buf[ctx->cvar];
  a.nk = m;
  for (int q = 0; q < m; q++) {
    a.k[q] = ctx->k[q];
    a.c[q] = ctx->tab.a[m][q] * dt;
  }
  a.gsrc = a.gdst = ctx->W;
  a.kout = ctx->k[m];
  a.add_bg = add;
  a.sub_bg = sub;
  launch_ghost(ctx->g, a, ctx->V, ctx->st);
  HIPCHK(hipGetLastError());
  launch_rhs(ctx->g, a, ctx->V, ctx->st);
  HIPCHK(hipGetLastError());
  return 0;
}

// timedisc_base.f90:759-814
double adjust_timestep(fv_ctx *ctx, double maxerr, double dtold) {
  double dttmp, dtnew;
  if (ctx->c.tol_rel >= 1.0) return DBL_MAX;
  if (maxerr > 0.)
    dttmp = 0.9 * dtold * std::exp(-std::log(maxerr) / ctx->c.order);
  else
    dttmp = 4. * dtold;
  if (ctx->maxerrold > 0.)
    dtnew = dttmp * std::exp(-ctx->c.beta * std::log(ctx->maxerrold / maxerr));
  else
    dtnew = dttmp;
  if (maxerr < 1.0) {
    dtnew = (4. * dtold < dtnew) ? 4. * dtold : dtnew;
    if (dtnew < dtold) ctx->dtcause = DTCAUSE_SMALLERR;
  } else {
    if (dtnew > dtold) dtnew = dttmp;
    dtnew = (0.25 * dtold > dtnew) ? 0.25 * dtold : dtnew;
  }
  ctx->maxerrold = maxerr;
  return dtnew;
}

// timedisc_rkfehlberg.f90:170-268 (boundary-flux accounting omitted: it only
// feeds diagnostics and accretion, neither enabled)
int solve_ode(fv_ctx *ctx, double *dt, double *err) {
  for (int m = 1; m < ctx->tab.m; m++) {
    int rc = rhs_stage(ctx, m, *dt);
    if (rc) return rc;
  }
  ctx->fl_ctmp = ctx->fl_cvar;  // ctmp = cvar
  const int outb = pick_free(ctx, ctx->cvar, ctx->cold, -1);
  CombineArgs a;
  std::memset(&a, 0, sizeof(a));
  a.Y = ctx->buf[ctx->cvar];
  a.nk = ctx->tab.m;
  for (int q = 0; q < ctx->tab.m; q++) {
    a.k[q] = ctx->k[q];
    a.ch[q] = ctx->tab.bh[q] * *dt;
    a.cl[q] = ctx->tab.bl[q] * *dt;
  }
  a.ynew = ctx->buf[outb];
  a.tol_rel = ctx->c.tol_rel;
  for (int l = 0; l < 4; l++) a.tol_abs[l] = ctx->c.tol_abs[l];
  a.part = ctx->part;
  launch_combine(ctx->g, a, ctx->V, ctx->st);
  HIPCHK(hipGetLastError());
  ctx->cvar = outb;
  double e;
  int f;
  int rc = reduce(ctx, ctx->nb_comb, &e, &f);
  if (rc) return rc;
  if (ctx->c.tol_rel <= 1.0) {
    *err = e;
    *dt = adjust_timestep(ctx, e, *dt);
  }
  return 0;
}

double calc_timestep(fv_ctx *ctx) {
  ctx->dtcause = 0;
  return ctx->c.cfl / ctx->invdt;  // gravity: dt_src = HUGE (sources_c_accel.f90:125)
}

// out-of-place frame change of cvar (Physics%Add/SubtractBackgroundVelocityY)
int frame_cvar(fv_ctx *ctx, int sign) {
  const int outb = pick_free(ctx, ctx->cvar, ctx->cold, -1);
  launch_frame(ctx->g, ctx->buf[ctx->cvar], ctx->buf[outb], sign, ctx->V, ctx->st);
  HIPCHK(hipGetLastError());
  ctx->cvar = outb;
  return 0;
}
int add_bg_main(fv_ctx *ctx) {  // physics_eulerisotherm.f90:1263-1303
  if (!ctx->fl_cvar || !ctx->fl_pvar) return 0;
  int rc = frame_cvar(ctx, +1);
  ctx->fl_cvar = ctx->fl_pvar = 0;
  return rc;
}
int sub_bg_main(fv_ctx *ctx) {  // physics_eulerisotherm.f90:1373-1413
  if (ctx->fl_cvar || ctx->fl_pvar) return 0;
  int rc = frame_cvar(ctx, -1);
  ctx->fl_cvar = ctx->fl_pvar = 1;
  return rc;
}

// timedisc_base.f90:649-750
int integration_step(fv_ctx *ctx) {
  int rc;
  double err = 0.0, dtold, dt, time;
  if (ctx->c.fargo == 2 && (rc = sub_bg_main(ctx))) return rc;
  time = ctx->time;
  dt = ctx->dt;
  if (dt < ctx->dtmin && ctx->dtcause != DTCAUSE_FILEIO) ctx->dtmin = dt;
  while (!ctx->brk && time + dt <= ctx->time + ctx->dt) {
    dtold = dt;
    if ((rc = solve_ode(ctx, &dt, &err))) return rc;
    if (err != err || err > DBL_MAX) {
      ctx->brk = 1;
    } else if (err >= 1.0) {  // RejectSolution (timedisc_base.f90:961-1009)
      ctx->cvar = ctx->cold;
      ctx->fl_cvar = ctx->fl_cold;
      if ((rc = rhs_main(ctx, CHECK_NOTHING))) return rc;
      ctx->n_adj++;
      ctx->dtcause = DTCAUSE_ERRADJ;
      if (dt < ctx->dtmin) ctx->dtmin = dt;
    } else {  // AcceptSolution (timedisc_base.f90:911-959)
      time = time + dtold;
      if (time < ctx->time + ctx->dt) {
        if ((rc = rhs_main(ctx, CHECK_ALL))) return rc;
        ctx->cold = ctx->cvar;
        ctx->fl_cold = ctx->fl_cvar;
      }
      ctx->iter++;
    }
    if (dt < ctx->dtlimit) ctx->brk = 1;
  }
  if (!ctx->brk) {
    ctx->dt = time - ctx->time;
    ctx->time = time;
    ctx->dtold = dt;
    if (ctx->c.fargo == 2) {
      // CenterBoundary in the comoving frame, then FargoAdvectionY
      // (timedisc_base.f90:722-743, 1935-2069)
      StageArgs a = base_args(ctx);
      a.Y = ctx->buf[ctx->cvar];
      a.gsrc = a.gdst = ctx->buf[ctx->cvar];
      ctx->fl_pvar = ctx->fl_cvar;
      launch_ghost(ctx->g, a, ctx->V, ctx->st);
      HIPCHK(hipGetLastError());
      const int outb = pick_free(ctx, ctx->cvar, ctx->cold, -1);
      launch_remap(ctx->g, ctx->buf[ctx->cvar], ctx->buf[outb], ctx->dt, ctx->V, ctx->st);
      HIPCHK(hipGetLastError());
      ctx->cvar = outb;
      ctx->fl_pvar = ctx->fl_cvar;  // Convert2Primitive(cvar, pvar)
    }
    if ((rc = rhs_main(ctx, CHECK_ALL))) return rc;
    ctx->cold = ctx->cvar;
    ctx->fl_cold = ctx->fl_cvar;
  }
  return 0;
}

int upload_geom(fv_ctx *ctx) {
  const int NI = ctx->NI, NJ = ctx->NJ;
  HIPCHK(hipMemcpy(ctx->d_ring, ctx->ring_h.data(), sizeof(double) * R_COUNT * NI, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_col, ctx->col_h.data(), sizeof(double) * C_COUNT * NJ, hipMemcpyHostToDevice));
  return 0;
}

// fills ring[id] from a cell array that must depend on i only
bool ring_from(fv_ctx *ctx, int id, const double *cells) {
  const int NI = ctx->NI, NJ = ctx->NJ;
  for (int j = 1; j < NJ; j++)
    if (std::memcmp(cells + (size_t)j * NI, cells, sizeof(double) * NI)) return false;
  std::memcpy(&ctx->ring_h[(size_t)id * NI], cells, sizeof(double) * NI);
  return true;
}

}  // namespace

// ================================================================== C ABI

extern "C" int fv_create(fv_ctx **out, const fv_config *cfg) {
  if (!out || !cfg) return -1;
  *out = nullptr;
  fv_ctx *ctx = new (std::nothrow) fv_ctx();
  if (!ctx) return -1;
  const fv_config &c = *cfg;
  ctx->c = c;
  auto bail = [&](const char *m) {
    std::fprintf(stderr, "fv_create: %s\n", m);
    delete ctx;
    return -1;
  };
  if (c.physics != EULER_ISOTHERM && c.physics != EULER) return bail("physics must be EULER_ISOTHERM (16) or EULER (17)");
  if (c.geometry != CARTESIAN && c.geometry != LOGCYLINDRICAL) return bail("geometry must be CARTESIAN (1) or LOGCYLINDRICAL (31)");
  if (c.inum < 4 || c.jnum < 4) return bail("inum and jnum must be >= 4 (2-D meshes only)");
  if (c.recvars != 2) return bail("only CONSERVATIVE reconstruction (variables = 2) is supported");
  if (c.dtype != 0) return bail("only dtype 0 (FP64) is supported");
  if (c.fargo != 0 && c.fargo != 2) return bail("fargo must be 0 or 2");
  if (c.fargo == 2 && (c.physics != EULER_ISOTHERM || c.geometry != LOGCYLINDRICAL))
    return bail("FARGO requires isothermal physics on a logcylindrical mesh");
  switch (c.limiter) {
    case MINMOD: case MONOCENT: case SWEBY: case SUPERBEE: case OSPRE: case VANLEER: case NOLIMIT: break;
    default: return bail("unknown limiter");
  }
  for (int d = 0; d < 4; d++) {
    const int b = c.bc[d];
    if (b != NO_GRADIENTS && b != PERIODIC && b != ABSORBING) return bail("boundary must be NO_GRADIENTS, PERIODIC or ABSORBING");
    if (b == ABSORBING && (d >= 2 || c.physics != EULER_ISOTHERM))
      return bail("ABSORBING boundaries are supported in x for isothermal physics");
  }
  // boundary_generic.f90: a periodic boundary needs a periodic opposite
  if ((c.bc[0] == PERIODIC) != (c.bc[1] == PERIODIC) || (c.bc[2] == PERIODIC) != (c.bc[3] == PERIODIC))
    return bail("opposite boundary should be periodic");
  if (c.gravity && c.geometry != LOGCYLINDRICAL) return bail("point-mass gravity requires a logcylindrical mesh");
  if (!set_tableau(c.method, c.order, ctx->tab)) return bail("unsupported time integrator (method/order)");
  if (mesh_build(c, ctx->mesh)) return bail("mesh construction failed");
  ctx->V = c.physics == EULER ? 4 : 3;
  ctx->NI = ctx->mesh.NI;
  ctx->NJ = ctx->mesh.NJ;
  ctx->P = (ctx->NI + 15) / 16 * 16;
  ctx->FS = (long)ctx->NJ * ctx->P;
  // per-ring metric slices
  ctx->ring_h.assign((size_t)R_COUNT * ctx->NI, 0.0);
  ctx->col_h.assign((size_t)C_COUNT * ctx->NJ, 0.0);
  const struct { int r, m; } rm[] = {{R_DAXW, M_DAXW}, {R_DAYS, M_DAYS}, {R_DYDZDV, M_DYDZDV}, {R_DZDXDV, M_DZDXDV},
                                     {R_DXDYDV, M_DXDYDV}, {R_DLX, M_DLX}, {R_DLY, M_DLY}, {R_CXYX, M_CXYX},
                                     {R_CYXY, M_CYXY}, {R_CZXZ, M_CZXZ}, {R_CZYZ, M_CZYZ}, {R_RDXW, M_RDXW},
                                     {R_RDXE, M_RDXE}, {R_ACCX, M_ACCX}, {R_ACCY, M_ACCY}};
  for (auto &p : rm)
    if (!ring_from(ctx, p.r, ctx->mesh.a[p.m].data())) return bail("mesh metric depends on j");
  for (int j = 0; j < ctx->NJ; j++) {
    ctx->col_h[(size_t)C_RDYS * ctx->NJ + j] = ctx->mesh.a[M_RDYS][(size_t)j * ctx->NI];
    ctx->col_h[(size_t)C_RDYN * ctx->NJ + j] = ctx->mesh.a[M_RDYN][(size_t)j * ctx->NI];
    for (int i = 1; i < ctx->NI; i++)
      if (ctx->mesh.a[M_RDYS][(size_t)j * ctx->NI + i] != ctx->col_h[(size_t)C_RDYS * ctx->NJ + j])
        return bail("reconstruction offsets in y depend on i");
  }
  // device resources
  const size_t sbytes = sizeof(double) * ctx->V * ctx->FS;
  auto dalloc = [&](double **p, size_t b) { return hipMalloc((void **)p, b) == hipSuccess; };
  bool ok = hipStreamCreate(&ctx->st) == hipSuccess && hipEventCreate(&ctx->e0) == hipSuccess &&
            hipEventCreate(&ctx->e1) == hipSuccess;
  for (int q = 0; q < NBUF && ok; q++) ok = dalloc(&ctx->buf[q], sbytes) && hipMemset(ctx->buf[q], 0, sbytes) == hipSuccess;
  ok = ok && dalloc(&ctx->W, sbytes) && hipMemset(ctx->W, 0, sbytes) == hipSuccess;
  for (int q = 0; q < ctx->tab.m && ok; q++) ok = dalloc(&ctx->k[q], sbytes) && hipMemset(ctx->k[q], 0, sbytes) == hipSuccess;
  ok = ok && dalloc(&ctx->d_ring, sizeof(double) * R_COUNT * ctx->NI) && dalloc(&ctx->d_col, sizeof(double) * C_COUNT * ctx->NJ);
  Geom &g = ctx->g;
  g.NI = ctx->NI; g.NJ = ctx->NJ; g.P = ctx->P; g.FS = ctx->FS;
  g.inum = c.inum; g.jnum = c.jnum;
  g.invdx = ctx->mesh.invdx; g.invdy = ctx->mesh.invdy;
  g.gamma = c.gamma; g.pmin = DBL_MIN; g.rhomin = DBL_MIN;
  g.limiter = c.limiter; g.theta = c.theta;
  for (int d = 0; d < 4; d++) g.bc[d] = c.bc[d];
  g.fargo = c.fargo == 2; g.gravity = c.gravity != 0; g.curvilinear = c.geometry != CARTESIAN;
  ctx->nb_rhs = rhs_blocks(g);
  ctx->nb_ghost = ghost_blocks(g);
  ctx->nb_comb = combine_blocks(g);
  const int np = std::max(ctx->nb_rhs + ctx->nb_ghost, ctx->nb_comb);
  ok = ok && dalloc(&ctx->part, sizeof(double) * 2 * np) && dalloc(&ctx->d_red, sizeof(double) * 2);
  ok = ok && hipHostMalloc((void **)&ctx->h_red, sizeof(double) * 2) == hipSuccess;
  if (!ok) {
    fv_destroy(ctx);
    std::fprintf(stderr, "fv_create: device allocation failed\n");
    return -3;
  }
  g.ring = ctx->d_ring;
  g.col = ctx->d_col;
  if (upload_geom(ctx)) { fv_destroy(ctx); return -3; }
  // timedisc_base.f90:315-352 defaults
  ctx->time = 0.0;
  ctx->dt = c.stoptime;
  ctx->dtold = ctx->dtmin = c.stoptime;
  ctx->dtlimit = c.dtlimit > 0 ? c.dtlimit : DBL_EPSILON * c.stoptime;
  ctx->t_out = c.stoptime;  // one output at the stop time
  ctx->iter = -1;
  *out = ctx;
  return 0;
}

extern "C" int fv_destroy(fv_ctx *ctx) {
  if (!ctx) return 0;
  for (auto &b : ctx->buf) if (b) (void)hipFree(b);
  for (auto &b : ctx->k) if (b) (void)hipFree(b);
  if (ctx->W) (void)hipFree(ctx->W);
  if (ctx->part) (void)hipFree(ctx->part);
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  if (ctx->d_ring) (void)hipFree(ctx->d_ring);
  if (ctx->d_col) (void)hipFree(ctx->d_col);
  if (ctx->h_red) (void)hipHostFree(ctx->h_red);
  if (ctx->e0) (void)hipEventDestroy(ctx->e0);
  if (ctx->e1) (void)hipEventDestroy(ctx->e1);
  if (ctx->st) (void)hipStreamDestroy(ctx->st);
  delete ctx;
  return 0;
}

extern "C" int fv_dims(const fv_ctx *ctx, int *ni, int *nj, int *vnum) {
  if (!ctx) return -1;
  if (ni) *ni = ctx->NI;
  if (nj) *nj = ctx->NJ;
  if (vnum) *vnum = ctx->V;
  return 0;
}

extern "C" int fv_set_sound_speeds(fv_ctx *ctx, const double *bccsound, const double *fcsound) {
  if (!ctx || !bccsound || !fcsound) return -1;
  if (ctx->V != 3) return fail(ctx, "sound speeds are set only for isothermal physics");
  const size_t N = (size_t)ctx->NI * ctx->NJ;
  if (!ring_from(ctx, R_CS, bccsound)) return fail(ctx, "bccsound must depend on the x index only");
  const int fid[4] = {R_CSW, R_CSE, R_CSS, R_CSN};
  for (int f = 0; f < 4; f++)
    if (!ring_from(ctx, fid[f], fcsound + f * N)) return fail(ctx, "fcsound must depend on the x index only");
  ctx->cs_min = DBL_MAX;
  for (size_t k = 0; k < N; k++) ctx->cs_min = bccsound[k] < ctx->cs_min ? bccsound[k] : ctx->cs_min;
  ctx->have_cs = true;
  return upload_geom(ctx);
}

extern "C" int fv_set_background_velocity(fv_ctx *ctx, const double *w) {
  if (!ctx || !w) return -1;
  if (ctx->c.fargo != 2) return fail(ctx, "background velocity needs fargo = 2");
  std::memcpy(&ctx->ring_h[(size_t)R_W * ctx->NI], w, sizeof(double) * ctx->NI);
  ctx->have_w = true;
  return upload_geom(ctx);
}

extern "C" int fv_set_pvar(fv_ctx *ctx, const double *pvar) {
  if (!ctx || !pvar) return -1;
  const int V = ctx->V;
  const size_t N = (size_t)ctx->NI * ctx->NJ;
  std::vector<double> cv(V * N);
  const double g = ctx->c.gamma;
  for (size_t k = 0; k < N; k++) {  // Prim2Cons2d, all cells
    const double rho = pvar[k], u = pvar[N + k], v = pvar[2 * N + k];
    cv[k] = rho;
    cv[N + k] = rho * u;
    cv[2 * N + k] = rho * v;
    if (V == 4) cv[3 * N + k] = pvar[3 * N + k] / (g - 1.) + 0.5 * rho * (u * u + v * v);
  }
  ctx->cvar = ctx->cold = 0;
  for (int l = 0; l < V; l++)
    HIPCHK(hipMemcpy2D(ctx->buf[0] + l * ctx->FS, sizeof(double) * ctx->P, cv.data() + l * N,
                       sizeof(double) * ctx->NI, sizeof(double) * ctx->NI, ctx->NJ, hipMemcpyHostToDevice));
  ctx->fl_pvar = ctx->fl_cvar = 0;
  ctx->have_state = true;
  return 0;
}

extern "C" int fv_first_step(fv_ctx *ctx, int *brk) {
  if (!ctx) return -1;
  if (!ctx->have_state) return fail(ctx, "initial data not set");
  if (ctx->V == 3 && !ctx->have_cs) return fail(ctx, "isothermal physics needs sound speeds");
  if (ctx->c.fargo == 2 && !ctx->have_w) return fail(ctx, "fargo = 2 needs the background velocity");
  if (ctx->iter > -1) return fail(ctx, "FirstStep should only be called once");
  ctx->iter = 0;
  int rc = rhs_main(ctx, CHECK_ALL);
  if (rc) return rc;
  ctx->dt = calc_timestep(ctx);
  ctx->cold = ctx->cvar;
  ctx->fl_cold = ctx->fl_cvar;
  // initial WriteDataset moves the state to the real frame (io/fileio_base.f90:523-533)
  if (ctx->time == 0.0 && ctx->c.fargo == 2 && (rc = add_bg_main(ctx))) return rc;
  if (brk) *brk = ctx->brk;
  return 0;
}

extern "C" int fv_step(fv_ctx *ctx, int *done) {
  if (!ctx) return -1;
  int rc, b;
  if (done) *done = 0;
  if (ctx->iter == -1 && (rc = fv_first_step(ctx, &b))) return rc;
  if (std::fabs(ctx->c.stoptime - ctx->time) <= 1.0E-05 * ctx->c.stoptime) {
    if (done) *done = 1;
    return 0;
  }
  // io/fileio_base.f90:604-620
  if ((ctx->time + ctx->dt) / ctx->t_out > 1.0) {
    ctx->dt = ctx->t_out - ctx->time;
    ctx->dtcause = DTCAUSE_FILEIO;
  } else if ((ctx->time + 1.5 * ctx->dt) / ctx->t_out > 1.0) {
    ctx->dt = 0.5 * (ctx->t_out - ctx->time);
    ctx->dtcause = DTCAUSE_FILEIO;
  }
  if ((rc = integration_step(ctx))) return rc;
  if (std::fabs(ctx->t_out - ctx->time) <= 1.0E-5 * ctx->t_out || ctx->brk) {
    if (ctx->c.fargo == 2 && (rc = add_bg_main(ctx))) return rc;  // WriteDataset
    if (ctx->brk) {
      if (done) *done = 1;
      return 0;
    }
    ctx->t_out = ctx->t_out + std::fabs(ctx->c.stoptime - 0.0) / 1;
    ctx->dtmin = ctx->c.stoptime;
    ctx->n_adj = 0;
  }
  ctx->dt = calc_timestep(ctx);
  return 0;
}

extern "C" int fv_run(fv_ctx *ctx, int nsteps, int *taken) {
  if (!ctx) return -1;
  int n = 0, done = 0, rc = 0;
  HIPCHK(hipEventRecord(ctx->e0, ctx->st));
  for (; n < nsteps && !done; n++)
    if ((rc = fv_step(ctx, &done))) break;
  HIPCHK(hipEventRecord(ctx->e1, ctx->st));
  HIPCHK(hipEventSynchronize(ctx->e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
  ctx->last_ms = ms;
  if (taken) *taken = done ? n - 1 : n;
  return rc;
}

extern "C" int fv_get_state(fv_ctx *ctx, double *cvar, double *pvar) {
  if (!ctx) return -1;
  const int V = ctx->V;
  const size_t N = (size_t)ctx->NI * ctx->NJ;
  std::vector<double> cv(V * N);
  HIPCHK(hipStreamSynchronize(ctx->st));
  for (int l = 0; l < V; l++)
    HIPCHK(hipMemcpy2D(cv.data() + l * N, sizeof(double) * ctx->NI, ctx->buf[ctx->cvar] + l * ctx->FS,
                       sizeof(double) * ctx->P, sizeof(double) * ctx->NI, ctx->NJ, hipMemcpyDeviceToHost));
  if (cvar) std::memcpy(cvar, cv.data(), sizeof(double) * V * N);
  if (pvar) {
    const double g = ctx->c.gamma;
    for (size_t k = 0; k < N; k++) {  // Cons2Prim2d
      const double rho = cv[k], inv = 1. / rho, mu = cv[N + k], mv = cv[2 * N + k];
      pvar[k] = rho;
      pvar[N + k] = mu * inv;
      pvar[2 * N + k] = mv * inv;
      if (V == 4) pvar[3 * N + k] = (g - 1.) * (cv[3 * N + k] - 0.5 * inv * (mu * mu + mv * mv));
    }
  }
  return 0;
}

extern "C" int fv_get_scalars(const fv_ctx *ctx, fv_scalars *o) {
  if (!ctx || !o) return -1;
  o->iter = ctx->iter;
  o->n_adj = ctx->n_adj;
  o->brk = ctx->brk;
  o->dtcause = ctx->dtcause;
  o->time = ctx->time;
  o->dt = ctx->dt;
  o->dtold = ctx->dtold;
  o->maxerrold = ctx->maxerrold;
  return 0;
}

extern "C" double fv_last_run_ms(const fv_ctx *ctx) { return ctx ? ctx->last_ms : -1.0; }

extern "C" const char *fv_last_error(const fv_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }
