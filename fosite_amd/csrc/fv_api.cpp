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
#include <cstddef>
#include <cstdlib>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "fv_host.h"

using namespace fv;

namespace {

constexpr int NBUF = 4;
// kernel classes of the timing statistics (fv_kernel_stats)
constexpr int K_RHS = 0, K_RHS_END = 1, K_RHS_COMBINE = 2, K_FRAME = 3, K_CONTROL = 4, K_REMAP = 5,
              K_ACCRETE = 6, K_VISC = 7, K_CLASSES = 8;
// timedisc_base.f90:193-203
constexpr int CHECK_NOTHING = 0, CHECK_ALL = ~0, CHECK_CSOUND = 1, CHECK_PMIN = 2;
constexpr int DTCAUSE_ERRADJ = -1, DTCAUSE_SMALLERR = -2, DTCAUSE_FILEIO = -4;
constexpr int DTCAUSE_VISCOSITY = 3;  // its source's position in sources_generic.f90:74
enum { RK_FEHLBERG = 2, CASH_KARP = 3, DORMAND_PRINCE = 4, SSPRK = 5 };

struct Tableau {
  int m;
  double a[MAXK][MAXK], bh[MAXK], bl[MAXK];
  double c[MAXK];  // stage times: ComputeRHS(time + c(m)*dt) (timedisc_rkfehlberg.f90:277-302)
};

// Butcher tableaux: timedisc_ssprk.f90:93-126, timedisc_rkfehlberg.f90:305-335,
// timedisc_dormand_prince.f90:91-114, timedisc_cashkarp.f90:101-113 (same
// rational literals, FP64)
bool set_tableau(int method, int order, Tableau &t) {
  std::memset(&t, 0, sizeof(t));
  if (method == SSPRK && order == 3) {
    t.m = 3;
    t.a[1][0] = 1.0;
    t.a[2][0] = 0.25; t.a[2][1] = 0.25;
    double bh[] = {1. / 6., 1. / 6., 2. / 3.}, bl[] = {0.5, 0.5, 0.}, cc[] = {0., 1., 0.5};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl)); std::memcpy(t.c, cc, sizeof(cc));
  } else if (method == SSPRK && order == 5) {
    t.m = 5;
    double a[5][5] = {{0., 0., 0., 0., 0.}, {0.36717, 0., 0., 0., 0.}, {0.26802, 0.3172, 0., 0., 0.},
                      {0.11606, 0.13735, 0.18816, 0., 0.}, {0.11212, 0.13269, 0.18178, 0.4198, 0.}};
    for (int i = 0; i < 5; i++) for (int j = 0; j < 5; j++) t.a[i][j] = a[i][j];
    double bh[] = {0.17279, 0.094505, 0.12947, 0.29899, 0.30424};
    double bl[] = {0.12293, 0.31981, -0.15316, 0.31887, 0.39155};
    double cc[] = {0., 0.36717, 0.58522, 0.44156, 0.8464};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl)); std::memcpy(t.c, cc, sizeof(cc));
  } else if (method == RK_FEHLBERG && order == 3) {
    t.m = 3;
    t.a[1][0] = 0.5;
    t.a[2][0] = -1.0; t.a[2][1] = 2.0;
    double bh[] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}, bl[] = {0.0, 1.0, 0.0}, cc[] = {0.0, 0.5, 1.0};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl)); std::memcpy(t.c, cc, sizeof(cc));
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
    double cc[] = {0.0, 0.25, 3.0 / 8.0, 12.0 / 13.0, 1.0, 0.5};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl)); std::memcpy(t.c, cc, sizeof(cc));
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
    double cc[] = {0., .2, 3. / 10., 4. / 5., 8. / 9., 1., 1.};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl)); std::memcpy(t.c, cc, sizeof(cc));
  } else if (method == CASH_KARP && order == 5) {  // timedisc_cashkarp.f90:101-113
    t.m = 6;
    double a[6][6] = {{0, 0, 0, 0, 0, 0},
                      {0.2, 0, 0, 0, 0, 0},
                      {0.075, 0.225, 0, 0, 0, 0},
                      {0.3, -0.9, 1.2, 0, 0, 0},
                      {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0, 0},
                      {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0, 0}};
    for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) t.a[i][j] = a[i][j];
    double bh[] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
    double bl[] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 0.25};
    double cc[] = {0.0, 0.2, 0.3, 0.6, 1.0, 7.0 / 8.0};
    std::memcpy(t.bh, bh, sizeof(bh)); std::memcpy(t.bl, bl, sizeof(bl)); std::memcpy(t.c, cc, sizeof(cc));
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
  float *d_ringf = nullptr, *d_colf = nullptr;  // FP32 copies (dtype 1)
  double *buf[NBUF] = {};
  int cvar = 0, cold = 0;
  double *k[MAXK] = {};     // stage rhs k_2..k_s (k[0] unused)
  double *k0[2] = {};       // coeff(1) = Timedisc%rhs, double-buffered for rollback
  int k0cur = 0;
  // boundary fluxes Fluxes%bxflux/byflux (fluxes_base.f90:142-143): bf[bcur]
  // holds the accepted values (bxfold), the last stage of a step writes the
  // other copy; AcceptSolution is a role swap, RejectSolution keeps bcur
  double *bf[2] = {};
  int bcur = 0;
  double *bks = nullptr;  // boundary rows of the last stage's rhs
  // Timedisc%pvar as the last host-path ComputeRHS left it (pvalid; its FARGO
  // frame flag in pflag); device rounds do not write it
  double *pbuf = nullptr;  // element type by dtype, like the state buffers
  bool pvalid = false;
  int pflag = 0;
  // point-mass accretion (c.outbound): state buffers [PM_HDR + 2*NI] and their
  // FP32 rings; every UpdateGravity reads grav[gcur], writes the other and
  // makes it current
  double *grav[2] = {};
  float *gravf[2] = {};
  int gcur = 0;
  // FARGO type 1: azimuthal velocities the last ComputeRHS left for the next
  // CalcBackgroundVelocity ([NJ][P], FP64) in vbufs[vcur]; w changes every
  // step, so the per-ring arrays exist twice and differ only in the w row:
  // rings(wcur) is current.  A device step reads rings(wcur) and vbufs[vcur]
  // and writes the other ones, so a halted step leaves the host's intact.
  double *vbufs[2] = {};
  int vcur = 0, wcur = 0;
  double *d_ring1 = nullptr;
  float *d_ringf1 = nullptr;
  // viscosity (c.vismodel): the source of the evaluation in flight, and for
  // ALPHA the dynamic viscosity of the last update and e_phi/r; vis_time is
  // the time of that update (sources_viscosity.f90:371, -1 before the first)
  double *d_vsrc = nullptr, *d_dyn = nullptr, *d_ephir = nullptr;
  double vis_time = -1.0;
  // device-controlled steps: per-workgroup viscous 1/dt of the end-of-step k_visc
  double *d_vpart = nullptr;
  int nb_vis = 0;
  // ALPHA with Euler physics: the sound speed of Timedisc%pvar, in d_cs[k0cur]
  // (a device step's stages read one, its end-of-step evaluation writes the
  // other, like the coeff(1) pair)
  double *d_cs[2] = {};
  Ctl *d_ctl = nullptr, *h_ctl = nullptr;
  struct GraphEntry { int x0, x1, t, k0, b0, g0, w0, v0, nsteps; hipGraphExec_t exec; };
  std::vector<GraphEntry> graphs;
  bool fast = true;         // device-controlled batches allowed
  int batch_steps = 0, batch_halts = 0;
  fv_batch *batch = nullptr;  // batch this context belongs to
  double *part_r = nullptr, *part_c = nullptr;  // per-block partials: CFL/checks, embedded error
  double *d_red = nullptr, *h_red = nullptr;
  int nb_rhs = 0;
  hipStream_t st = nullptr;
  int dev = 0;  // the device current at fv_create, which holds the state
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Geom g{};
  // FARGO frame flags (fargo_transformation_applied) of each state vector
  int fl_pvar = 0, fl_cvar = 0, fl_ptmp = 0, fl_ctmp = 0, fl_cold = 0;
  double time = 0, dt = 0, dtold = 0, dtmin = 0, maxerrold = 0, t_out = 0, dtlimit = 0;
  int iter = -1, n_adj = 0, brk = 0, dtcause = 0;
  int count = 1;  // datafile outputs between the start and the stop time
  bool have_cs = false, have_w = false, have_state = false;
  double cs_min = 0;
  double last_ms = 0;
  double invdt = 0;
  double invdt_v = 0;  // viscous 1/dt of the last device-controlled step (Ctl::invdt_v)
  int flags = 0;
  // per-kernel timing (HIP events around each launch on the context stream)
  bool prof = false;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  struct Pending { int kind; size_t e; double bytes; };
  std::vector<Pending> pend;
  int st_launch[K_CLASSES] = {};
  double st_ms[K_CLASSES] = {}, st_bytes[K_CLASSES] = {};
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
// the per-ring arrays with w of parity q (FARGO type 1 keeps two)
double *ring_tab(const fv_ctx *ctx, int q) { return q ? ctx->d_ring1 : ctx->d_ring; }
float *ringf_tab(const fv_ctx *ctx, int q) { return q ? ctx->d_ringf1 : ctx->d_ringf; }
void set_wcur(fv_ctx *ctx, int q) {
  ctx->wcur = q;
  ctx->g.ring = ring_tab(ctx, q);
  ctx->g.ringf = ringf_tab(ctx, q);
}
}  // namespace

namespace {

int fail(fv_ctx *ctx, const std::string &m, int code = -1) {
  ctx->err = m;
  return code;
}

int pick_free(const fv_ctx *, int a, int b, int c) {
  for (int q = 0; q < NBUF; q++)
    if (q != a && q != b && q != c) return q;
  return -1;
}

// kernel timing: returns the event slot recorded before a launch, or -1
long prof_begin(fv_ctx *ctx) {
  if (!ctx->prof) return -1;
  while (ctx->evused + 2 > ctx->evpool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    ctx->evpool.push_back(e);
  }
  size_t k = ctx->evused;
  ctx->evused += 2;
  (void)hipEventRecord(ctx->evpool[k], ctx->st);
  return (long)k;
}
void prof_end(fv_ctx *ctx, long k, int kind, double bytes) {
  if (k < 0) return;
  (void)hipEventRecord(ctx->evpool[k + 1], ctx->st);
  ctx->pend.push_back({kind, (size_t)k, bytes});
}
// after a synchronisation: fold the recorded intervals into the statistics
void prof_collect(fv_ctx *ctx) {
  for (auto &p : ctx->pend) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->evpool[p.e], ctx->evpool[p.e + 1]) == hipSuccess) {
      ctx->st_launch[p.kind]++;
      ctx->st_ms[p.kind] += ms;
      ctx->st_bytes[p.kind] += p.bytes;
    }
  }
  ctx->pend.clear();
  ctx->evused = 0;
}
double cells(const fv_ctx *ctx) { return (double)ctx->c.inum * ctx->c.jnum; }
size_t elem_bytes(const fv_ctx *ctx) { return ctx->c.dtype == 1 ? sizeof(float) : sizeof(double); }
double word_bytes(const fv_ctx *ctx) { return (double)elem_bytes(ctx) * ctx->V * cells(ctx); }
// address of field l of a state buffer (element type by dtype)
char *field(const fv_ctx *ctx, double *buf, int l) {
  return reinterpret_cast<char *>(buf) + (size_t)l * ctx->FS * elem_bytes(ctx);
}

// reduce the per-block partials and fetch {max, flags} to the host
int reduce(fv_ctx *ctx, const double *part, int n, double *mx, int *flags) {
  launch_reduce(part, n, ctx->d_red, ctx->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ctx->h_red, ctx->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  prof_collect(ctx);
  *mx = ctx->h_red[0];
  *flags = (int)ctx->h_red[1];
  return 0;
}

// gravitational acceleration rings of an evaluation: point-mass state q
// (accretion), else the static rings
void set_acc(const fv_ctx *ctx, StageArgs &a, int q) {
  if (ctx->grav[0]) {
    a.acc = ctx->grav[q] + PM_HDR;
    a.accf = ctx->gravf[q];
  } else {
    a.acc = ctx->d_ring + (size_t)R_ACCX * ctx->NI;
    a.accf = ctx->d_ringf ? ctx->d_ringf + (size_t)R_ACCX * ctx->NI : nullptr;
  }
}

StageArgs base_args(fv_ctx *ctx) {
  StageArgs a;
  std::memset(&a, 0, sizeof(a));
  a.part = ctx->part_r;
  a.bflux = 1;
  set_acc(ctx, a, ctx->gcur);
  return a;
}

VisArgs vis_args(const fv_ctx *ctx, int update) {
  VisArgs v;
  std::memset(&v, 0, sizeof(v));
  v.model = ctx->c.vismodel;
  v.update = update;
  v.dynconst = ctx->c.dynconst;
  v.bulkconst = ctx->c.bulkconst;
  v.omega = ctx->c.omega;
  v.ephir = ctx->d_ephir;
  v.dyn = ctx->d_dyn;
  v.src = ctx->d_vsrc;
  v.cs_in = v.cs_out = ctx->d_cs[ctx->k0cur];
  return v;
}
// the viscous source of the evaluation a at time t, ahead of its k_rhs
// launch (UpdateViscosity recomputes the dynamic viscosity when t differs
// from the last update's time or is 0, sources_viscosity.f90:371)
int visc(fv_ctx *ctx, StageArgs &a, double t) {
  if (!ctx->c.vismodel) return 0;
  const int update = t != ctx->vis_time || t == 0.0;
  ctx->vis_time = t;
  long pk = prof_begin(ctx);
  launch_visc(ctx->g, a, vis_args(ctx, update), ctx->V, ctx->st);
  prof_end(ctx, pk, K_VISC, 8.0 * cells(ctx) * (ctx->V * (a.nk + 1.0) + ctx->V - 1.0));
  HIPCHK(hipGetLastError());
  a.vsrc = ctx->d_vsrc;
  return 0;
}

// UpdateGravity_single of one ComputeRHS (gravity_pointmass.f90:383-428):
// point-mass state gin -> 1 - gin, with the boundary fluxes bf[b] and the
// call's dt argument
AccreteArgs accrete_args(const fv_ctx *ctx, int gin, int b, double dt) {
  AccreteArgs p;
  std::memset(&p, 0, sizeof(p));
  const int ob = ctx->c.outbound, xd = ob <= 2, side = (ob - 1) & 1;
  p.gin = ctx->grav[gin];
  p.gout = ctx->grav[1 - gin];
  p.gfout = ctx->gravf[1 - gin];
  // density rows: x part [V][2 sides][NJ], then y part [V][2 sides][NI]
  p.bf = ctx->bf[b] + (xd ? (size_t)side * ctx->NJ : 2 * (size_t)ctx->V * ctx->NJ + (size_t)side * ctx->NI);
  p.lo = GH;
  p.hi = GH + (xd ? ctx->c.jnum : ctx->c.inum) - 1;
  p.NI = ctx->NI;
  p.acclimit = ctx->c.acclimit;
  p.dt = dt;
  return p;
}
// the UpdateGravity that ComputeRHS performs first; the reference runs it at
// every evaluation, here on the host-controlled path as well
int accrete(fv_ctx *ctx, int b, double dt) {
  if (!ctx->grav[0]) return 0;
  launch_accrete(accrete_args(ctx, ctx->gcur, b, dt), ctx->st);
  HIPCHK(hipGetLastError());
  ctx->gcur = 1 - ctx->gcur;
  return 0;
}

// FARGO frame flags of one ComputeRHS call (timedisc_base.f90:1098-1208):
// returns add/sub and updates the flags of (p, c) like the reference does
void rhs_flags(const fv_ctx *ctx, int &fp, int &fc, int &add, int &sub) {
  add = sub = 0;
  if (!ctx->c.fargo) return;
  if (fc && fp) {
    add = 1;
    fc = fp = 0;
  }
  fp = fc;  // CenterBoundary
  if (!fc && !fp) {
    sub = 1;
    fc = fp = 1;
  }
}

// ComputeRHS on the main state (pvar, cvar) -> coeff(1) and the new cvar
// (fresh buffer, ghost cells included); also the CFL inverse time step and
// the CheckData flags (timedisc_base.f90:1666-1742).  remap = 1 evaluates it
// on FargoAdvectionY(cvar) (the post-step remap, timedisc_base.f90:722-743).
int rhs_main(fv_ctx *ctx, int checkbm, int remap, double t) {
  int add, sub;
  rhs_flags(ctx, ctx->fl_pvar, ctx->fl_cvar, add, sub);
  const int pmin = (ctx->V == 4 && (checkbm & CHECK_PMIN)) ? 1 : 0;
  int in = ctx->cvar;
  if (remap) {  // FargoAdvectionY into a free buffer
    const int rb = pick_free(ctx, ctx->cvar, ctx->cold, -1);
    RemapArgs r{ctx->buf[ctx->cvar], ctx->buf[rb], ctx->dt, nullptr};
    long pk = prof_begin(ctx);
    launch_remap(ctx->g, r, ctx->V, ctx->st);
    prof_end(ctx, pk, K_REMAP, 2.0 * word_bytes(ctx));
    HIPCHK(hipGetLastError());
    in = rb;
  }
  const int outb = pick_free(ctx, ctx->cvar, ctx->cold, in);
  StageArgs a = base_args(ctx);
  a.Y = ctx->buf[in];
  a.nk = 0;
  a.kout = ctx->k0[ctx->k0cur];
  a.yout = ctx->buf[outb];
  a.add_bg = add;
  a.sub_bg = sub;
  a.pmin_clamp = pmin;
  a.want_cfl = 1;
  a.want_check = checkbm != CHECK_NOTHING;
  a.vout = ctx->vbufs[ctx->vcur];
  a.pout = ctx->pbuf;
  if (int rc = visc(ctx, a, t)) return rc;
  long pk = prof_begin(ctx);
  launch_rhs(ctx->g, a, ctx->V, ctx->st);
  prof_end(ctx, pk, K_RHS_END, 3.0 * word_bytes(ctx));
  HIPCHK(hipGetLastError());
  ctx->pvalid = true;
  ctx->pflag = ctx->fl_pvar;
  ctx->cvar = outb;
  int rc = reduce(ctx, ctx->part_r, ctx->nb_rhs, &ctx->invdt, &ctx->flags);
  if (rc) return rc;
  if (checkbm != CHECK_NOTHING) {
    if ((checkbm & CHECK_CSOUND) && ctx->V == 3 && ctx->cs_min <= 0.) ctx->brk = 1;
    if (ctx->flags) ctx->brk = 1;
  }
  return 0;
}

// arguments of stage m (1..s-1) on the state y with coeff(1) in k0[k0],
// the last stage (m = s-1) fused with the final combination into buffer out
StageArgs stage_args(fv_ctx *ctx, int m, int y, int k0, int out) {
  StageArgs a = base_args(ctx);
  a.Y = ctx->buf[y];
  a.nk = m;
  for (int q = 0; q < m; q++) {
    a.k[q] = q ? ctx->k[q] : ctx->k0[k0];
    a.a[q] = ctx->tab.a[m][q];
  }
  if (m == ctx->tab.m - 1) {
    a.ynew = ctx->buf[out];
    for (int q = 0; q < ctx->tab.m; q++) {
      a.bh[q] = ctx->tab.bh[q];
      a.bl[q] = ctx->tab.bl[q];
    }
    a.tol_rel = ctx->c.tol_rel;
    for (int l = 0; l < 4; l++) a.tol_abs[l] = ctx->c.tol_abs[l];
    a.part_c = ctx->part_c;
    a.bks = ctx->bks;
  } else {
    a.kout = ctx->k[m];
  }
  return a;
}
// boundary-flux accumulation of a step with coeff(1) in k0[k0]: bf[bin] ->
// bf[1 - bin] (dt: host-controlled steps; device rounds read Ctl::dt)
BfluxArgs bflux_args(const fv_ctx *ctx, int k0, int bin, double dt) {
  BfluxArgs b;
  std::memset(&b, 0, sizeof(b));
  for (int q = 0; q < ctx->tab.m - 1; q++) b.k[q] = q ? ctx->k[q] : ctx->k0[k0];
  for (int q = 0; q < ctx->tab.m; q++) b.bh[q] = ctx->tab.bh[q];
  b.ks = ctx->bks;
  b.in = ctx->bf[bin];
  b.out = ctx->bf[1 - bin];
  b.dt = dt;
  b.nk = ctx->tab.m - 1;
  b.V = ctx->V;
  b.NI = ctx->NI;
  b.NJ = ctx->NJ;
  b.P = ctx->P;
  b.inum = ctx->c.inum;
  b.jnum = ctx->c.jnum;
  b.dtype = ctx->c.dtype;
  b.FS = ctx->FS;
  return b;
}

void launch_stage(fv_ctx *ctx, const Geom &g, const StageArgs &a, int m) {
  const bool last = m == ctx->tab.m - 1;
  long pk = prof_begin(ctx);
  launch_rhs(g, a, ctx->V, ctx->st);
  // stage m reads the state and m slopes and writes one slope; the last one
  // writes the combined state instead (m + 2 words per cell either way)
  prof_end(ctx, pk, last ? K_RHS_COMBINE : K_RHS, (m + 2.0) * word_bytes(ctx));
}

// ComputeRHS on the stage state (ptmp, ctmp), ctmp = cvar - sum a*dt*k, at
// time t
int rhs_stage(fv_ctx *ctx, int m, double dt, int out, double t) {
  if (int rc = accrete(ctx, ctx->bcur, dt)) return rc;
  int add, sub;
  rhs_flags(ctx, ctx->fl_ptmp, ctx->fl_ctmp, add, sub);
  StageArgs a = stage_args(ctx, m, ctx->cvar, ctx->k0cur, out);
  a.dt = dt;
  a.add_bg = add;
  a.sub_bg = sub;
  if (int rc = visc(ctx, a, t)) return rc;
  launch_stage(ctx, ctx->g, a, m);
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

// timedisc_rkfehlberg.f90:170-268; the boundary fluxes of the step go to
// bf[1 - bcur] (the last stage accumulates them)
int solve_ode(fv_ctx *ctx, double time, double *dt, double *err) {
  const int outb = pick_free(ctx, ctx->cvar, ctx->cold, -1);
  for (int m = 1; m < ctx->tab.m; m++) {  // the last stage writes the combination
    int rc = rhs_stage(ctx, m, *dt, outb, time + ctx->tab.c[m] * *dt);
    if (rc) return rc;
  }
  launch_bflux(bflux_args(ctx, ctx->k0cur, ctx->bcur, *dt), ctx->st);
  HIPCHK(hipGetLastError());
  ctx->fl_ctmp = ctx->fl_cvar;  // ctmp = cvar
  ctx->cvar = outb;
  double e;
  int f;
  int rc = reduce(ctx, ctx->part_c, ctx->nb_rhs, &e, &f);
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

// CalcTimestep reads the wave speeds of Timedisc%pvar (timedisc_base.f90:
// 844-858).  After an output, WriteDataset has moved pvar to the real frame
// (v + w), so the next CFL step is that of the real-frame velocities, not the
// comoving ones the end-of-step evaluation reduced: recompute invdt from the
// stored pvar of that evaluation with w added.
bool pvar_moved_to_real_frame(const fv_ctx *ctx) {
  return ctx->c.fargo && ctx->pvalid && ctx->pflag && !ctx->fl_pvar;
}
int cfl_real_frame(fv_ctx *ctx) {
  const long cells = (long)ctx->NI * ctx->NJ;
  const int nb = (int)std::min<long>(ctx->nb_rhs, (cells + 255) / 256);
  launch_cfl_pvar(ctx->g, ctx->pbuf, 1, ctx->part_r, nb, ctx->V, ctx->st);
  HIPCHK(hipGetLastError());
  int f = 0;
  return reduce(ctx, ctx->part_r, nb, &ctx->invdt, &f);
}

// CalcTimestep (timedisc_base.f90:818-909) into ctx->dt: the CFL limit, then
// the source list's limits (sources_generic.f90:264-293), of which only the
// viscosity's is finite: its advection-diffusion limit over every cell
// (sources_viscosity.f90:507-577) on Timedisc%pvar, after UpdateViscosity at
// the current time
int timestep(fv_ctx *ctx) {
  ctx->dt = calc_timestep(ctx);
  if (!ctx->c.vismodel) return 0;
  double iv = ctx->invdt_v;  // after device-controlled steps: the last one's, from the step control
  if (ctx->pvalid) {
    const int update = ctx->c.vismodel >= VIS_ALPHA && (ctx->time != ctx->vis_time || ctx->time == 0.0);
    ctx->vis_time = ctx->time;
    const long cells = (long)ctx->NI * ctx->NJ;
    const int nb = (int)std::min<long>(ctx->nb_rhs, (cells + 255) / 256);
    launch_visc_cfl(ctx->g, vis_args(ctx, update), ctx->pbuf, pvar_moved_to_real_frame(ctx) ? 1 : 0, ctx->part_r,
                    nb, ctx->V, ctx->st);
    HIPCHK(hipGetLastError());
    int f = 0;
    if (int rc = reduce(ctx, ctx->part_r, nb, &iv, &f)) return rc;
  }
  const double dtv = iv > DBL_MIN ? ctx->c.cvis / iv : DBL_MAX;
  if (dtv < ctx->dt) ctx->dtcause = DTCAUSE_VISCOSITY;
  ctx->dt = dtv < ctx->dt ? dtv : ctx->dt;
  return 0;
}

// out-of-place frame change of cvar (Physics%Add/SubtractBackgroundVelocityY)
int frame_cvar(fv_ctx *ctx, int sign) {
  const int outb = pick_free(ctx, ctx->cvar, ctx->cold, -1);
  long pk = prof_begin(ctx);
  launch_frame(ctx->g, ctx->buf[ctx->cvar], ctx->buf[outb], sign, ctx->V, ctx->st);
  prof_end(ctx, pk, K_FRAME, 2.0 * 8.0 * ctx->V * (double)ctx->NI * ctx->NJ);
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
// FARGO type 1 CalcBackgroundVelocity (timedisc_base.f90:2243-2326): back to
// the real frame with the old w, then w from the azimuthal velocities the
// last ComputeRHS produced (kept in vbuf: the pvar that AddBackgroundVelocityY
// restores is (v - w) + w of that evaluation)
int calc_background_velocity(fv_ctx *ctx) {
  int rc = add_bg_main(ctx);
  if (rc) return rc;
  launch_wsum(ctx->g, ctx->vbufs[ctx->vcur], ring_tab(ctx, 1 - ctx->wcur), ringf_tab(ctx, 1 - ctx->wcur), ctx->st);
  HIPCHK(hipGetLastError());
  set_wcur(ctx, 1 - ctx->wcur);
  return 0;
}

// timedisc_base.f90:649-750
int integration_step(fv_ctx *ctx) {
  int rc;
  double err = 0.0, dtold = ctx->dt, dt, time;
  if (ctx->c.fargo == 1 && (rc = calc_background_velocity(ctx))) return rc;
  if (ctx->c.fargo && (rc = sub_bg_main(ctx))) return rc;
  time = ctx->time;
  dt = ctx->dt;
  if (dt < ctx->dtmin && ctx->dtcause != DTCAUSE_FILEIO) ctx->dtmin = dt;
  while (!ctx->brk && time + dt <= ctx->time + ctx->dt) {
    dtold = dt;
    if ((rc = solve_ode(ctx, time, &dt, &err))) return rc;
    if (err != err || err > DBL_MAX) {
      ctx->brk = 1;
    } else if (err >= 1.0) {  // RejectSolution (timedisc_base.f90:961-1009)
      ctx->cvar = ctx->cold;
      ctx->fl_cvar = ctx->fl_cold;
      // its ComputeRHS runs before bxflux = bxfold: the point mass sees the
      // rejected step's boundary fluxes
      if ((rc = accrete(ctx, 1 - ctx->bcur, dt))) return rc;
      if ((rc = rhs_main(ctx, CHECK_NOTHING, 0, time))) return rc;
      ctx->n_adj++;
      ctx->dtcause = DTCAUSE_ERRADJ;
      if (dt < ctx->dtmin) ctx->dtmin = dt;
    } else {  // AcceptSolution (timedisc_base.f90:911-959)
      time = time + dtold;
      ctx->bcur = 1 - ctx->bcur;  // bxfold = bxflux
      if (time < ctx->time + ctx->dt) {
        if ((rc = accrete(ctx, ctx->bcur, dtold))) return rc;
        if ((rc = rhs_main(ctx, CHECK_ALL, 0, time))) return rc;
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
    const int remap = ctx->c.fargo != 0;
    // FARGO: CenterBoundary in the comoving frame, FargoAdvectionY and
    // Convert2Primitive (timedisc_base.f90:722-741) are folded into the
    // end-of-step RHS evaluation, which reads the remapped state on the fly
    if (remap) ctx->fl_pvar = ctx->fl_cvar;
    // ComputeRHS(..., this%time, 0. with FARGO, else dtold, ...) (timedisc_base.f90:742-746)
    if ((rc = accrete(ctx, ctx->bcur, remap ? 0.0 : dtold))) return rc;
    if ((rc = rhs_main(ctx, CHECK_ALL, remap, ctx->time))) return rc;
    ctx->cold = ctx->cvar;
    ctx->fl_cold = ctx->fl_cvar;
  }
  return 0;
}

int upload_geom(fv_ctx *ctx) {
  const int NI = ctx->NI, NJ = ctx->NJ;
  // copies on the context's own (non-blocking) stream: no null-stream
  // synchronisation with work other contexts have in flight
  for (int q = 0; q < 2; q++)
    if (ring_tab(ctx, q))
      HIPCHK(hipMemcpyAsync(ring_tab(ctx, q), ctx->ring_h.data(), sizeof(double) * R_COUNT * NI,
                            hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(ctx->d_col, ctx->col_h.data(), sizeof(double) * C_COUNT * NJ, hipMemcpyHostToDevice,
                        ctx->st));
  std::vector<float> rf, cf;
  if (ctx->d_ringf) {
    rf.assign(ctx->ring_h.begin(), ctx->ring_h.end());
    cf.assign(ctx->col_h.begin(), ctx->col_h.end());
    for (int q = 0; q < 2; q++)
      if (ringf_tab(ctx, q))
        HIPCHK(hipMemcpyAsync(ringf_tab(ctx, q), rf.data(), sizeof(float) * R_COUNT * NI, hipMemcpyHostToDevice,
                              ctx->st));
    HIPCHK(hipMemcpyAsync(ctx->d_colf, cf.data(), sizeof(float) * C_COUNT * NJ, hipMemcpyHostToDevice, ctx->st));
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
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

// ------------------------------------------------- device-controlled batches
// In the steady state of the integration (frame flags settled, cold = cvar,
// no pending break), a step that is accepted at the first attempt needs no
// host decision: the device runs stages, combination, accept test, end-of-step
// RHS and CFL update, and a 2-step HIP graph (buffers alternate between two
// state arrays) is launched repeatedly.  Steps that leave this path halt the
// batch and are redone by the host path above.
// (batch: members of a batch, whose shared launches have no FARGO type-1
// step, take host-controlled steps for it)
bool fast_eligible(const fv_ctx *ctx, bool batch = false) {
  if (!ctx->fast || ctx->iter < 1 || ctx->brk || ctx->cvar != ctx->cold) return false;
  if (batch && ctx->c.fargo == 1) return false;
  if (!(ctx->c.tol_rel < 1.0)) return false;  // AdjustTimestep active
  // the stop test of the next Step() is the host's (fosite.f90:370-373)
  if (std::fabs(ctx->c.stoptime - ctx->time) <= 1.0E-05 * ctx->c.stoptime) return false;
  const int f = ctx->c.fargo ? 1 : 0;
  return ctx->fl_cvar == f && ctx->fl_pvar == f && ctx->fl_ctmp == f && ctx->fl_ptmp == f &&
         ctx->fl_cold == f;
}

// The launches of one device-controlled step: input state x_in/k0_in ->
// x_out/k0_out (tb: combined state, rb: remapped state), step control in ctl
struct FastStep {
  Geom g;              // geometry of the step's evaluations (FARGO type 1: with the new w)
  // FARGO type 1, CalcBackgroundVelocity at the start of the step: w from vin
  // into the other ring arrays, then x_in to the real frame with the old w
  // and back to the comoving frame with the new one, in one pass (into x_out,
  // the stages' input)
  bool cbv;
  Geom gold;
  const double *vin;
  double *wring;
  float *wringf;
  const void *fsrc, *wnew;
  void *fdst;
  StageArgs st[MAXK];  // stages 2..s (index m-1); the last one fused with the combination
  bool remap;
  RemapArgs rm;
  bool accrete;        // point-mass accretion: boundary fluxes, then UpdateGravity
  BfluxArgs bfl;
  AccreteArgs acc;
  StageArgs end;       // end-of-step ComputeRHS: CFL, data checks, new state with ghosts
};
void fast_step_args(fv_ctx *ctx, Ctl *ctl, int xin, int xout, int tb, int rb, int k0in, int k0out, int bin,
                    int gin, int win, int vin, FastStep &f) {
  const int steady = ctx->c.fargo ? 1 : 0;
  f.g = ctx->g;
  f.cbv = ctx->c.fargo == 1;
  int ystage = xin;
  if (f.cbv) {
    f.gold = ctx->g;
    f.gold.ring = ring_tab(ctx, win);
    f.gold.ringf = ringf_tab(ctx, win);
    f.g.ring = ring_tab(ctx, 1 - win);
    f.g.ringf = ringf_tab(ctx, 1 - win);
    f.vin = ctx->vbufs[vin];
    f.wring = ring_tab(ctx, 1 - win);
    f.wringf = ringf_tab(ctx, 1 - win);
    f.fsrc = ctx->buf[xin];
    f.fdst = ctx->buf[xout];
    f.wnew = ctx->c.dtype == 1 ? (const void *)(f.wringf + (size_t)R_W * ctx->NI)
                               : (const void *)(f.wring + (size_t)R_W * ctx->NI);
    ystage = xout;  // free until the end-of-step evaluation writes it
  }
  for (int m = 1; m < ctx->tab.m; m++) {
    StageArgs a = stage_args(ctx, m, ystage, k0in, tb);
    a.ctl = ctl;
    a.add_bg = a.sub_bg = steady;
    set_acc(ctx, a, gin);
    f.st[m - 1] = a;
  }
  // accretion: the stage evaluations see unchanged boundary fluxes (their
  // UpdateGravity only sets accrate = 0); the end-of-step evaluation sees the
  // step's, so they are accumulated before it
  f.accrete = ctx->grav[0] != nullptr;
  if (f.accrete) {
    f.bfl = bflux_args(ctx, k0in, bin, 0.0);
    f.bfl.ctl = ctl;
    f.acc = accrete_args(ctx, gin, 1 - bin, 0.0);
    f.acc.dt_ctl = ctx->c.fargo ? 0 : 1;
    f.acc.zero_accrate = 1;
    f.acc.ctl = ctl;
  }
  int y = tb;
  f.remap = ctx->c.fargo != 0;
  if (f.remap) {
    f.rm = RemapArgs{ctx->buf[tb], ctx->buf[rb], 0.0, ctl};
    y = rb;
  }
  StageArgs a = base_args(ctx);
  a.Y = ctx->buf[y];
  a.nk = 0;
  a.ctl = ctl;
  a.kout = ctx->k0[k0out];
  a.yout = ctx->buf[xout];
  a.add_bg = a.sub_bg = steady;
  a.pmin_clamp = ctx->V == 4 ? 1 : 0;
  a.want_cfl = a.want_check = 1;
  if (f.accrete)
    set_acc(ctx, a, 1 - gin);
  else
    a.bfa = bflux_args(ctx, k0in, bin, 0.0);  // the step's boundary fluxes
  if (f.cbv) a.vout = ctx->vbufs[1 - vin];  // for the next step's CalcBackgroundVelocity
  f.end = a;
}

// the viscous source of a device-controlled evaluation at ctl->time +
// c_cur*ctl->dt, the previous evaluation being at ctl->time + c_prev*ctl->dt
// (part: the end-of-step one's viscous 1/dt partials)
void visc_device(fv_ctx *ctx, const Geom &g, StageArgs &a, double c_cur, double c_prev, double *part, int csin,
                 int csout) {
  VisArgs v = vis_args(ctx, 0);
  v.c_cur = c_cur;
  v.c_prev = c_prev;
  v.part = part;
  v.cs_in = ctx->d_cs[csin];
  v.cs_out = ctx->d_cs[csout];
  long pk = prof_begin(ctx);
  launch_visc(g, a, v, ctx->V, ctx->st);
  prof_end(ctx, pk, K_VISC, 8.0 * cells(ctx) * (ctx->V * (a.nk + 1.0) + ctx->V - 1.0));
  a.vsrc = ctx->d_vsrc;
}

void enqueue_fast_step(fv_ctx *ctx, int xin, int xout, int tb, int rb, int k0in, int k0out, int bin, int gin,
                       int win, int vin) {
  FastStep f;
  fast_step_args(ctx, ctx->d_ctl, xin, xout, tb, rb, k0in, k0out, bin, gin, win, vin, f);
  const bool vis = ctx->c.vismodel != 0;
  if (f.cbv) {
    long pk = prof_begin(ctx);
    launch_wsum(f.gold, f.vin, f.wring, f.wringf, ctx->st, ctx->d_ctl);
    launch_frame(f.gold, f.fsrc, f.fdst, +1, ctx->V, ctx->st, ctx->d_ctl, f.wnew);
    prof_end(ctx, pk, K_FRAME, 2.0 * word_bytes(ctx));
  }
  for (int m = 1; m < ctx->tab.m; m++) {
    // stage m runs at time + c(m)*dt; stage 1 follows the end-of-step
    // evaluation of the previous step, at time
    if (vis) visc_device(ctx, f.g, f.st[m - 1], ctx->tab.c[m], m > 1 ? ctx->tab.c[m - 1] : 0.0, nullptr, k0in, k0in);
    launch_stage(ctx, f.g, f.st[m - 1], m);
  }
  if (f.remap) {
    long pk = prof_begin(ctx);
    launch_remap(f.g, f.rm, ctx->V, ctx->st);
    prof_end(ctx, pk, K_REMAP, 2.0 * word_bytes(ctx));
  }
  if (f.accrete) {  // the step's boundary fluxes, then UpdateGravity
    long pk = prof_begin(ctx);
    launch_bflux(f.bfl, ctx->st);
    launch_accrete(f.acc, ctx->st);
    prof_end(ctx, pk, K_ACCRETE, 0.0);
  }
  // the end of the step at time + dt (as time + 1*dt)
  if (vis) visc_device(ctx, f.g, f.end, 1.0, ctx->tab.c[ctx->tab.m - 1], ctx->d_vpart, k0out, k0out);
  long pk = prof_begin(ctx);
  launch_rhs(f.g, f.end, ctx->V, ctx->st);
  prof_end(ctx, pk, K_RHS_END, 3.0 * word_bytes(ctx));
  pk = prof_begin(ctx);
  launch_control(ctx->d_ctl, ctx->part_c, ctx->part_r, ctx->nb_rhs, ctx->st, vis ? ctx->d_vpart : nullptr,
                 ctx->nb_vis);
  prof_end(ctx, pk, K_CONTROL, 32.0 * ctx->nb_rhs);
}

// step-control header of a device-controlled batch of at most nmax steps.
// ctx->dt is the CFL step of the last CalcTimestep; the first device step
// takes it after the output-time clipping the host applies at the start of
// every Step (io/fileio_base.f90:604-620), as k_control does for the others.
void prime_ctl(const fv_ctx *ctx, Ctl *h, int nmax) {
  std::memset(h, 0, offsetof(Ctl, err_hist));
  double dt = ctx->dt;
  if ((ctx->time + dt) / ctx->t_out > 1.0)
    dt = ctx->t_out - ctx->time;
  else if ((ctx->time + 1.5 * dt) / ctx->t_out > 1.0)
    dt = 0.5 * (ctx->t_out - ctx->time);
  h->time = ctx->time;
  h->dt = dt;
  h->t_out = ctx->t_out;
  h->stoptime = ctx->c.stoptime;
  h->cfl = ctx->c.cfl;
  h->dtlimit = ctx->dtlimit;
  h->maxerrold = ctx->maxerrold;
  h->beta = ctx->c.beta;
  h->order = ctx->c.order;
  h->nsteps = nmax;
  h->vis = ctx->c.vismodel ? 1 : 0;
  h->cvis = ctx->c.cvis;
}

// replays the host scalars of the h->step steps the device completed, exactly
// as the host path would have computed them (fosite.f90:367-464,
// timedisc_base.f90:649-750); the state alternated x0 -> x1 -> x0 ...
int replay_device_steps(fv_ctx *ctx, const Ctl *h, int x0, int x1, int *halted) {
  const int k = h->step;
  if (k > 0) ctx->pvalid = false;  // device steps leave no pvar
  for (int s = 0; s < k; s++) {
    const double dts = h->dt_hist[s];
    if (dts < ctx->dtmin && ctx->dtcause != DTCAUSE_FILEIO) ctx->dtmin = dts;
    const double dtnew = adjust_timestep(ctx, h->err_hist[s], dts);
    const double t1 = ctx->time + dts;
    ctx->iter++;
    if (dtnew < ctx->dtlimit) return fail(ctx, "device step control missed a dtlimit break", -4);
    ctx->dt = t1 - ctx->time;
    ctx->time = t1;
    ctx->dtold = dtnew;
    ctx->dtcause = 0;
    ctx->cvar = ctx->cold = (s % 2 == 0) ? x1 : x0;
    ctx->k0cur = 1 - ctx->k0cur;
    ctx->bcur = 1 - ctx->bcur;
    if (ctx->grav[0]) ctx->gcur = 1 - ctx->gcur;
    if (ctx->c.fargo == 1) {
      set_wcur(ctx, 1 - ctx->wcur);
      ctx->vcur = 1 - ctx->vcur;
    }
  }
  if (k > 0) {  // CalcTimestep of the last accepted step (before output clipping)
    ctx->invdt = h->invdt;
    ctx->dt = calc_timestep(ctx);
    if (ctx->c.vismodel) {  // and its viscous limit, as timestep() forms it
      ctx->invdt_v = h->invdt_v;
      const double dtv = h->invdt_v > DBL_MIN ? ctx->c.cvis / h->invdt_v : DBL_MAX;
      if (dtv < ctx->dt) ctx->dtcause = DTCAUSE_VISCOSITY;
      ctx->dt = dtv < ctx->dt ? dtv : ctx->dt;
    }
  }
  // the device steps (a halted one included) updated the stored viscosity on
  // their own schedule: the next host evaluation recomputes it
  ctx->vis_time = std::numeric_limits<double>::quiet_NaN();
  *halted = h->reason != H_BATCH;
  ctx->batch_steps += k;
  ctx->batch_halts += *halted;
  return 0;
}

int get_graph(fv_ctx *ctx, int x0, int x1, int tb, int rb, int k0, int b0, int g0, int w0, int v0,
              hipGraphExec_t *out) {
  for (auto &e : ctx->graphs)
    if (e.x0 == x0 && e.x1 == x1 && e.t == tb && e.k0 == k0 && e.b0 == b0 && e.g0 == g0 && e.w0 == w0 &&
        e.v0 == v0) {
      *out = e.exec;
      return 0;
    }
  hipGraph_t graph;
  HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
  const int g1 = ctx->grav[0] ? 1 - g0 : g0, t1 = ctx->c.fargo == 1;
  enqueue_fast_step(ctx, x0, x1, tb, rb, k0, 1 - k0, b0, g0, w0, v0);
  enqueue_fast_step(ctx, x1, x0, tb, rb, 1 - k0, k0, 1 - b0, g1, t1 ? 1 - w0 : w0, t1 ? 1 - v0 : v0);
  HIPCHK(hipStreamEndCapture(ctx->st, &graph));
  hipGraphExec_t exec;
  HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  (void)hipGraphDestroy(graph);
  ctx->graphs.push_back({x0, x1, tb, k0, b0, g0, w0, v0, 2, exec});
  *out = exec;
  return 0;
}

// FV_EAGER=1: device rounds issue their launches one by one instead of
// replaying HIP graphs (same kernels, same arguments, same order).  For runs
// under a profiler: two profiled runs whose process had destroyed one
// context's or batch's graphs before launching another's crashed in host
// code inside a HIP call (no GPU fault; the same runs complete unprofiled).
bool eager_launches() {
  static const bool e = [] {
    const char *v = std::getenv("FV_EAGER");
    return v && std::atoi(v) != 0;
  }();
  return e;
}

// runs up to nmax steps on the device; *done_steps = steps completed there,
// *halted = 1 if the batch stopped on a step the host must redo
int run_batch(fv_ctx *ctx, int nmax, int *done_steps, int *halted) {
  *done_steps = 0;
  *halted = 0;
  const int x0 = ctx->cvar;
  const int x1 = pick_free(ctx, x0, -1, -1);
  const int tb = pick_free(ctx, x0, x1, -1);
  const int rb = pick_free(ctx, x0, x1, tb);
  const int k0 = ctx->k0cur, b0 = ctx->bcur, g0 = ctx->gcur, g1 = ctx->grav[0] ? 1 - g0 : g0;
  const int w0 = ctx->wcur, v0 = ctx->vcur, t1 = ctx->c.fargo == 1;
  Ctl *h = ctx->h_ctl;
  prime_ctl(ctx, h, nmax);
  hipGraphExec_t exec = nullptr;
  if (!ctx->prof && !eager_launches()) {
    int rc = get_graph(ctx, x0, x1, tb, rb, k0, b0, g0, w0, v0, &exec);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpyAsync(ctx->d_ctl, h, offsetof(Ctl, err_hist), hipMemcpyHostToDevice, ctx->st));
  if (exec) {
    for (int s = 0; s < nmax; s += 2) HIPCHK(hipGraphLaunch(exec, ctx->st));
  } else {  // timed or eager: the same launches, issued one by one
    for (int s = 0; s < nmax; s++)
      enqueue_fast_step(ctx, s % 2 ? x1 : x0, s % 2 ? x0 : x1, tb, rb, s % 2 ? 1 - k0 : k0, s % 2 ? k0 : 1 - k0,
                        s % 2 ? 1 - b0 : b0, s % 2 ? g1 : g0, (s % 2 && t1) ? 1 - w0 : w0,
                        (s % 2 && t1) ? 1 - v0 : v0);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(h, ctx->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  prof_collect(ctx);
  int rc = replay_device_steps(ctx, h, x0, x1, halted);
  if (rc) return rc;
  *done_steps = h->step;
  return 0;
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
  if (c.recvars != 1 && c.recvars != 2) return bail("reconstruction variables must be 1 PRIMITIVE or 2 CONSERVATIVE");
  if (c.dtype != 0 && c.dtype != 1) return bail("dtype must be 0 (FP64) or 1 (FP32 precision form)");
  if (c.fargo < 0 || c.fargo > 2) return bail("fargo must be 0, 1 or 2");
  if (c.fargo && c.physics == EULER && c.dtype == 1)
    return bail("FARGO with Euler physics runs in FP64 only (dtype 0)");
  switch (c.limiter) {
    case MINMOD: case MONOCENT: case SWEBY: case SUPERBEE: case OSPRE: case VANLEER: case NOLIMIT: break;
    default: return bail("unknown limiter");
  }
  for (int d = 0; d < 4; d++) {
    const int b = c.bc[d];
    if (b != NO_GRADIENTS && b != PERIODIC && b != ABSORBING) return bail("boundary must be NO_GRADIENTS, PERIODIC or ABSORBING");
    if (b == ABSORBING && d >= 2) return bail("ABSORBING boundaries are supported in x");
  }
  // boundary_generic.f90: a periodic boundary needs a periodic opposite
  if ((c.bc[0] == PERIODIC) != (c.bc[1] == PERIODIC) || (c.bc[2] == PERIODIC) != (c.bc[3] == PERIODIC))
    return bail("opposite boundary should be periodic");
  if (c.gravity && c.geometry != LOGCYLINDRICAL) return bail("point-mass gravity requires a logcylindrical mesh");
  if (c.outbound < 0 || c.outbound > 4) return bail("outbound must be 0 (no accretion) or a boundary 1..4");
  if (c.outbound && !c.gravity) return bail("accretion (outbound) needs the point mass (gravity = 1)");
  if (c.rotframe < 0 || c.rotframe > 2) return bail("rotframe must be 0 (off), 1 or 2 (disable_centaccel)");
  if (c.vismodel < 0 || c.vismodel > 3) return bail("vismodel must be 0 (off), 1 MOLECULAR, 2 ALPHA or 3 BETA");
  if (c.vismodel && c.dtype != 0) return bail("viscosity is supported with the FP64 state (dtype 0)");
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
                                     {R_RDXE, M_RDXE}, {R_ACCX, M_ACCX}, {R_ACCY, M_ACCY}, {R_WFAC, M_WFAC},
                                     {R_DAXE, M_DAXE}, {R_DAYN, M_DAYN}};
  for (auto &p : rm)
    if (!ring_from(ctx, p.r, ctx->mesh.a[p.m].data())) return bail("mesh metric depends on j");
  for (int j = 0; j < ctx->NJ; j++) {
    ctx->col_h[(size_t)C_RDYS * ctx->NJ + j] = ctx->mesh.a[M_RDYS][(size_t)j * ctx->NI];
    ctx->col_h[(size_t)C_RDYN * ctx->NJ + j] = ctx->mesh.a[M_RDYN][(size_t)j * ctx->NI];
    for (int i = 1; i < ctx->NI; i++)
      if (ctx->mesh.a[M_RDYS][(size_t)j * ctx->NI + i] != ctx->col_h[(size_t)C_RDYS * ctx->NJ + j])
        return bail("reconstruction offsets in y depend on i");
  }
  // rotating frame, centre at the origin (sources_rotframe.f90:110-175): the
  // centrifugal acceleration caccel = (Omega*e_z x r) x Omega*e_z of the
  // bary-centre position r = (posvec_x(i), posvec_y(j), 0); with the zero
  // components of e_z the cross products reduce exactly to (Omega*x)*Omega
  // and (Omega*y)*Omega.  Only interior cells read it (Omega*e_z is zero on
  // ghost cells there).
  if (c.rotframe) {
    const double om = c.omega;
    for (int i = 0; i < ctx->NI; i++)
      ctx->ring_h[(size_t)R_CACX * ctx->NI + i] = (om * ctx->mesh.a[M_POSX][(size_t)GH * ctx->NI + i]) * om;
    for (int j = 0; j < ctx->NJ; j++)
      ctx->col_h[(size_t)C_CACY * ctx->NJ + j] = (om * ctx->mesh.a[M_POSY][(size_t)j * ctx->NI + GH]) * om;
  }
  if (const char *e = std::getenv("FV_HOST_CONTROL")) ctx->fast = std::atoi(e) == 0;
  // device resources
  const size_t sbytes = elem_bytes(ctx) * ctx->V * ctx->FS;
  auto dalloc = [&](double **p, size_t b) { return hipMalloc((void **)p, b) == hipSuccess; };
  bool ok = hipGetDevice(&ctx->dev) == hipSuccess &&
            hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&ctx->e0) == hipSuccess &&
            hipEventCreate(&ctx->e1) == hipSuccess;
  auto zalloc = [&](double **p) {
    return dalloc(p, sbytes) && hipMemsetAsync(*p, 0, sbytes, ctx->st) == hipSuccess;
  };
  for (int q = 0; q < NBUF && ok; q++) ok = zalloc(&ctx->buf[q]);
  ok = ok && zalloc(&ctx->pbuf);
  for (int q = 1; q < ctx->tab.m && ok; q++) ok = zalloc(&ctx->k[q]);
  for (int q = 0; q < 2 && ok; q++) ok = zalloc(&ctx->k0[q]);
  const size_t bfbytes = sizeof(double) * 2 * ctx->V * (ctx->NJ + ctx->NI);
  for (int q = 0; q < 2 && c.fargo == 1; q++) ok = ok && dalloc(&ctx->vbufs[q], sizeof(double) * ctx->FS);
  if (c.fargo == 1) {
    ok = ok && dalloc(&ctx->d_ring1, sizeof(double) * R_COUNT * ctx->NI);
    if (c.dtype == 1) ok = ok && hipMalloc((void **)&ctx->d_ringf1, sizeof(float) * R_COUNT * ctx->NI) == hipSuccess;
  }
  ok = ok && dalloc(&ctx->bks, sizeof(double) * 2 * ctx->V * (ctx->NJ + ctx->NI));
  for (int q = 0; q < 2 && ok; q++)  // InitFluxes: zero (fluxes_base.f90:209-214)
    ok = dalloc(&ctx->bf[q], bfbytes) && hipMemsetAsync(ctx->bf[q], 0, bfbytes, ctx->st) == hipSuccess;
  ok = ok && dalloc(&ctx->d_ring, sizeof(double) * R_COUNT * ctx->NI) && dalloc(&ctx->d_col, sizeof(double) * C_COUNT * ctx->NJ);
  if (c.dtype == 1)
    ok = ok && hipMalloc((void **)&ctx->d_ringf, sizeof(float) * R_COUNT * ctx->NI) == hipSuccess &&
         hipMalloc((void **)&ctx->d_colf, sizeof(float) * C_COUNT * ctx->NJ) == hipSuccess;
  Geom &g = ctx->g;
  g.NI = ctx->NI; g.NJ = ctx->NJ; g.P = ctx->P; g.FS = ctx->FS;
  g.inum = c.inum; g.jnum = c.jnum;
  g.invdx = ctx->mesh.invdx; g.invdy = ctx->mesh.invdy;
  // CheckData limits and the Euler pressure floor (timedisc_base.f90:353-359)
  g.gamma = c.gamma;
  g.pmin = c.pmin > 0 ? c.pmin : DBL_MIN;
  g.rhomin = c.rhomin > 0 ? c.rhomin : DBL_MIN;
  g.limiter = c.limiter; g.theta = c.theta;
  g.recon_prim = c.recvars == 1;
  for (int d = 0; d < 4; d++) g.bc[d] = c.bc[d];
  g.bdydz = ctx->mesh.dy * ctx->mesh.dz;
  g.bdzdx = ctx->mesh.dz * ctx->mesh.dx;
  g.fargo = c.fargo != 0; g.gravity = c.gravity != 0; g.curvilinear = c.geometry != CARTESIAN;
  g.rotframe = c.rotframe;
  g.twoomega = 2 * c.omega;
  g.dtype = c.dtype;
  g.stages = ctx->tab.m;
  g.tile_x = 64;
  if (const char *e = std::getenv("FV_RHS_TILE")) g.tile_x = std::atoi(e) == 32 ? 32 : 64;
  ctx->nb_rhs = rhs_blocks(g);
  Geom g32 = g;  // partials sized for the finest tiling, which batches may choose
  g32.tile_x = 32;
  const int nb_part = std::max(ctx->nb_rhs, rhs_blocks(g32));
  ok = ok && dalloc(&ctx->part_r, sizeof(double) * 2 * nb_part) &&
       dalloc(&ctx->part_c, sizeof(double) * 2 * nb_part) && dalloc(&ctx->d_red, sizeof(double) * 2);
  ok = ok && hipMalloc((void **)&ctx->d_ctl, sizeof(Ctl)) == hipSuccess &&
       hipMemsetAsync(ctx->d_ctl, 0, sizeof(Ctl), ctx->st) == hipSuccess && hipStreamSynchronize(ctx->st) == hipSuccess;
  ok = ok && hipHostMalloc((void **)&ctx->h_red, sizeof(double) * 2) == hipSuccess;
  ok = ok && hipHostMalloc((void **)&ctx->h_ctl, sizeof(Ctl)) == hipSuccess;
  for (int q = 0; q < 2 && ok && c.outbound; q++) {
    ok = dalloc(&ctx->grav[q], sizeof(double) * (PM_HDR + 2 * ctx->NI));
    if (ok && c.dtype == 1) ok = hipMalloc((void **)&ctx->gravf[q], sizeof(float) * 2 * ctx->NI) == hipSuccess;
  }
  if (c.vismodel) {
    ctx->nb_vis = visc_blocks(g);
    ok = ok && dalloc(&ctx->d_vsrc, sizeof(double) * (ctx->V - 1) * ctx->FS) &&
         dalloc(&ctx->d_vpart, sizeof(double) * 2 * ctx->nb_vis);
  }
  if (c.vismodel >= VIS_ALPHA)
    ok = ok && dalloc(&ctx->d_dyn, sizeof(double) * ctx->FS) && dalloc(&ctx->d_ephir, sizeof(double) * 3 * ctx->FS);
  for (int q = 0; q < 2 && c.vismodel == VIS_ALPHA && ctx->V == 4; q++)
    ok = ok && dalloc(&ctx->d_cs[q], sizeof(double) * ctx->FS);
  if (!ok) {
    fv_destroy(ctx);
    std::fprintf(stderr, "fv_create: device allocation failed\n");
    return -3;
  }
  g.ring = ctx->d_ring;
  g.col = ctx->d_col;
  g.ringf = ctx->d_ringf;
  g.colf = ctx->d_colf;
  if (upload_geom(ctx)) { fv_destroy(ctx); return -3; }
  if (c.vismodel >= VIS_ALPHA) {
    // e_phi/r = (-y/r^2, x/r^2) at the Cartesian bary centres, r^2 = radius^2
    // (kept as the third array, BETA's r**2)
    // (sources_viscosity.f90:196-223; geometry_logcylindrical.f90:344-356),
    // in curvilinear components (geometry_cylindrical.f90:485-494; the
    // Cartesian geometry keeps them)
    std::vector<double> e(3 * (size_t)ctx->FS, 0.0);
    for (int j = 0; j < ctx->NJ; j++)
      for (int i = 0; i < ctx->NI; i++) {
        const size_t k = (size_t)j * ctx->NI + i, ix = (size_t)j * ctx->P + i;
        const double phi = ctx->mesh.a[M_BCY][k], rad = ctx->mesh.a[M_RAD][k];
        double x = ctx->mesh.a[M_BCX][k], y = phi;
        if (c.geometry == LOGCYLINDRICAL) {
          const double r = c.gparam * std::exp(x);
          x = r * std::cos(phi);
          y = r * std::sin(phi);
        }
        const double r2 = rad * rad, ex = -(y / r2), ey = x / r2;
        e[2 * ctx->FS + ix] = r2;
        if (c.geometry == LOGCYLINDRICAL) {
          e[ix] = ex * std::cos(phi) + ey * std::sin(phi);
          e[ctx->FS + ix] = -(ex * std::sin(phi)) + ey * std::cos(phi);
        } else {
          e[ix] = ex;
          e[ctx->FS + ix] = ey;
        }
      }
    if (hipMemcpyAsync(ctx->d_ephir, e.data(), sizeof(double) * e.size(), hipMemcpyHostToDevice, ctx->st) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->st) != hipSuccess) {
      fv_destroy(ctx);
      return -3;
    }
  }
  if (c.outbound) {  // InitGravity_pointmass: mass; mdot, massloss, accrate = 0 (gravity_pointmass.f90:206-209)
    std::vector<double> gh(PM_HDR + 2 * (size_t)ctx->NI, 0.0);
    gh[PM_MASS] = c.mass;
    std::memcpy(&gh[PM_HDR], &ctx->ring_h[(size_t)R_ACCX * ctx->NI], sizeof(double) * 2 * ctx->NI);
    std::vector<float> gf(gh.begin() + PM_HDR, gh.end());
    bool up = true;
    for (int q = 0; q < 2; q++) {
      up = up && hipMemcpy(ctx->grav[q], gh.data(), sizeof(double) * gh.size(), hipMemcpyHostToDevice) == hipSuccess;
      if (ctx->gravf[q])
        up = up && hipMemcpy(ctx->gravf[q], gf.data(), sizeof(float) * gf.size(), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!up) { fv_destroy(ctx); return -3; }
  }
  // timedisc_base.f90:295-352 defaults
  ctx->time = c.starttime;
  ctx->dt = c.stoptime;
  ctx->dtold = ctx->dtmin = c.stoptime;
  ctx->dtlimit = c.dtlimit > 0 ? c.dtlimit : DBL_EPSILON * c.stoptime;
  // the output clock (io/fileio_base.f90:474-487, 591-600) starts at the
  // start time; one IncTime happens before the first Step (in InitFileIO
  // after a restart, after FirstStep's initial dataset otherwise)
  ctx->count = c.count > 0 ? c.count : 1;
  ctx->t_out = ctx->time + std::fabs(c.stoptime - ctx->time) / ctx->count;
  ctx->iter = -1;
  *out = ctx;
  return 0;
}

#ifndef FV_BUILD_ID
#define FV_BUILD_ID "unknown"
#endif
extern "C" const char *fv_build_id(void) { return FV_BUILD_ID; }

extern "C" int fv_destroy(fv_ctx *ctx) {
  if (!ctx) return 0;
  if (ctx->batch) (void)fv_batch_remove(ctx->batch, ctx);
  if (ctx->st) (void)hipStreamSynchronize(ctx->st);
  for (auto &e : ctx->graphs) (void)hipGraphExecDestroy(e.exec);
  for (auto &b : ctx->buf) if (b) (void)hipFree(b);
  for (auto &b : ctx->k) if (b) (void)hipFree(b);
  for (auto &b : ctx->k0) if (b) (void)hipFree(b);
  for (auto &b : ctx->bf) if (b) (void)hipFree(b);
  for (auto &b : ctx->grav) if (b) (void)hipFree(b);
  for (auto &b : ctx->gravf) if (b) (void)hipFree(b);
  for (auto &b : ctx->vbufs) if (b) (void)hipFree(b);
  if (ctx->pbuf) (void)hipFree(ctx->pbuf);
  if (ctx->d_vsrc) (void)hipFree(ctx->d_vsrc);
  if (ctx->d_vpart) (void)hipFree(ctx->d_vpart);
  for (auto &b : ctx->d_cs) if (b) (void)hipFree(b);
  if (ctx->d_dyn) (void)hipFree(ctx->d_dyn);
  if (ctx->d_ephir) (void)hipFree(ctx->d_ephir);
  if (ctx->d_ring1) (void)hipFree(ctx->d_ring1);
  if (ctx->d_ringf1) (void)hipFree(ctx->d_ringf1);
  if (ctx->bks) (void)hipFree(ctx->bks);
  if (ctx->part_r) (void)hipFree(ctx->part_r);
  if (ctx->part_c) (void)hipFree(ctx->part_c);
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  if (ctx->d_ctl) (void)hipFree(ctx->d_ctl);
  if (ctx->d_ring) (void)hipFree(ctx->d_ring);
  if (ctx->d_col) (void)hipFree(ctx->d_col);
  if (ctx->d_ringf) (void)hipFree(ctx->d_ringf);
  if (ctx->d_colf) (void)hipFree(ctx->d_colf);
  if (ctx->h_red) (void)hipHostFree(ctx->h_red);
  if (ctx->h_ctl) (void)hipHostFree(ctx->h_ctl);
  for (auto e : ctx->evpool) (void)hipEventDestroy(e);
  if (ctx->e0) (void)hipEventDestroy(ctx->e0);
  if (ctx->e1) (void)hipEventDestroy(ctx->e1);
  if (ctx->st) (void)hipStreamDestroy(ctx->st);
  delete ctx;
  return 0;
}

extern "C" int fv_device_bytes(const fv_config *c, double *bytes) {
  if (!c || !bytes) return -1;
  Tableau t;
  if (c->inum < 4 || c->jnum < 4 || (c->dtype != 0 && c->dtype != 1) || !set_tableau(c->method, c->order, t))
    return -1;
  const double NI = c->inum + 4.0, NJ = c->jnum + 4.0, P = std::ceil(NI / 16.0) * 16.0;
  const double V = c->physics == EULER ? 4.0 : 3.0, es = c->dtype == 1 ? 4.0 : 8.0;
  const double state = es * V * NJ * P;  // one state buffer, ghost cells and row padding included
  const double nb32 = std::ceil(c->inum / 32.0) * std::ceil(c->jnum / (double)TY_TILE);
  double b = state * (NBUF + (t.m - 1) + 2 + 1);      // states, stage rhs, coeff(1) pair, pvar
  b += 3.0 * 8.0 * 2.0 * V * (NI + NJ);                // boundary fluxes (two copies), last-stage rows
  if (c->fargo == 1)                                    // FARGO type 1: velocities, second ring arrays
    b += 2.0 * 8.0 * NJ * P + (8.0 + (c->dtype == 1 ? 4.0 : 0.0)) * R_COUNT * NI;
  if (c->outbound) b += 2.0 * (8.0 * (PM_HDR + 2.0 * NI) + (c->dtype == 1 ? 4.0 * 2.0 * NI : 0.0));
  if (c->vismodel)                                      // viscous source, its step partials
    b += 8.0 * (V - 1.0) * NJ * P + 16.0 * std::ceil(NI / 64.0) * std::ceil(NJ / 8.0);
  if (c->vismodel >= 2)                                 // viscosity, e_phi/r, r^2; Euler ALPHA: sound speeds
    b += (c->vismodel == 2 && V == 4 ? 6.0 : 4.0) * 8.0 * NJ * P;
  b += 2.0 * 2.0 * 8.0 * nb32;                         // CFL and error partials
  b += 8.0 * (R_COUNT * NI + C_COUNT * NJ);             // metric arrays
  if (c->dtype == 1) b += 4.0 * (R_COUNT * NI + C_COUNT * NJ);
  b += sizeof(Ctl) + 16.0;
  *bytes = b;
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

// the conservative state of all cells into buffer 0 (rounded to FP32 in the
// precision form); the stored state then starts in the real frame
static int upload_state(fv_ctx *ctx, const double *cv_in) {
  const int V = ctx->V;
  const size_t N = (size_t)ctx->NI * ctx->NJ;
  ctx->cvar = ctx->cold = 0;
  const size_t es = elem_bytes(ctx);
  std::vector<float> cf;
  const void *src = cv_in;
  if (ctx->c.dtype == 1) {  // precision form: conserved variables from FP64, stored rounded
    cf.assign(cv_in, cv_in + V * N);
    src = cf.data();
  }
  for (int l = 0; l < V; l++)
    HIPCHK(hipMemcpy2DAsync(field(ctx, ctx->buf[0], l), es * ctx->P, static_cast<const char *>(src) + l * N * es,
                            es * ctx->NI, es * ctx->NI, ctx->NJ, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->fl_pvar = ctx->fl_cvar = 0;
  ctx->have_state = true;
  ctx->pvalid = false;
  return 0;
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
  return upload_state(ctx, cv.data());
}

extern "C" int fv_set_cvar(fv_ctx *ctx, const double *cvar) {
  if (!ctx || !cvar) return -1;
  return upload_state(ctx, cvar);
}

extern "C" int fv_first_rhs(fv_ctx *ctx, int *brk) {
  if (!ctx) return -1;
  if (!ctx->have_state) return fail(ctx, "initial data not set");
  if (ctx->V == 3 && !ctx->have_cs) return fail(ctx, "isothermal physics needs sound speeds");
  if (ctx->c.fargo == 2 && !ctx->have_w) return fail(ctx, "fargo = 2 needs the background velocity");
  if (ctx->iter > -1) return fail(ctx, "FirstStep should only be called once");
  ctx->iter = 0;
  int rc;
  if (ctx->c.fargo == 1) {
    // CenterBoundary, then CalcBackgroundVelocity (fosite.f90:305-311): one
    // evaluation with w = 0 on the initial state only for its velocities
    // (the frame flags are clear, so the frame change is a no-op)
    StageArgs a = base_args(ctx);
    a.Y = ctx->buf[ctx->cvar];
    a.kout = ctx->k0[1 - ctx->k0cur];
    a.sub_bg = 1;
    a.bflux = 0;
    a.vout = ctx->vbufs[ctx->vcur];
    launch_rhs(ctx->g, a, ctx->V, ctx->st);
    HIPCHK(hipGetLastError());
    if ((rc = calc_background_velocity(ctx))) return rc;
  }
  if ((rc = accrete(ctx, ctx->bcur, 0.0))) return rc;  // ComputeRHS(..., 0.0, ...) (fosite.f90:315-317)
  rc = rhs_main(ctx, CHECK_ALL, 0, ctx->time);
  if (rc) return rc;
  if ((rc = timestep(ctx))) return rc;
  ctx->cold = ctx->cvar;
  ctx->fl_cold = ctx->fl_cvar;
  if (brk) *brk = ctx->brk;
  return 0;
}

extern "C" int fv_first_step(fv_ctx *ctx, int *brk) {
  int rc = fv_first_rhs(ctx, brk);
  if (rc) return rc;
  // initial WriteDataset moves the state to the real frame (io/fileio_base.f90:523-533)
  if (ctx->time == 0.0 && ctx->c.fargo && (rc = add_bg_main(ctx))) return rc;
  return 0;
}

extern "C" int fv_integration_step(fv_ctx *ctx, double dt, int dtcause, int *brk) {
  if (!ctx) return -1;
  if (ctx->iter < 0) return fail(ctx, "fv_integration_step before the first step");
  if (!(dt > 0.0)) return fail(ctx, "fv_integration_step: dt must be positive");
  ctx->dt = dt;
  ctx->dtcause = dtcause;
  int rc = integration_step(ctx);
  if (brk) *brk = ctx->brk;
  return rc;
}

extern "C" int fv_calc_timestep(fv_ctx *ctx, double *dt, int *dtcause) {
  if (!ctx) return -1;
  if (ctx->iter < 0) return fail(ctx, "fv_calc_timestep before the first step");
  int rc;
  if (pvar_moved_to_real_frame(ctx) && (rc = cfl_real_frame(ctx))) return rc;
  if ((rc = timestep(ctx))) return rc;
  if (dt) *dt = ctx->dt;
  if (dtcause) *dtcause = ctx->dtcause;
  return 0;
}

extern "C" int fv_add_background_velocity(fv_ctx *ctx) {
  if (!ctx) return -1;
  return ctx->c.fargo ? add_bg_main(ctx) : 0;
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
    if (ctx->c.fargo && (rc = add_bg_main(ctx))) return rc;  // WriteDataset
    if (ctx->brk) {
      if (done) *done = 1;
      return 0;
    }
    if (pvar_moved_to_real_frame(ctx) && (rc = cfl_real_frame(ctx))) return rc;
    ctx->t_out = ctx->t_out + std::fabs(ctx->c.stoptime - ctx->c.starttime) / ctx->count;  // IncTime
    ctx->dtmin = ctx->c.stoptime;
    ctx->n_adj = 0;
  }
  return timestep(ctx);
}

// fv_step with the reduction scratch of its kernels (per-workgroup partials
// of the end-of-step evaluation and of the last stage, the reduction result)
// in the library's static device array instead of the context's buffers: the
// call allocates no device or pinned memory.  Every call on one device shares
// that array; the partials are written by every workgroup before they are
// reduced, so nothing in it has to start at zero.
extern "C" int fv_step_noalloc(fv_ctx *ctx, int *done) {
  if (!ctx) return -1;
  if (ctx->prof) return fail(ctx, "fv_step_noalloc: kernel profiling (fv_profile) creates events; disable it");
  if (ctx->nb_rhs > SCRATCH_MAX_BLOCKS)
    return fail(ctx, "fv_step_noalloc: " + std::to_string(ctx->nb_rhs) + " workgroups per RHS launch exceed the " +
                         std::to_string(SCRATCH_MAX_BLOCKS) + " the static scratch holds (use fv_step)");
  int cur = -1;
  HIPCHK(hipGetDevice(&cur));
  if (cur != ctx->dev) return fail(ctx, "fv_step_noalloc: the context's device is not the current device");
  double *s = nullptr;
  HIPCHK(scratch_address(ctx->dev, &s));
  double *const own[3] = {ctx->part_r, ctx->part_c, ctx->d_red};
  ctx->part_r = s + SCRATCH_PART_R;
  ctx->part_c = s + SCRATCH_PART_C;
  ctx->d_red = s + SCRATCH_RED;
  const int rc = fv_step(ctx, done);
  ctx->part_r = own[0];
  ctx->part_c = own[1];
  ctx->d_red = own[2];
  return rc;
}

extern "C" int fv_run(fv_ctx *ctx, int nsteps, int *taken) {
  if (!ctx) return -1;
  int n = 0, done = 0, rc = 0;
  HIPCHK(hipEventRecord(ctx->e0, ctx->st));
  while (n < nsteps && !done) {
    if (fast_eligible(ctx) && nsteps - n >= 2) {
      int k = 0, halted = 0;
      const int want = std::min(nsteps - n, CTL_HIST);
      if ((rc = run_batch(ctx, want, &k, &halted))) break;
      n += k;
      if (!halted || n >= nsteps) continue;
    }
    if ((rc = fv_step(ctx, &done))) break;
    if (!done) n++;
  }
  HIPCHK(hipEventRecord(ctx->e1, ctx->st));
  HIPCHK(hipEventSynchronize(ctx->e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->e0, ctx->e1));
  ctx->last_ms = ms;
  if (taken) *taken = n;
  return rc;
}

extern "C" int fv_get_state(fv_ctx *ctx, double *cvar, double *pvar) {
  if (!ctx) return -1;
  const int V = ctx->V;
  const size_t N = (size_t)ctx->NI * ctx->NJ;
  std::vector<double> cv(V * N);
  std::vector<float> cf(ctx->c.dtype == 1 ? V * N : 0);
  const size_t es = elem_bytes(ctx);
  char *dst = ctx->c.dtype == 1 ? reinterpret_cast<char *>(cf.data()) : reinterpret_cast<char *>(cv.data());
  HIPCHK(hipStreamSynchronize(ctx->st));
  for (int l = 0; l < V; l++)
    HIPCHK(hipMemcpy2DAsync(dst + l * N * es, es * ctx->NI, field(ctx, ctx->buf[ctx->cvar], l), es * ctx->P,
                            es * ctx->NI, ctx->NJ, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (ctx->c.dtype == 1) std::copy(cf.begin(), cf.end(), cv.begin());
  if (cvar) std::memcpy(cvar, cv.data(), sizeof(double) * V * N);
  if (pvar && ctx->pvalid) {
    // Timedisc%pvar of the last ComputeRHS, in the frame of the state: after
    // an output-time AddBackgroundVelocityY (physics_eulerisotherm.f90:
    // 1263-1303) the same + w as the reference applies to its pvar
    std::vector<double> pv(V * N);
    std::vector<float> pf(ctx->c.dtype == 1 ? V * N : 0);
    char *pd = ctx->c.dtype == 1 ? reinterpret_cast<char *>(pf.data()) : reinterpret_cast<char *>(pv.data());
    for (int l = 0; l < V; l++)
      HIPCHK(hipMemcpy2DAsync(pd + l * N * es, es * ctx->NI, field(ctx, ctx->pbuf, l), es * ctx->P, es * ctx->NI,
                              ctx->NJ, hipMemcpyDeviceToHost, ctx->st));
    std::vector<double> w(ctx->NI, 0.0);
    const bool addw = ctx->c.fargo && ctx->pflag && !ctx->fl_pvar;
    if (addw)
      HIPCHK(hipMemcpyAsync(w.data(), ring_tab(ctx, ctx->wcur) + (size_t)R_W * ctx->NI, sizeof(double) * ctx->NI,
                            hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    if (ctx->c.dtype == 1) std::copy(pf.begin(), pf.end(), pv.begin());
    if (addw)
      for (int j = 0; j < ctx->NJ; j++)
        for (int i = 0; i < ctx->NI; i++) pv[2 * N + (size_t)j * ctx->NI + i] += w[i];
    std::memcpy(pvar, pv.data(), sizeof(double) * V * N);
  } else if (pvar) {  // before FirstStep: Convert2Primitive of the state
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

extern "C" int fv_get_background_velocity(fv_ctx *ctx, double *w) {
  if (!ctx || !w) return -1;
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipMemcpyAsync(w, ring_tab(ctx, ctx->wcur) + (size_t)R_W * ctx->NI, sizeof(double) * ctx->NI, hipMemcpyDeviceToHost,
                        ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

extern "C" int fv_get_boundary_flux(fv_ctx *ctx, double *bx, double *by) {
  if (!ctx) return -1;
  const size_t nx = (size_t)2 * ctx->V * ctx->NJ, ny = (size_t)2 * ctx->V * ctx->NI;
  HIPCHK(hipStreamSynchronize(ctx->st));
  // the device layout is the reference's (ghost entries stay zero)
  if (bx) HIPCHK(hipMemcpyAsync(bx, ctx->bf[ctx->bcur], sizeof(double) * nx, hipMemcpyDeviceToHost, ctx->st));
  if (by) HIPCHK(hipMemcpyAsync(by, ctx->bf[ctx->bcur] + nx, sizeof(double) * ny, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

extern "C" int fv_get_point_mass(fv_ctx *ctx, double *out) {
  if (!ctx) return -1;
  if (!out) return fail(ctx, "fv_get_point_mass: null output");
  if (!ctx->c.gravity) return fail(ctx, "fv_get_point_mass: no point mass (gravity = 0)");
  if (!ctx->grav[0]) {
    out[PM_MASS] = ctx->c.mass;
    out[PM_ACCRATE] = out[PM_MASSLOSS] = out[PM_MDOT] = 0.0;
    return 0;
  }
  HIPCHK(hipMemcpyAsync(out, ctx->grav[ctx->gcur], sizeof(double) * PM_HDR, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

extern "C" int fv_get_boundary_flux_total(fv_ctx *ctx, int direction, double *bflux) {
  if (!ctx) return -1;
  if (!bflux || direction < 1 || direction > 4) return fail(ctx, "fv_get_boundary_flux_total: direction 1..4");
  const int V = ctx->V, NI = ctx->NI, NJ = ctx->NJ;
  std::vector<double> bx((size_t)2 * V * NJ), by((size_t)2 * V * NI);
  int rc = fv_get_boundary_flux(ctx, bx.data(), by.data());
  if (rc) return rc;
  const bool xd = direction <= 2;
  const int side = (direction - 1) & 1, n = xd ? NJ : NI, lo = 2, hi = (xd ? ctx->c.jnum : ctx->c.inum) + 1;
  for (int l = 0; l < V; l++) {  // SUM over JMIN:JMAX (IMIN:IMAX), in index order
    const double *a = (xd ? bx.data() : by.data()) + (size_t)(l * 2 + side) * n;
    double acc = 0.0;
    for (int k = lo; k <= hi; k++) acc = acc + a[k];
    bflux[l] = acc;
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

extern "C" int fv_profile(fv_ctx *ctx, int enable) {
  if (!ctx) return -1;
  HIPCHK(hipStreamSynchronize(ctx->st));
  prof_collect(ctx);
  ctx->prof = enable != 0;
  for (int q = 0; q < K_CLASSES; q++) ctx->st_launch[q] = 0, ctx->st_ms[q] = ctx->st_bytes[q] = 0;
  return 0;
}

extern "C" int fv_kernel_stats(const fv_ctx *ctx, int kernel, int *launches, double *ms, double *alg_bytes) {
  if (!ctx || kernel < 0 || kernel >= K_CLASSES) return -1;
  if (launches) *launches = ctx->st_launch[kernel];
  if (ms) *ms = ctx->st_ms[kernel];
  if (alg_bytes) *alg_bytes = ctx->st_bytes[kernel];
  return 0;
}

extern "C" const char *fv_last_error(const fv_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// ================================================================ batches
// Many contexts advanced together.  A round of device-controlled steps
// launches every kernel of a step once for all members: the per-member
// arguments of each launch sit in tables in device memory (rewritten before
// every round, so a captured graph stays valid while members' buffer roles
// change), and block b of a launch runs the tile tiles[b] of member
// tile_disk(tiles[b]).  Each member has its own Ctl; a member that reaches
// its step count or a step the device cannot decide halts on its own, its
// blocks then return at once, and the host replays or redoes its steps with
// the single-context code above.

namespace {
constexpr int NCAND = 4;     // RHS tilings: 64 row-major, 64 XCD-chunked, 32 row-major, 32 XCD-chunked
constexpr int XCDS = 8;      // MI355X: workgroup b of a launch is dispatched to XCD b mod 8
constexpr int TUNE_REPS = 2;
int cand_tx(int c) { return c < 2 ? 64 : 32; }
}  // namespace

struct fv_batch {
  std::string err;
  fv_batch_opts opt{};
  std::vector<fv_ctx *> sims;
  int V = 0, dtype = 0, m = 0, limiter = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr, t0 = nullptr, t1 = nullptr;
  int cap = 0;                       // members the tables hold
  Ctl *d_ctl = nullptr, *h_ctl = nullptr;
  RhsEntry *d_rhs = nullptr, *h_rhs = nullptr;          // [2 parities][m launches][cap]
  RemapEntry *d_remap = nullptr, *h_remap = nullptr;    // [2][cap]
  ControlEntry *d_cent = nullptr, *h_cent = nullptr;    // [NCAND][cap]
  // point-mass accretion members: the step's boundary-flux accumulation and
  // UpdateGravity before the end-of-step evaluation, [2 parities][cap] each,
  // acc_n entries used (members in order), acc_bpm workgroups per member
  BfluxArgs *d_bfl = nullptr, *h_bfl = nullptr;
  AccreteArgs *d_acc = nullptr, *h_acc = nullptr;
  int acc_n = 0, acc_bpm = 0;
  unsigned *d_tiles = nullptr;
  size_t tiles_cap = 0;
  int cand_off[NCAND] = {}, cand_n[NCAND] = {}, remap_off = 0, remap_n = 0;
  bool dirty = true;
  int cand = 0, tuned = 0, trial = 0;
  double tune_ms[NCAND] = {};
  struct G { int K, cand; hipGraphExec_t exec; };
  std::vector<G> graphs;
  struct Roles { int x0, x1, tb, rb, k0, b0; };
  std::vector<Roles> roles;
  long dev_steps = 0, host_steps = 0, halts = 0, rounds = 0;
  double last_ms = 0;
};

#define BCHK(x)                                                                \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      b->err = std::string(#x) + ": " + hipGetErrorString(e_);                 \
      return -3;                                                               \
    }                                                                          \
  } while (0)

namespace {

// limiter a batch launch compiles in for this member: 0 (the run-time
// switch) for PRIMITIVE reconstruction, which only that instantiation carries
int launch_limiter(const fv_ctx *x) { return x->c.recvars == 1 ? 0 : x->c.limiter; }

int bfail(fv_batch *b, const std::string &m, int code = -1) {
  b->err = m;
  return code;
}

void drop_graphs(fv_batch *b) {
  for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.exec);
  b->graphs.clear();
}

void free_tables(fv_batch *b) {
  if (b->d_ctl) (void)hipFree(b->d_ctl);
  if (b->d_rhs) (void)hipFree(b->d_rhs);
  if (b->d_remap) (void)hipFree(b->d_remap);
  if (b->d_cent) (void)hipFree(b->d_cent);
  if (b->h_ctl) (void)hipHostFree(b->h_ctl);
  if (b->h_rhs) (void)hipHostFree(b->h_rhs);
  if (b->h_remap) (void)hipHostFree(b->h_remap);
  if (b->h_cent) (void)hipHostFree(b->h_cent);
  if (b->d_bfl) (void)hipFree(b->d_bfl);
  if (b->h_bfl) (void)hipHostFree(b->h_bfl);
  if (b->d_acc) (void)hipFree(b->d_acc);
  if (b->h_acc) (void)hipHostFree(b->h_acc);
  b->d_bfl = b->h_bfl = nullptr;
  b->d_acc = b->h_acc = nullptr;
  b->d_ctl = b->h_ctl = nullptr;
  b->d_rhs = b->h_rhs = nullptr;
  b->d_remap = b->h_remap = nullptr;
  b->d_cent = b->h_cent = nullptr;
  b->cap = 0;
}

// k_rhs tile list of tiling c over all members
std::vector<unsigned> rhs_tiles(const fv_batch *b, int c) {
  const int tx = cand_tx(c);
  std::vector<unsigned> lin;
  for (size_t d = 0; d < b->sims.size(); d++) {
    const fv_ctx *x = b->sims[d];
    const int nbx = (x->c.inum + tx - 1) / tx, nby = (x->c.jnum + TY_TILE - 1) / TY_TILE;
    for (int by = 0; by < nby; by++)
      for (int bx = 0; bx < nbx; bx++) lin.push_back(tile_code((int)d, bx, by));
  }
  if (!(c & 1)) return lin;
  // XCD-chunked: tile t goes to XCD t / Q, at that XCD's slot t % Q
  const size_t N = lin.size(), Q = (N + XCDS - 1) / XCDS;
  std::vector<unsigned> out(Q * XCDS, NO_TILE);
  for (size_t t = 0; t < N; t++) out[(t % Q) * XCDS + t / Q] = lin[t];
  return out;
}

// tables and tile lists for the current members; graphs are recaptured
int rebuild(fv_batch *b) {
  drop_graphs(b);
  const int n = (int)b->sims.size();
  if (n > b->cap) {
    free_tables(b);
    const int cap = std::max(n, 8);
    const int m = b->m;
    BCHK(hipMalloc((void **)&b->d_ctl, sizeof(Ctl) * cap));
    BCHK(hipMalloc((void **)&b->d_rhs, sizeof(RhsEntry) * 2 * m * cap));
    BCHK(hipMalloc((void **)&b->d_remap, sizeof(RemapEntry) * 2 * cap));
    BCHK(hipMalloc((void **)&b->d_cent, sizeof(ControlEntry) * NCAND * cap));
    BCHK(hipHostMalloc((void **)&b->h_ctl, sizeof(Ctl) * cap));
    BCHK(hipHostMalloc((void **)&b->h_rhs, sizeof(RhsEntry) * 2 * m * cap));
    BCHK(hipHostMalloc((void **)&b->h_remap, sizeof(RemapEntry) * 2 * cap));
    BCHK(hipHostMalloc((void **)&b->h_cent, sizeof(ControlEntry) * NCAND * cap));
    BCHK(hipMalloc((void **)&b->d_bfl, sizeof(BfluxArgs) * 2 * cap));
    BCHK(hipMalloc((void **)&b->d_acc, sizeof(AccreteArgs) * 2 * cap));
    BCHK(hipHostMalloc((void **)&b->h_bfl, sizeof(BfluxArgs) * 2 * cap));
    BCHK(hipHostMalloc((void **)&b->h_acc, sizeof(AccreteArgs) * 2 * cap));
    BCHK(hipMemsetAsync(b->d_ctl, 0, sizeof(Ctl) * cap, b->st));
    b->cap = cap;
  }
  b->acc_n = b->acc_bpm = 0;
  for (int d = 0; d < n; d++) {
    const fv_ctx *x = b->sims[d];
    if (!x->grav[0]) continue;
    b->acc_n++;
    b->acc_bpm = std::max(b->acc_bpm, (x->V * 2 * (x->c.inum + x->c.jnum) + 255) / 256);
  }
  std::vector<unsigned> all;
  for (int c = 0; c < NCAND; c++) {
    std::vector<unsigned> t = rhs_tiles(b, c);
    b->cand_off[c] = (int)all.size();
    b->cand_n[c] = (int)t.size();
    all.insert(all.end(), t.begin(), t.end());
  }
  b->remap_off = (int)all.size();
  for (int d = 0; d < n; d++) {
    const fv_ctx *x = b->sims[d];
    if (x->c.fargo != 2) continue;
    const int nbx = (x->c.inum + 63) / 64, nby = remap_rows_blocks(x->c.jnum);
    for (int by = 0; by < nby; by++)
      for (int bx = 0; bx < nbx; bx++) all.push_back(tile_code(d, bx, by));
  }
  b->remap_n = (int)all.size() - b->remap_off;
  if (all.size() > b->tiles_cap) {
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    b->d_tiles = nullptr;
    b->tiles_cap = 0;
    BCHK(hipMalloc((void **)&b->d_tiles, sizeof(unsigned) * all.size()));
    b->tiles_cap = all.size();
  }
  if (!all.empty())
    BCHK(hipMemcpyAsync(b->d_tiles, all.data(), sizeof(unsigned) * all.size(), hipMemcpyHostToDevice, b->st));
  BCHK(hipStreamSynchronize(b->st));
  b->dirty = false;
  return 0;
}

// the launches of one device step of parity p under tiling c
void enqueue_batch_step(fv_batch *b, int p, int c) {
  const int m = b->m, cap = b->cap, n = (int)b->sims.size();
  const unsigned *tl = b->d_tiles + b->cand_off[c];
  for (int q = 0; q < m - 1; q++)  // stages 2..s, the last fused with the combination
    launch_rhs_multi(b->d_rhs + (size_t)(p * m + q) * cap, tl, b->cand_n[c], b->V, b->dtype, b->limiter,
                     cand_tx(c), b->m, q == m - 2 ? 1 : 0, b->st);
  if (b->remap_n)
    launch_remap_multi(b->d_remap + (size_t)p * cap, b->d_tiles + b->remap_off, b->remap_n, b->V, b->dtype, b->st);
  if (b->acc_n) {  // the step's boundary fluxes, then UpdateGravity
    launch_bflux_multi(b->d_bfl + (size_t)p * cap, b->acc_n, b->acc_bpm, b->st);
    launch_accrete_multi(b->d_acc + (size_t)p * cap, b->acc_n, b->st);
  }
  launch_rhs_multi(b->d_rhs + (size_t)(p * m + m - 1) * cap, tl, b->cand_n[c], b->V, b->dtype, b->limiter,
                   cand_tx(c), b->m, 2, b->st);
  launch_control_multi(b->d_cent + (size_t)c * cap, n, b->st);
}

int batch_graph(fv_batch *b, int K, int c, hipGraphExec_t *out) {
  for (auto &g : b->graphs)
    if (g.K == K && g.cand == c) {
      *out = g.exec;
      return 0;
    }
  hipGraph_t graph;
  BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < K; s++) enqueue_batch_step(b, s % 2, c);
  BCHK(hipStreamEndCapture(b->st, &graph));
  hipGraphExec_t exec;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  BCHK(e);
  b->graphs.push_back({K, c, exec});
  *out = exec;
  return 0;
}

// the autotuner's choices by batch class: element type, fields, stages, and
// the log2 of the members' total cells and of their number; a later batch of
// a class already tuned in this process takes the choice without trials
std::mutex tune_mu;
std::map<long, int> tune_cache;
long tune_key(const fv_batch *b) {
  double cells = 0;
  for (const fv_ctx *x : b->sims) cells += (double)x->c.inum * x->c.jnum;
  const long lc = (long)std::floor(std::log2(std::max(cells, 1.0)));
  const long ln = (long)std::floor(std::log2((double)std::max<size_t>(b->sims.size(), 1)));
  return ((((long)b->dtype * 8 + b->V) * 16 + b->m) * 64 + lc) * 16 + ln;
}

// one round: want[d] device steps of member d (0 = not in this round), K =
// steps launched (>= every want), tiling c; timed rounds are issued eagerly
// between events and return their device time in *ms.  done[d] = steps the
// device completed, halted[d] = 1 if member d stopped on a step the host must redo.
int batch_round(fv_batch *b, const std::vector<int> &want, int K, int c, bool timed, double *ms,
                std::vector<int> &done, std::vector<int> &halted) {
  const int n = (int)b->sims.size(), m = b->m, cap = b->cap;
  b->roles.resize(n);
  int ai = 0;  // entry of the next accreting member
  for (int d = 0; d < n; d++) {
    fv_ctx *x = b->sims[d];
    Ctl *h = &b->h_ctl[d];
    Ctl *dc = b->d_ctl + d;
    if (want[d] > 0) {
      prime_ctl(x, h, want[d]);
    } else {
      std::memset(h, 0, offsetof(Ctl, err_hist));
      h->halt = 1;
    }
    fv_batch::Roles &r = b->roles[d];
    r.x0 = x->cvar;
    r.x1 = pick_free(x, r.x0, -1, -1);
    r.tb = pick_free(x, r.x0, r.x1, -1);
    r.rb = pick_free(x, r.x0, r.x1, r.tb);
    r.k0 = x->k0cur;
    r.b0 = x->bcur;
    for (int p = 0; p < 2; p++) {
      FastStep f;
      const int gin = (p && x->grav[0]) ? 1 - x->gcur : x->gcur;  // the point mass alternates as the state
      fast_step_args(x, dc, p ? r.x1 : r.x0, p ? r.x0 : r.x1, r.tb, r.rb, p ? 1 - r.k0 : r.k0,
                     p ? r.k0 : 1 - r.k0, p ? 1 - r.b0 : r.b0, gin, x->wcur, x->vcur, f);  // no FARGO type 1
      if (f.accrete) {
        b->h_bfl[(size_t)p * cap + ai] = f.bfl;
        b->h_acc[(size_t)p * cap + ai] = f.acc;
      }
      for (int q = 0; q < m - 1; q++) b->h_rhs[(size_t)(p * m + q) * cap + d] = RhsEntry{x->g, f.st[q]};
      b->h_rhs[(size_t)(p * m + m - 1) * cap + d] = RhsEntry{x->g, f.end};
      RemapEntry re;
      std::memset(&re, 0, sizeof(re));
      re.g = x->g;
      re.a = f.remap ? f.rm : RemapArgs{nullptr, nullptr, 0.0, dc};
      b->h_remap[(size_t)p * cap + d] = re;
    }
    for (int cc = 0; cc < NCAND; cc++) {
      const int tx = cand_tx(cc);
      const int nb = ((x->c.inum + tx - 1) / tx) * ((x->c.jnum + TY_TILE - 1) / TY_TILE);
      b->h_cent[(size_t)cc * cap + d] = ControlEntry{dc, x->part_c, x->part_r, nb};
    }
    if (x->grav[0]) ai++;
  }
  BCHK(hipMemcpy2DAsync(b->d_ctl, sizeof(Ctl), b->h_ctl, sizeof(Ctl), offsetof(Ctl, err_hist), n,
                        hipMemcpyHostToDevice, b->st));
  BCHK(hipMemcpyAsync(b->d_rhs, b->h_rhs, sizeof(RhsEntry) * 2 * m * cap, hipMemcpyHostToDevice, b->st));
  BCHK(hipMemcpyAsync(b->d_remap, b->h_remap, sizeof(RemapEntry) * 2 * cap, hipMemcpyHostToDevice, b->st));
  BCHK(hipMemcpyAsync(b->d_cent, b->h_cent, sizeof(ControlEntry) * NCAND * cap, hipMemcpyHostToDevice, b->st));
  if (b->acc_n) {
    BCHK(hipMemcpyAsync(b->d_bfl, b->h_bfl, sizeof(BfluxArgs) * 2 * cap, hipMemcpyHostToDevice, b->st));
    BCHK(hipMemcpyAsync(b->d_acc, b->h_acc, sizeof(AccreteArgs) * 2 * cap, hipMemcpyHostToDevice, b->st));
  }
  if (timed) {
    BCHK(hipEventRecord(b->t0, b->st));
    for (int s = 0; s < K; s++) enqueue_batch_step(b, s % 2, c);
    BCHK(hipGetLastError());
    BCHK(hipEventRecord(b->t1, b->st));
  } else if (eager_launches()) {
    for (int s = 0; s < K; s++) enqueue_batch_step(b, s % 2, c);
    BCHK(hipGetLastError());
  } else {
    hipGraphExec_t exec;
    int rc = batch_graph(b, K, c, &exec);
    if (rc) return rc;
    BCHK(hipGraphLaunch(exec, b->st));
  }
  BCHK(hipMemcpyAsync(b->h_ctl, b->d_ctl, sizeof(Ctl) * n, hipMemcpyDeviceToHost, b->st));
  BCHK(hipStreamSynchronize(b->st));
  b->rounds++;
  if (timed) {
    float f = 0;
    BCHK(hipEventElapsedTime(&f, b->t0, b->t1));
    *ms = f;
  }
  done.assign(n, 0);
  halted.assign(n, 0);
  for (int d = 0; d < n; d++) {
    if (want[d] <= 0) continue;
    fv_ctx *x = b->sims[d];
    int hl = 0;
    int rc = replay_device_steps(x, &b->h_ctl[d], b->roles[d].x0, b->roles[d].x1, &hl);
    if (rc) return bfail(b, "member " + std::to_string(d) + ": " + x->err, rc);
    done[d] = b->h_ctl[d].step;
    halted[d] = hl;
  }
  return 0;
}

}  // namespace

extern "C" int fv_batch_create(fv_batch **out, const fv_batch_opts *opts) {
  if (!out) return -1;
  *out = nullptr;
  fv_batch *b = new (std::nothrow) fv_batch();
  if (!b) return -1;
  if (opts) b->opt = *opts;
  if (b->opt.tiling < 0 || b->opt.tiling >= NCAND || b->opt.max_graph_steps < 0 || b->opt.max_graph_steps > CTL_HIST) {
    delete b;
    std::fprintf(stderr, "fv_batch_create: tiling must be 0..3 and max_graph_steps 0..%d\n", CTL_HIST);
    return -1;
  }
  if (b->opt.max_graph_steps == 0) b->opt.max_graph_steps = CTL_HIST;
  b->cand = b->opt.tiling;
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&b->e0) != hipSuccess ||
      hipEventCreate(&b->e1) != hipSuccess || hipEventCreate(&b->t0) != hipSuccess ||
      hipEventCreate(&b->t1) != hipSuccess) {
    fv_batch_destroy(b);
    std::fprintf(stderr, "fv_batch_create: stream or event creation failed\n");
    return -3;
  }
  *out = b;
  return 0;
}

extern "C" int fv_batch_add(fv_batch *b, fv_ctx *x) {
  if (!b || !x) return -1;
  if (x->batch) return bfail(b, "context already belongs to a batch");
  if ((int)b->sims.size() >= MAX_BATCH) return bfail(b, "batch is full (" + std::to_string(MAX_BATCH) + " members)");
  if (x->c.vismodel) return bfail(b, "viscosity is supported on single contexts, not in batches");
  const int nbx = (x->c.inum + 31) / 32, nby = (x->c.jnum + TY_TILE - 1) / TY_TILE, rby = remap_rows_blocks(x->c.jnum);
  if (nbx >= (1 << TILE_BX_BITS) || nby >= (1 << TILE_BY_BITS) || rby >= (1 << TILE_BY_BITS))
    return bfail(b, "mesh too large for a batch tile code");
  if (b->sims.empty()) {
    b->V = x->V;
    b->dtype = x->c.dtype;
    b->m = x->tab.m;
    b->limiter = launch_limiter(x);
  } else {
    if (x->V != b->V || x->c.dtype != b->dtype || x->tab.m != b->m ||
        x->c.method != b->sims[0]->c.method || x->c.order != b->sims[0]->c.order)
      return bfail(b, "batch members must agree in physics, dtype and time integrator");
    if (launch_limiter(x) != b->limiter) b->limiter = 0;  // run-time limiter switch
  }
  b->sims.push_back(x);
  x->batch = b;
  b->dirty = true;
  return 0;
}

extern "C" int fv_batch_remove(fv_batch *b, fv_ctx *x) {
  if (!b || !x) return -1;
  auto it = std::find(b->sims.begin(), b->sims.end(), x);
  if (it == b->sims.end()) return bfail(b, "context is not a member of this batch");
  (void)hipStreamSynchronize(b->st);
  b->sims.erase(it);
  x->batch = nullptr;
  b->dirty = true;
  if (!b->sims.empty()) {
    b->limiter = launch_limiter(b->sims[0]);
    for (auto *y : b->sims)
      if (launch_limiter(y) != b->limiter) b->limiter = 0;
  }
  return 0;
}

extern "C" int fv_batch_run(fv_batch *b, int nsteps, int *taken) {
  if (!b || nsteps < 0) return -1;
  const int n = (int)b->sims.size();
  std::vector<int> rem(n, nsteps), tk(n, 0), fin(n, 0), want(n), done, halted;
  int rc = 0;
  if (n == 0) {
    b->last_ms = 0;
    return 0;
  }
  if (b->dirty && (rc = rebuild(b))) return rc;
  if (b->opt.autotune && b->opt.reuse_tuning && !b->tuned) {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tune_cache.find(tune_key(b));
    if (it != tune_cache.end()) {
      b->cand = it->second;
      b->tuned = 2;
    }
  }
  BCHK(hipEventRecord(b->e0, b->st));
  auto host_step = [&](int d) -> int {  // one Step on the host path
    int dn = 0;
    const int r = fv_step(b->sims[d], &dn);
    if (r) return bfail(b, "member " + std::to_string(d) + ": " + b->sims[d]->err, r);
    if (dn) {
      fin[d] = 1;
    } else {
      rem[d]--;
      tk[d]++;
      b->host_steps++;
    }
    return 0;
  };
  for (;;) {
    // members the device path cannot take yet step on the host path
    for (int d = 0; d < n; d++)
      while (rem[d] > 0 && !fin[d] && !fast_eligible(b->sims[d], true))
        if ((rc = host_step(d))) return rc;
    int kmax = 0, kmin = CTL_HIST;
    for (int d = 0; d < n; d++) {
      want[d] = (rem[d] > 0 && !fin[d]) ? std::min(rem[d], b->opt.max_graph_steps) : 0;
      if (want[d] > 0) {
        kmax = std::max(kmax, want[d]);
        kmin = std::min(kmin, want[d]);
        BCHK(hipStreamSynchronize(b->sims[d]->st));  // host-path work on the member's stream is done
      }
    }
    if (kmax == 0) break;
    const bool tuning = b->opt.autotune && !b->tuned && kmin >= 2;
    int K = kmax, c = b->cand;
    double ms = 0;
    if (tuning) {
      K = 2;
      c = b->trial % NCAND;
      for (int d = 0; d < n; d++) want[d] = std::min(want[d], 2);
    }
    if ((rc = batch_round(b, want, K, c, tuning, &ms, done, halted))) return rc;
    if (tuning) {
      const double per = ms / K;
      if (b->tune_ms[c] == 0 || per < b->tune_ms[c]) b->tune_ms[c] = per;
      if (++b->trial == NCAND * TUNE_REPS) {
        int best = 0;
        for (int q = 1; q < NCAND; q++)
          if (b->tune_ms[q] < b->tune_ms[best]) best = q;
        b->cand = best;
        b->tuned = 1;
        if (b->opt.reuse_tuning) {
          std::lock_guard<std::mutex> lk(tune_mu);
          tune_cache[tune_key(b)] = best;
        }
      }
    }
    for (int d = 0; d < n; d++) {
      if (want[d] <= 0) continue;
      rem[d] -= done[d];
      tk[d] += done[d];
      b->dev_steps += done[d];
      if (halted[d]) {
        b->halts++;
        if (rem[d] > 0 && (rc = host_step(d))) return rc;  // the step the device left
      }
    }
  }
  BCHK(hipEventRecord(b->e1, b->st));
  BCHK(hipEventSynchronize(b->e1));
  float ms = 0;
  BCHK(hipEventElapsedTime(&ms, b->e0, b->e1));
  b->last_ms = ms;
  if (taken)
    for (int d = 0; d < n; d++) taken[d] = tk[d];
  return 0;
}

extern "C" int fv_batch_get_info(const fv_batch *b, fv_batch_info *o) {
  if (!b || !o) return -1;
  o->members = (int)b->sims.size();
  o->tiling = b->cand;
  o->tuned = b->tuned;
  for (int c = 0; c < NCAND; c++) o->tune_ms[c] = b->tune_ms[c];
  o->device_steps = b->dev_steps;
  o->host_steps = b->host_steps;
  o->halts = b->halts;
  o->rounds = b->rounds;
  o->last_run_ms = b->last_ms;
  return 0;
}

extern "C" int fv_batch_destroy(fv_batch *b) {
  if (!b) return 0;
  if (b->st) (void)hipStreamSynchronize(b->st);
  for (auto *x : b->sims) x->batch = nullptr;
  drop_graphs(b);
  free_tables(b);
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
  return 0;
}

extern "C" const char *fv_batch_last_error(const fv_batch *b) { return b ? b->err.c_str() : "null batch"; }
