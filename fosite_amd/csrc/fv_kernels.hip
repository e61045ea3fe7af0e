// fv_kernels.hip -- step-control and reduction kernels, and the launchers the
// host code calls: they pick the element type (state dtype) and hand over to
// the typed launchers instantiated in fv_inst_*.hip.
#include <cstring>
#include <mutex>

#include "fv_kernels.h"

namespace fv {

// Step control of a device-controlled batch (one block): embedded error of
// the step from the partials of the last stage, accept test, then CFL and
// data checks from the end-of-step k_rhs partials and the next time step.
__device__ __forceinline__ void control_block(Ctl *ctl, const double *part_c, const double *part_r, int nb,
                                              const double *part_v = nullptr, int nbv = 0) {
  __shared__ double rede[4], redm[4], redv[4];
  __shared__ int redi[4];
  if (ctl->halt) return;
  double e = -DBL_MAX, m = -DBL_MAX, v = -DBL_MAX;
  int f = 0;
#pragma unroll 4
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    e = fmax2(e, part_c[2 * b]);
    m = fmax2(m, part_r[2 * b]);
    f |= (int)part_r[2 * b + 1];
  }
  if (part_v) {
    for (int b = threadIdx.x; b < nbv; b += blockDim.x) v = fmax2(v, part_v[2 * b]);
  }
  // one barrier: wave results to LDS, thread 0 combines
  e = wave_max(e);
  m = wave_max(m);
  v = wave_max(v);
  f = wave_or(f);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    rede[w] = e;
    redm[w] = m;
    redv[w] = v;
    redi[w] = f;
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // from -DBL_MAX in wave order, as blk_max
    e = m = v = -DBL_MAX;
    f = 0;
    for (int k = 0; k < nw; k++) {
      e = fmax2(e, rede[k]);
      m = fmax2(m, redm[k]);
      v = fmax2(v, redv[k]);
      f |= redi[k];
    }
    control_accept(ctl, e);
    if (!ctl->halt) control_next(ctl, m, f, v);
  }
}
__global__ __launch_bounds__(256) void k_control(Ctl *ctl, const double *part_c, const double *part_r, int nb,
                                                 const double *part_v, int nbv) {
  control_block(ctl, part_c, part_r, nb, part_v, nbv);
}
// host-controlled steps: the boundary-flux accumulation of a step
__global__ __launch_bounds__(256) void k_bflux(BfluxArgs bf) {
  if (bf.ctl && bf.ctl->halt) return;
  bflux_accumulate(bf, bf.ctl ? bf.ctl->dt : bf.dt, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// UpdateGravity_single of one ComputeRHS (gravity_pointmass.f90:383-428; no
// switch-on phase), one workgroup: GetBoundaryFlux's density sum of the
// outflow boundary in index order, then the point-mass update.  A strictly
// ordered chain of FP64 adds: about 12 us for 1024 terms on MI355X.
__global__ __launch_bounds__(256) void k_accrete(AccreteArgs p) {
  if (p.ctl && p.ctl->halt) return;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const double sum = ordered_row_sum(p.bf + p.lo, p.hi - p.lo + 1, tid);
  if (tid < 64) accrete_update(p, sum, tid, 64);
}
// batches: the boundary-flux accumulation of the accreting members, bpm
// workgroups per member striding over its boundary elements
__global__ __launch_bounds__(256) void k_bflux_multi(const BfluxArgs *__restrict__ tab, int bpm) {
  const BfluxArgs &bf = tab[blockIdx.x / bpm];
  if (bf.ctl && bf.ctl->halt) return;
  const int r = blockIdx.x % bpm;
  bflux_accumulate(bf, bf.ctl ? bf.ctl->dt : bf.dt, r * blockDim.x + threadIdx.x, bpm * blockDim.x);
}
// batches: UpdateGravity of the accreting members, one workgroup each (the
// same ordered chain as k_accrete)
__global__ __launch_bounds__(256) void k_accrete_multi(const AccreteArgs *__restrict__ tab) {
  const AccreteArgs &p = tab[blockIdx.x];
  if (p.ctl && p.ctl->halt) return;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const double sum = ordered_row_sum(p.bf + p.lo, p.hi - p.lo + 1, tid);
  if (tid < 64) accrete_update(p, sum, tid, 64);
}
// one block per mesh of a batch
__global__ __launch_bounds__(256) void k_control_multi(const ControlEntry *__restrict__ tab) {
  const ControlEntry &E = tab[blockIdx.x];
  control_block(E.ctl, E.part_c, E.part_r, E.nb);
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

// FARGO type 1 background velocity (timedisc_base.f90:2243-2326, direction
// 2): w(i) = SUM(v(i,JMIN:JMAX)*wfactor(i)) / JNUM for every ring, the sum in
// index order.  A block covers WS_RINGS = 16 rings; waves 1-3 stage the next
// chunk of WS_ROWS rows in LDS (lane = ring + WS_RINGS * row group) while
// wave 0 adds the current one sequentially, one chain per ring.  Few rings
// per block and long chunks: more workgroups and fewer chunks, each chunk
// costing one memory latency.
constexpr int WS_RINGS = 16, WS_ROWS = 256, WS_RG = 64 / WS_RINGS;
__global__ __launch_bounds__(256) void k_wsum(Geom g, const double *__restrict__ v, double *ring, float *ringf,
                                              const Ctl *ctl) {
  __shared__ double buf[2][WS_ROWS][WS_RINGS];
  if (ctl && ctl->halt) return;  // a device-controlled step after the round halted
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int rl = lane % WS_RINGS, rg = lane / WS_RINGS;
  const int i = blockIdx.x * WS_RINGS + rl, jmin = GH, J = g.jnum;
  const int nchunk = (J + WS_ROWS - 1) / WS_ROWS;
  // waves 1..3 stage rows WS_RG*m + rg, m = wv-1, wv+2, ..., of chunk c into
  // buf[c & 1] (each row exactly once), every load issued before the first
  // store (one memory latency per chunk)
  constexpr int RPW = (WS_ROWS / WS_RG + 2) / 3;
  auto stage = [&](int c) {
    double t[RPW];
#pragma unroll
    for (int q = 0; q < RPW; q++) {
      const int r = WS_RG * (wv - 1 + 3 * q) + rg, j = c * WS_ROWS + r;
      t[q] = (r < WS_ROWS && i < g.NI && j < J) ? v[(long)(jmin + j) * g.P + i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < RPW; q++) {
      const int r = WS_RG * (wv - 1 + 3 * q) + rg;
      if (r < WS_ROWS) buf[c & 1][r][rl] = t[q];
    }
  };
  if (wv > 0) stage(0);
  __syncthreads();
  double acc = 0.0;
  const double f = i < g.NI ? g.ring[(long)R_WFAC * g.NI + i] : 0.0;
  for (int c = 0; c < nchunk; c++) {
    if (wv > 0) {
      if (c + 1 < nchunk) stage(c + 1);
    } else {
      const int rows = J - c * WS_ROWS < WS_ROWS ? J - c * WS_ROWS : WS_ROWS;
      // acc + v*f term by term in row order (lanes from WS_RINGS on repeat
      // the first WS_RINGS);
      // the LDS reads and products of 16 rows are issued ahead of their adds
      acc = ordered_sum<16>(acc, rows, [&](int r) { return buf[c & 1][r][rl] * f; });
    }
    __syncthreads();
  }
  if (wv == 0 && rg == 0 && i < g.NI) {
    const double w = acc / (double)J;
    ring[(long)R_W * g.NI + i] = w;
    if (ringf) ringf[(long)R_W * g.NI + i] = (float)w;
  }
}

// CalcTimestep on a stored pvar (timedisc_base.f90:818-909 with
// CalcWaveSpeeds_center, physics_eulerisotherm.f90:530-552) for the step
// after an output: WriteDataset has moved fosite's pvar to the real frame
// (v + w, physics_eulerisotherm.f90:1263-1303), and the wave speeds are those
// of that pvar.  pv is the comoving pvar of the last evaluation, all cells;
// grid-stride over the cells, one max per block into part[2b].
template <typename T, int V>
__global__ __launch_bounds__(256) void k_cfl_pvar(Geom g, const T *__restrict__ pv, int add_w, double *part) {
  __shared__ double red[4];
  double m = -DBL_MAX;
  const long n = (long)g.NI * g.NJ;
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long)gridDim.x * 256) {
    const int i = (int)(k % g.NI), j = (int)(k / g.NI);
    const unsigned ix = (unsigned)j * (unsigned)g.P + (unsigned)i;
    const T u = (pv + 1 * g.FS)[ix];
    T v = (pv + 2 * g.FS)[ix];
    if (add_w) v = v + rg<T>(g, R_W, i);
    const T cs = (V == 4) ? sound_speed((T)g.gamma, (pv + 0 * g.FS)[ix], (pv + 3 * g.FS)[ix]) : rg<T>(g, R_CS, i);
    m = fmax2(m, (double)cell_invdt(g, i, u, v, cs));
  }
  m = blk_max(m, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = m;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

void launch_cfl_pvar(const Geom &g, const void *pv, int add_w, double *part, int nblocks, int V, hipStream_t s) {
  if (g.dtype == 1) {
    if (V == 3) hipLaunchKernelGGL((k_cfl_pvar<float, 3>), dim3(nblocks), dim3(256), 0, s, g, (const float *)pv, add_w, part);
    else hipLaunchKernelGGL((k_cfl_pvar<float, 4>), dim3(nblocks), dim3(256), 0, s, g, (const float *)pv, add_w, part);
  } else {
    if (V == 3) hipLaunchKernelGGL((k_cfl_pvar<double, 3>), dim3(nblocks), dim3(256), 0, s, g, (const double *)pv, add_w, part);
    else hipLaunchKernelGGL((k_cfl_pvar<double, 4>), dim3(nblocks), dim3(256), 0, s, g, (const double *)pv, add_w, part);
  }
}

void launch_visc(const Geom &g, const StageArgs &a, const VisArgs &v, int V, hipStream_t s) {
  const dim3 grid((g.NI + VT_X - 1) / VT_X, (g.NJ + VT_Y - 1) / VT_Y), block(VT_X, VT_Y);
  const bool small = km_for(g.stages) == KM_SMALL;
  if (V == 3) {
    if (small) hipLaunchKernelGGL((k_visc<3, KM_SMALL>), grid, block, 0, s, g, a, v);
    else hipLaunchKernelGGL((k_visc<3, MAXK>), grid, block, 0, s, g, a, v);
  } else {
    if (small) hipLaunchKernelGGL((k_visc<4, KM_SMALL>), grid, block, 0, s, g, a, v);
    else hipLaunchKernelGGL((k_visc<4, MAXK>), grid, block, 0, s, g, a, v);
  }
}
void launch_visc_cfl(const Geom &g, const VisArgs &v, const double *pv, int add_w, double *part, int nblocks, int V,
                     hipStream_t s) {
  if (V == 3) hipLaunchKernelGGL((k_visc_cfl<3>), dim3(nblocks), dim3(256), 0, s, g, v, pv, add_w, part);
  else hipLaunchKernelGGL((k_visc_cfl<4>), dim3(nblocks), dim3(256), 0, s, g, v, pv, add_w, part);
}

// ------------------------------------------------- static reduction scratch
// The array's name is only meaningful in device code; host code uses the
// address hipGetSymbolAddress returns for each device (each device holds its
// own copy at its own address).
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

hipError_t scratch_address(int dev, double **out) {
  constexpr int MAXDEV = 64;
  static std::mutex mu;
  static double *addr[MAXDEV] = {};
  if (dev < 0 || dev >= MAXDEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    double *p = nullptr;
    e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    addr[dev] = p;
  }
  *out = addr[dev];
  return hipSuccess;
}

// -------------------------------------------------------------- launchers
// The host fills StageArgs / RemapArgs with untyped buffer addresses; the
// FP32 form reads the same fields as their float counterparts (identical
// layout: only the pointee type of the state pointers differs).
static_assert(sizeof(StageArgsT<float>) == sizeof(StageArgsT<double>), "StageArgs layout");
static_assert(sizeof(RemapArgsT<float>) == sizeof(RemapArgsT<double>), "RemapArgs layout");

int rhs_blocks(const Geom &g) { return ((g.inum + g.tile_x - 1) / g.tile_x) * ((g.jnum + TY_RHS - 1) / TY_RHS); }

void launch_rhs(const Geom &g, const StageArgs &a, int V, hipStream_t s) {
  if (g.dtype == 1) {
    StageArgsT<float> af;
    std::memcpy(&af, &a, sizeof(af));
    if (km_for(g.stages) == KM_SMALL) rhs_single<float, KM_SMALL>(g, af, V, s);
    else rhs_single<float, MAXK>(g, af, V, s);
  } else {
    if (km_for(g.stages) == KM_SMALL) rhs_single<double, KM_SMALL>(g, a, V, s);
    else rhs_single<double, MAXK>(g, a, V, s);
  }
}
void launch_remap(const Geom &g, const RemapArgs &a, int V, hipStream_t s) {
  if (g.dtype == 1) {
    RemapArgsT<float> af;
    std::memcpy(&af, &a, sizeof(af));
    remap_single<float>(g, af, V, s);
  } else {
    remap_single<double>(g, a, V, s);
  }
}
void launch_frame(const Geom &g, const void *src, void *dst, int sign, int V, hipStream_t s, const Ctl *ctl,
                  const void *w2) {
  if (g.dtype == 1)
    frame_single<float>(g, src, dst, sign, V, ctl, w2, s);
  else
    frame_single<double>(g, src, dst, sign, V, ctl, w2, s);
}
void launch_control(Ctl *ctl, const double *part_c, const double *part_r, int nb, hipStream_t s,
                    const double *part_v, int nbv) {
  hipLaunchKernelGGL(k_control, dim3(1), dim3(256), 0, s, ctl, part_c, part_r, nb, part_v, nbv);
}
int visc_blocks(const Geom &g) { return ((g.NI + VT_X - 1) / VT_X) * ((g.NJ + VT_Y - 1) / VT_Y); }
void launch_rhs_multi(const RhsEntry *tab, const unsigned *tiles, int ntiles, int V, int dtype, int limiter,
                      int tx, int stages, int mode, hipStream_t s) {
  const bool small = km_for(stages) == KM_SMALL;
  if (dtype == 1) {
    if (small) rhs_multi<float, KM_SMALL>(tab, tiles, ntiles, V, limiter, tx, mode, s);
    else rhs_multi<float, MAXK>(tab, tiles, ntiles, V, limiter, tx, mode, s);
  } else {
    if (small) rhs_multi<double, KM_SMALL>(tab, tiles, ntiles, V, limiter, tx, mode, s);
    else rhs_multi<double, MAXK>(tab, tiles, ntiles, V, limiter, tx, mode, s);
  }
}
void launch_remap_multi(const RemapEntry *tab, const unsigned *tiles, int ntiles, int V, int dtype, hipStream_t s) {
  if (dtype == 1)
    remap_multi<float>(tab, tiles, ntiles, V, s);
  else
    remap_multi<double>(tab, tiles, ntiles, V, s);
}
void launch_control_multi(const ControlEntry *tab, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_control_multi, dim3(n), dim3(256), 0, s, tab);
}
void launch_bflux_multi(const BfluxArgs *tab, int n, int bpm, hipStream_t s) {
  hipLaunchKernelGGL(k_bflux_multi, dim3(n * bpm), dim3(256), 0, s, tab, bpm);
}
void launch_accrete_multi(const AccreteArgs *tab, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_accrete_multi, dim3(n), dim3(64, 4), 0, s, tab);
}
void launch_wsum(const Geom &g, const double *v, double *ring, float *ringf, hipStream_t s, const Ctl *ctl) {
  hipLaunchKernelGGL(k_wsum, dim3((g.NI + WS_RINGS - 1) / WS_RINGS), dim3(256), 0, s, g, v, ring, ringf, ctl);
}
void launch_bflux(const BfluxArgs &bf, hipStream_t s) {
  const int n = bf.V * 2 * (bf.inum + bf.jnum);
  hipLaunchKernelGGL(k_bflux, dim3((n + 255) / 256), dim3(256), 0, s, bf);
}
void launch_accrete(const AccreteArgs &p, hipStream_t s) {
  hipLaunchKernelGGL(k_accrete, dim3(1), dim3(64, 4), 0, s, p);
}
void launch_reduce(const double *part, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, part, n, out);
}

}  // namespace fv
