// fv_launch.h -- launch templates of the kernels in fv_kernels.h: the
// compile-time variants (limiter, tile width, field count) each launch picks.
#pragma once
#include "fv_kernels.h"

namespace fv {

template <typename T, int V, int LIM, int TX, int KM>
static void rhs_lim(const Geom &g, const StageArgsT<T> &a, hipStream_t s) {
  const dim3 grid(xcd_grid((g.inum + TX - 1) / TX, (g.jnum + TY_RHS - 1) / TY_RHS)), block(TX, TY_RHS);
  switch (rhs_mode(a)) {
    case M_STAGE: hipLaunchKernelGGL((k_rhs<T, V, LIM, TX, KM, M_STAGE>), grid, block, 0, s, g, a); break;
    case M_LAST: hipLaunchKernelGGL((k_rhs<T, V, LIM, TX, KM, M_LAST>), grid, block, 0, s, g, a); break;
    default: hipLaunchKernelGGL((k_rhs<T, V, LIM, TX, KM, M_END>), grid, block, 0, s, g, a); break;
  }
}
template <typename T, int V, int TX, int KM>
static void rhs_v(const Geom &g, const StageArgsT<T> &a, hipStream_t s) {
  // compile-time limiters for the common choices, run-time switch otherwise
  // (and for PRIMITIVE reconstruction, which only that instantiation carries)
  if (g.recon_prim)
    rhs_lim<T, V, 0, TX, KM>(g, a, s);
  else if (g.limiter == VANLEER)
    rhs_lim<T, V, VANLEER, TX, KM>(g, a, s);
  else if (g.limiter == MINMOD)
    rhs_lim<T, V, MINMOD, TX, KM>(g, a, s);
  else
    rhs_lim<T, V, 0, TX, KM>(g, a, s);
}
template <typename T, int KM>
void rhs_single(const Geom &g, const StageArgsT<T> &a, int V, hipStream_t s) {
  if (V == 3) {
    if (g.tile_x == 32) rhs_v<T, 3, 32, KM>(g, a, s);
    else rhs_v<T, 3, 64, KM>(g, a, s);
  } else {
    if (g.tile_x == 32) rhs_v<T, 4, 32, KM>(g, a, s);
    else rhs_v<T, 4, 64, KM>(g, a, s);
  }
}
template <typename T>
void remap_single(const Geom &g, const RemapArgsT<T> &a, int V, hipStream_t s) {
  const dim3 grid(xcd_grid((g.inum + 63) / 64, remap_rows_blocks(g.jnum))), block(64, 4);
  if (V == 3)
    hipLaunchKernelGGL((k_remap<T, 3>), grid, block, 0, s, g, a);
  else
    hipLaunchKernelGGL((k_remap<T, 4>), grid, block, 0, s, g, a);
}
template <typename T>
void frame_single(const Geom &g, const void *src, void *dst, int sign, int V, const Ctl *ctl, const void *w2,
                  hipStream_t s) {
  dim3 grid((g.NI + 63) / 64, (g.NJ + 3) / 4), block(64, 4);
  const T *sp = static_cast<const T *>(src);
  T *dp = static_cast<T *>(dst);
  if (V == 3)
    hipLaunchKernelGGL((k_frame<T, 3>), grid, block, 0, s, g, sp, dp, sign, ctl, static_cast<const T *>(w2));
  else
    hipLaunchKernelGGL((k_frame<T, 4>), grid, block, 0, s, g, sp, dp, sign, ctl, static_cast<const T *>(w2));
}

// batches: van Leer compiled in, every other limiter (and mixed limiters) by
// the run-time switch
template <typename T, int V, int LIM, int TX, int KM>
static void rhs_multi_lim(const RhsEntry *tab, const unsigned *tiles, int n, int mode, hipStream_t s) {
  const dim3 block(TX, TY_RHS);
  switch (mode) {
    case M_STAGE: hipLaunchKernelGGL((k_rhs_multi<T, V, LIM, TX, KM, M_STAGE>), dim3(n), block, 0, s, tab, tiles); break;
    case M_LAST: hipLaunchKernelGGL((k_rhs_multi<T, V, LIM, TX, KM, M_LAST>), dim3(n), block, 0, s, tab, tiles); break;
    default: hipLaunchKernelGGL((k_rhs_multi<T, V, LIM, TX, KM, M_END>), dim3(n), block, 0, s, tab, tiles); break;
  }
}
template <typename T, int V, int TX, int KM>
static void rhs_multi_v(const RhsEntry *tab, const unsigned *tiles, int n, int limiter, int mode, hipStream_t s) {
  if (limiter == VANLEER)
    rhs_multi_lim<T, V, VANLEER, TX, KM>(tab, tiles, n, mode, s);
  else
    rhs_multi_lim<T, V, 0, TX, KM>(tab, tiles, n, mode, s);
}
template <typename T, int KM>
void rhs_multi(const RhsEntry *tab, const unsigned *tiles, int n, int V, int limiter, int tx, int mode, hipStream_t s) {
  if (V == 3) {
    if (tx == 32) rhs_multi_v<T, 3, 32, KM>(tab, tiles, n, limiter, mode, s);
    else rhs_multi_v<T, 3, 64, KM>(tab, tiles, n, limiter, mode, s);
  } else {
    if (tx == 32) rhs_multi_v<T, 4, 32, KM>(tab, tiles, n, limiter, mode, s);
    else rhs_multi_v<T, 4, 64, KM>(tab, tiles, n, limiter, mode, s);
  }
}
template <typename T>
void remap_multi(const RemapEntry *tab, const unsigned *tiles, int n, int V, hipStream_t s) {
  const dim3 block(64, 4);
  if (V == 3)
    hipLaunchKernelGGL((k_remap_multi<T, 3>), dim3(n), block, 0, s, tab, tiles);
  else
    hipLaunchKernelGGL((k_remap_multi<T, 4>), dim3(n), block, 0, s, tab, tiles);
}

}  // namespace fv
