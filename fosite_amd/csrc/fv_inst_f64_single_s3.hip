// fv_inst_f64_single_s3.hip -- FP64 kernels of single-mesh launches, tableaux of up to three stages
#include "fv_launch.h"

namespace fv {
template void rhs_single<double, KM_SMALL>(const Geom &, const StageArgsT<double> &, int, hipStream_t);
template void remap_single<double>(const Geom &, const RemapArgsT<double> &, int, hipStream_t);
template void frame_single<double>(const Geom &, const void *, void *, int, int, const Ctl *, const void *,
                                   hipStream_t);
}  // namespace fv
