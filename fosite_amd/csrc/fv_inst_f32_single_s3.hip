// fv_inst_f32_single_s3.hip -- FP32 kernels of single-mesh launches, tableaux of up to three stages
#include "fv_launch.h"

namespace fv {
template void rhs_single<float, KM_SMALL>(const Geom &, const StageArgsT<float> &, int, hipStream_t);
template void remap_single<float>(const Geom &, const RemapArgsT<float> &, int, hipStream_t);
template void frame_single<float>(const Geom &, const void *, void *, int, int, const Ctl *, const void *,
                                   hipStream_t);
}  // namespace fv
