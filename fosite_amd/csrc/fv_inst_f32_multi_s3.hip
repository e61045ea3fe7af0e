// fv_inst_f32_multi_s3.hip -- FP32 kernels of batch launches (many meshes per launch), tableaux of up to three stages
#include "fv_launch.h"

namespace fv {
template void rhs_multi<float, KM_SMALL>(const RhsEntry *, const unsigned *, int, int, int, int, int, hipStream_t);
template void remap_multi<float>(const RemapEntry *, const unsigned *, int, int, hipStream_t);
}  // namespace fv
