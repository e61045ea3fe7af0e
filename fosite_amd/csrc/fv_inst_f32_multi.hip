// fv_inst_f32_multi.hip -- FP32 kernels of batch launches (many meshes per launch)
#include "fv_launch.h"

namespace fv {
template void rhs_multi<float>(const RhsEntry *, const unsigned *, int, int, int, int, hipStream_t);
template void remap_multi<float>(const RemapEntry *, const unsigned *, int, int, hipStream_t);
}  // namespace fv
