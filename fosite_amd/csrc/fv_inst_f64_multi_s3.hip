// fv_inst_f64_multi_s3.hip -- FP64 kernels of batch launches (many meshes per launch), tableaux of up to three stages
#include "fv_launch.h"

namespace fv {
template void rhs_multi<double, KM_SMALL>(const RhsEntry *, const unsigned *, int, int, int, int, int, hipStream_t);
template void remap_multi<double>(const RemapEntry *, const unsigned *, int, int, hipStream_t);
}  // namespace fv
