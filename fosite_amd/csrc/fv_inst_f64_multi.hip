// fv_inst_f64_multi.hip -- FP64 kernels of batch launches (many meshes per launch)
#include "fv_launch.h"

namespace fv {
template void rhs_multi<double>(const RhsEntry *, const unsigned *, int, int, int, int, hipStream_t);
template void remap_multi<double>(const RemapEntry *, const unsigned *, int, int, hipStream_t);
}  // namespace fv
