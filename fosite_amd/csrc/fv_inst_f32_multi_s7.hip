// fv_inst_f32_multi_s7.hip -- FP32 kernels of batch launches (many meshes per launch), tableaux of up to seven stages
#include "fv_launch.h"

namespace fv {
template void rhs_multi<float, MAXK>(const RhsEntry *, const unsigned *, int, int, int, int, int, hipStream_t);
}  // namespace fv
