// fv_inst_f64_multi_s7.hip -- FP64 kernels of batch launches (many meshes per launch), tableaux of up to seven stages
#include "fv_launch.h"

namespace fv {
template void rhs_multi<double, MAXK>(const RhsEntry *, const unsigned *, int, int, int, int, int, hipStream_t);
}  // namespace fv
