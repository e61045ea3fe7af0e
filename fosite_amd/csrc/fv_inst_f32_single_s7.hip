// fv_inst_f32_single_s7.hip -- FP32 kernels of single-mesh launches, tableaux of up to seven stages
#include "fv_launch.h"

namespace fv {
template void rhs_single<float, MAXK>(const Geom &, const StageArgsT<float> &, int, hipStream_t);
}  // namespace fv
