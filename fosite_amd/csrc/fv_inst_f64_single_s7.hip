// fv_inst_f64_single_s7.hip -- FP64 kernels of single-mesh launches, tableaux of up to seven stages
#include "fv_launch.h"

namespace fv {
template void rhs_single<double, MAXK>(const Geom &, const StageArgsT<double> &, int, hipStream_t);
}  // namespace fv
