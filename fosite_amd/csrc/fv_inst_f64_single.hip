// fv_inst_f64_single.hip -- FP64 kernels of single-mesh launches
#include "fv_launch.h"

namespace fv {
template void rhs_single<double>(const Geom &, const StageArgsT<double> &, int, hipStream_t);
template void remap_single<double>(const Geom &, const RemapArgsT<double> &, int, hipStream_t);
template void frame_single<double>(const Geom &, const void *, void *, int, int, hipStream_t);
}  // namespace fv
