// axis_pass_f64_col.hip — instantiations of the fused axis pass (double, column mode).
#include "axis_pass_impl.hpp"

namespace qsf {
int dispatch_f64_col(int64_t n, int v, const PassParams& p, hipStream_t s) { return dispatch_n<double, false>(n, v, p, s); }
int64_t tiles_f64_col(int64_t n, int v, const PassParams& p) { return tiles_n<double, false>(n, v, p); }
}  // namespace qsf
