// axis_pass_f64_row.hip — instantiations of the fused axis pass (double, row mode).
#include "axis_pass_impl.hpp"

namespace qsf {
int dispatch_f64_row(int64_t n, int v, const PassParams& p, hipStream_t s) { return dispatch_n<double, true>(n, v, p, s); }
int64_t tiles_f64_row(int64_t n, int v, const PassParams& p) { return tiles_n<double, true>(n, v, p); }
}  // namespace qsf
