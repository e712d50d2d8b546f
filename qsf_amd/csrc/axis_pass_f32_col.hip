// axis_pass_f32_col.hip — instantiations of the fused axis pass (float, column mode).
#include "axis_pass_impl.hpp"

namespace qsf {
int dispatch_f32_col(int64_t n, int v, const PassParams& p, hipStream_t s) { return dispatch_n<float, false>(n, v, p, s); }
int64_t tiles_f32_col(int64_t n, int v, const PassParams& p) { return tiles_n<float, false>(n, v, p); }
}  // namespace qsf
