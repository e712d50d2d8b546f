// axis_pass_f32_row.hip — instantiations of the fused axis pass (float, row mode).
#include "axis_pass_impl.hpp"

namespace qsf {
int dispatch_f32_row(int64_t n, int v, const PassParams& p, hipStream_t s) { return dispatch_n<float, true>(n, v, p, s); }
int64_t tiles_f32_row(int64_t n, int v, const PassParams& p) { return tiles_n<float, true>(n, v, p); }
}  // namespace qsf
