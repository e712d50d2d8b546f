// axis_pass.hip — public launchers of the fused axis pass (kernels in axis_pass_impl.hpp,
// instantiated per precision and mode in axis_pass_{f64,f32}_{row,col}.hip).
#include <hip/hip_runtime.h>

#include "qsf_launch.hpp"

namespace qsf {

int dispatch_f64_row(int64_t n, int v, const PassParams& p, hipStream_t s);
int dispatch_f64_col(int64_t n, int v, const PassParams& p, hipStream_t s);
int dispatch_f32_row(int64_t n, int v, const PassParams& p, hipStream_t s);
int dispatch_f32_col(int64_t n, int v, const PassParams& p, hipStream_t s);
int64_t tiles_f64_row(int64_t n, int v, const PassParams& p);
int64_t tiles_f64_col(int64_t n, int v, const PassParams& p);
int64_t tiles_f32_row(int64_t n, int v, const PassParams& p);
int64_t tiles_f32_col(int64_t n, int v, const PassParams& p);

bool axis_size_supported(int64_t n) {
  switch (n) {
    case 8: case 16: case 32: case 64: case 128: case 256: case 512: case 1024: case 2048:
    case 4096: case 12: case 24: case 48: case 96: case 192: case 384: case 768: case 1536:
    case 3072:
      return true;
    default: return false;
  }
}

int axis_variants(int64_t n) { return (n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096 ||
                                        n == 768 || n == 1536) ? 6 : 1; }

int launch_axis_pass(int precision, const PassParams& p, hipStream_t s, int v) {
  const int64_t n = p.n[p.axis];
  const bool row = (p.inner == 1);
  if (precision == 0) return row ? dispatch_f64_row(n, v, p, s) : dispatch_f64_col(n, v, p, s);
  return row ? dispatch_f32_row(n, v, p, s) : dispatch_f32_col(n, v, p, s);
}

int64_t axis_pass_tiles(int precision, const PassParams& p, int v) {
  const int64_t n = p.n[p.axis];
  const bool row = (p.inner == 1);
  if (precision == 0) return row ? tiles_f64_row(n, v, p) : tiles_f64_col(n, v, p);
  return row ? tiles_f32_row(n, v, p) : tiles_f32_col(n, v, p);
}

}  // namespace qsf
