// axis_pass.hip — public launchers of the fused axis pass. The kernels live in
// axis_pass_impl.hpp and are instantiated per (precision, mode, length group) by
// axis_pass_inst.hip; this file only dispatches on the runtime axis length.
#include <hip/hip_runtime.h>

#include "../../include/qsf_types.h"
#include "axis_pass_impl.hpp"

namespace qsf {

#define QSF_EXTERN_ONE(R, ROW, NN)                                                   \
  extern template int launch_size<R, ROW, NN>(int, const PassParams&, hipStream_t); \
  extern template int64_t tiles_size<R, ROW, NN>(int, const PassParams&);
#define QSF_EXTERN_D_ROW(NN) QSF_EXTERN_ONE(double, true, NN)
#define QSF_EXTERN_D_COL(NN) QSF_EXTERN_ONE(double, false, NN)
#define QSF_EXTERN_F_ROW(NN) QSF_EXTERN_ONE(float, true, NN)
#define QSF_EXTERN_F_COL(NN) QSF_EXTERN_ONE(float, false, NN)
QSF_SIZES(QSF_EXTERN_D_ROW)
QSF_SIZES(QSF_EXTERN_D_COL)
QSF_SIZES(QSF_EXTERN_F_ROW)
QSF_SIZES(QSF_EXTERN_F_COL)
#define QSF_EXTERN_S32(ROW, NN) extern template int launch_size<double, ROW, NN, true>(int, const PassParams&, hipStream_t);
#define QSF_EXTERN_S_ROW(NN) QSF_EXTERN_S32(true, NN)
#define QSF_EXTERN_S_COL(NN) QSF_EXTERN_S32(false, NN)
QSF_SIZES(QSF_EXTERN_S_ROW)
QSF_SIZES(QSF_EXTERN_S_COL)

template <typename Real, bool ROW, bool S32 = false>
static int dispatch(int64_t n, int v, const PassParams& p, hipStream_t s) {
  switch (n) {
#define QSF_CASE(NN) case NN: return launch_size<Real, ROW, NN, S32>(v, p, s);
    QSF_SIZES(QSF_CASE)
#undef QSF_CASE
    default: return -1;
  }
}

template <typename Real, bool ROW>
static int64_t tiles(int64_t n, int v, const PassParams& p) {
  switch (n) {
#define QSF_CASE(NN) case NN: return tiles_size<Real, ROW, NN>(v, p);
    QSF_SIZES(QSF_CASE)
#undef QSF_CASE
    default: return -1;
  }
}

bool axis_size_supported(int64_t n) {
  switch (n) {
#define QSF_CASE(NN) case NN:
    QSF_SIZES(QSF_CASE)
#undef QSF_CASE
      return true;
    default: return false;
  }
}

int axis_variants(int64_t n) { return tuned_length(n) ? kVariants : 1; }

int launch_axis_pass(int precision, const PassParams& p, hipStream_t s, int v) {
  const int64_t n = p.n[p.axis];
  const bool row = (p.inner == 1);
  if (v < 0 || v >= axis_variants(n)) v = 0;
  if (precision == QSF_FP64) return row ? dispatch<double, true>(n, v, p, s) : dispatch<double, false>(n, v, p, s);
  if (precision == QSF_FP32_STATE)
    return row ? dispatch<double, true, true>(n, v, p, s) : dispatch<double, false, true>(n, v, p, s);
  return row ? dispatch<float, true>(n, v, p, s) : dispatch<float, false>(n, v, p, s);
}

int64_t axis_pass_tiles(int precision, const PassParams& p, int v) {
  const int64_t n = p.n[p.axis];
  const bool row = (p.inner == 1);
  if (v < 0 || v >= axis_variants(n)) v = 0;
  // FP64 arithmetic over an FP32 state tiles like the FP64 kernels
  if (precision == QSF_FP64 || precision == QSF_FP32_STATE)
    return row ? tiles<double, true>(n, v, p) : tiles<double, false>(n, v, p);
  return row ? tiles<float, true>(n, v, p) : tiles<float, false>(n, v, p);
}

template <typename Real>
__global__ void gather_x_kernel(GatherXArgs a) {
  using T = typename cvec<Real>::T;
  const T* psi = reinterpret_cast<const T*>(a.psi);
  T* out = reinterpret_cast<T*>(a.out);
  const int axis = a.dim - 1;
  const int64_t len = a.n[axis];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.nvals;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = a.row[k];
    const int p = a.pos[k];
    T v = psi[r * len + p];
    if (a.has_xa) {
      // the line's coordinates (axis `axis` varies along the row)
      const int i0 = a.dim == 3 ? (int)(r / a.n[1]) : (a.dim == 2 ? (int)r : 0);
      const int i1 = a.dim == 3 ? (int)(r % a.n[1]) : 0;
      XEval<1> xa;
      xa.init(a.xa, axis, i0, i1, 0);
      v = capply(v, xa.at(p));
    }
    out[k] = v;
  }
}

void launch_gather_x(int precision, const GatherXArgs& a, hipStream_t s) {
  if (a.nvals <= 0) return;
  const int blocks = (int)((a.nvals + 255) / 256 < 65536 ? (a.nvals + 255) / 256 : 65536);
  if (precision == QSF_FP64) hipLaunchKernelGGL(gather_x_kernel<double>, dim3(blocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(gather_x_kernel<float>, dim3(blocks), dim3(256), 0, s, a);
}

}  // namespace qsf
