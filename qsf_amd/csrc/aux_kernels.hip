// aux_kernels.hip — small kernels around the axis passes: per-substep phase tables, pair
// tables, FFT roots, deterministic reductions, the surface-flux reduction and conversions.
#include <hip/hip_runtime.h>

#include "../../include/qsf_types.h"
#include "fft_core.hpp"
#include "qsf_launch.hpp"

namespace qsf {

struct AxisTableBatch {
  AxisTableArgs a[24];
};

template <typename Real>
__global__ void axis_tables_kernel(AxisTableBatch batch) {
  using T = typename cvec<Real>::T;
  const AxisTableArgs& a = batch.a[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double h;
  if (a.kind == 0) {  // X: potential (per-axis part) + x·F  (hamiltonian.h:44-79)
    const double x = a.xmin + i * a.dx;
    const double ne = a.use_ne ? a.NE / sqrt(x * x + a.Nsoft) : 0.0;
    h = ne + x * a.C;
  } else {            // P: ½p² + p·A
    const double p = a.dp * (double)(i < a.pn2 ? i : i - a.pn);
    h = (0.5 * p * p) + (p * a.C);
  }
  const double v = a.delta * h;
  double re, im;
  if (a.mode == QSF_MODE_IM) { re = exp(-v); im = 0.0; }
  else { sincos(-v, &im, &re); }
  if (a.kind == 0 && a.apply_abs && a.absorber != QSF_ABS_NONE) {
    const int64_t half = a.n / 2;
    const int64_t nd = i < half ? -i : i - a.n + 1;   // neg_dist_from_edge grid.h:181-185
    const int64_t node = a.nCAP + nd;
    if (node > 0) {
      double f;
      if (a.absorber == QSF_ABS_CAP) {
        const double s4 = (double)(node * (node * (node * node)));
        f = exp(-a.delta * s4 * a.eta);                  // absorbers.h:40-46, one axis
      } else {
        // cos(π/2 · node/nCAP) as the sine of the distance from the edge (never negative)
        const double c = sin(0.5 * M_PI * (double)(a.nCAP - node) / (double)a.nCAP);
        f = pow(pow(c, 0.125), a.delta / a.tau);         // cos^(1/8) per tau of time
      }
      re *= f;
      im *= f;
    }
  }
  re *= a.norm;
  im *= a.norm;
  reinterpret_cast<T*>(a.out)[i] = T{(Real)re, (Real)im};
}

void launch_axis_tables(int precision, const AxisTableArgs* a, int count, hipStream_t s) {
  for (int off = 0; off < count; off += 24) {
    AxisTableBatch b;
    int k = count - off < 24 ? count - off : 24;
    int64_t nmax = 0;
    for (int i = 0; i < k; ++i) {
      b.a[i] = a[off + i];
      nmax = a[off + i].n > nmax ? a[off + i].n : nmax;
    }
    dim3 grid((unsigned)((nmax + 255) / 256), (unsigned)k);
    // phase tables are FP64 in both precisions: the FP32 step applies them with one FP64
    // product and a single rounding (a table rounded to FP32 once would bias every step)
    (void)precision;
    hipLaunchKernelGGL(axis_tables_kernel<double>, grid, dim3(256), 0, s, b);
  }
}

template <typename Real>
__global__ void pair_table_kernel(PairTableArgs a) {
  using T = typename cvec<Real>::T;
  const int64_t total = a.kind == 1 ? a.n : a.n * a.n;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  double ee;
  if (a.kind == 1) {  // |i-j| = k (Eberly: depends on x-y only)
    const double d = (double)k * a.dx;
    ee = a.EE / sqrt(d * d + a.Esoft);
  } else {
    const int64_t i = k / a.n, j = k % a.n;
    const double x = a.xmin + i * a.dx, y = a.xmin + j * a.dx;
    if (a.model == QSF_POT_ECKHARDT_SACHA) ee = a.EE / sqrt((x - y) * (x - y) + x * y + a.Esoft);
    else ee = a.EE / sqrt((x - y) * (x - y) + a.Esoft);
  }
  const double v = a.delta * ee;
  double re, im;
  if (a.mode == QSF_MODE_IM) { re = exp(-v); im = 0.0; }
  else { sincos(-v, &im, &re); }
  reinterpret_cast<T*>(a.out)[k] = T{(Real)re, (Real)im};
}

template <typename Real>
__global__ void radial_x_kernel(RadialArgs a) {
  using T = typename cvec<Real>::T;
  const int64_t m = a.n[0] * a.n[1] * a.n[2];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const int64_t id[3] = {e / (a.n[1] * a.n[2]), (e / a.n[2]) % a.n[1], e % a.n[2]};
  // squares of the bounded coordinates folded from the right, as the parameter-pack fold of
  // potential.h:43 does for the coordinates in order
  double s = 0.0;
  bool first = true;
  for (int d = 2; d >= 0; --d) {
    if (!(a.bounded_mask & (1 << d))) continue;
    const double x = a.xmin[d] + id[d] * a.dx[d];
    s = first ? x * x : (x * x) + s;
    first = false;
  }
  const double v = a.delta * (a.NE / sqrt(s + a.Nsoft));
  double re, im;
  if (a.mode == QSF_MODE_IM) { re = exp(-v); im = 0.0; }
  else { sincos(-v, &im, &re); }
  T* p = reinterpret_cast<T*>(a.psi);
  const T z = p[e];
  const double zr = (double)z.x, zi = (double)z.y;
  p[e] = T{(Real)(zr * re - zi * im), (Real)(zr * im + zi * re)};
}

void launch_radial_x(const RadialArgs& a, hipStream_t s) {
  const int64_t m = a.n[0] * a.n[1] * a.n[2];
  const dim3 grid((unsigned)((m + 255) / 256));
  if (a.precision == QSF_FP64) hipLaunchKernelGGL(radial_x_kernel<double>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(radial_x_kernel<float>, grid, dim3(256), 0, s, a);
}

void launch_pair_table(int precision, const PairTableArgs& a, hipStream_t s) {
  const int64_t total = a.kind == 1 ? a.n : a.n * a.n;
  dim3 grid((unsigned)((total + 255) / 256));
  (void)precision;  // FP64 in both precisions (see launch_axis_tables)
  hipLaunchKernelGGL(pair_table_kernel<double>, grid, dim3(256), 0, s, a);
}

template <typename Real>
__global__ void roots_kernel(typename cvec<Real>::T* out, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s, c;
  sincospi(-2.0 * (double)j / (double)n, &s, &c);
  const Real hc = (Real)c, hs = (Real)s;
  if constexpr (sizeof(Real) == 8) {
    out[j] = typename cvec<Real>::T{hc, hs};
  } else {
    // FP32: {hi, lo} per root, adjacent (hi + lo carries the FP64 root; one 16-byte load)
    out[2 * j] = typename cvec<Real>::T{hc, hs};
    out[2 * j + 1] = typename cvec<Real>::T{(Real)(c - (double)hc), (Real)(s - (double)hs)};
  }
}

void launch_roots(int precision, void* out, int64_t n, hipStream_t s) {
  dim3 grid((unsigned)((n + 255) / 256));
  if (precision == QSF_FP64)
    hipLaunchKernelGGL(roots_kernel<double>, grid, dim3(256), 0, s, reinterpret_cast<double2*>(out), n);
  else
    hipLaunchKernelGGL(roots_kernel<float>, grid, dim3(256), 0, s, reinterpret_cast<float2*>(out), n);
}

// One block, fixed summation order: thread k sums blocks k, k+256, ... then a tree.
__global__ void reduce_partials_kernel(const double* partials, int64_t count, int width, int ncols, double* out) {
  __shared__ double red[256];
  for (int c = 0; c < ncols; ++c) {
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < count; k += 256) acc += partials[k * width + c];
    red[threadIdx.x] = acc;
    block_sync();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      block_sync();
    }
    if (threadIdx.x == 0) out[c] = red[0];
    block_sync();
  }
}

// First level: block b sums the contiguous chunk [b*chunk, (b+1)*chunk) of the partials.
__global__ void reduce_chunks_kernel(const double* partials, int64_t count, int width, int ncols, int64_t chunk,
                                     double* out) {
  __shared__ double red[256];
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < count ? lo + chunk : count;
  for (int c = 0; c < ncols; ++c) {
    double acc = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += 256) acc += partials[k * width + c];
    red[threadIdx.x] = acc;
    block_sync();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      block_sync();
    }
    if (threadIdx.x == 0) out[blockIdx.x * ncols + c] = red[0];
    block_sync();
  }
}

void launch_reduce_partials(const double* partials, int64_t count, int width, double* out, double* scratch,
                            hipStream_t s, int ncols) {
  if (ncols < 0) ncols = width;
  if (count <= 4096 || !scratch) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partials, count, width, ncols, out);
    return;
  }
  // two fixed-order levels: 256 chunk sums into `scratch` (256*ncols doubles), then one block
  const int64_t chunk = (count + 255) / 256;
  hipLaunchKernelGGL(reduce_chunks_kernel, dim3(256), dim3(256), 0, s, partials, count, width, ncols, chunk, scratch);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, scratch, (int64_t)256, ncols, ncols, out);
}

template <typename Real>
__global__ void flux_kernel(FluxArgs a) {
  using T = typename cvec<Real>::T;
  const T* g = reinterpret_cast<const T*>(a.gathered);
  const int D = a.D, nb = a.nb, stride = 1 + 2 * D;
  double acc[16];
  for (int b = 0; b < nb; ++b) acc[b] = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.npts;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* q = a.pts + p * stride;
    const T c = g[q[0]];
    const double cr = c.x, ci = c.y, nrm = cr * cr + ci * ci;
    double j[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
      const int32_t sp = q[1 + d], sm = q[1 + D + d];
      if (sp >= 0 && sm >= 0) {
        // inv_2dx * Im(conj(ψ_c) (ψ_+ - ψ_-))  (grid.h:467-481)
        const double dr = (double)g[sp].x - (double)g[sm].x, di = (double)g[sp].y - (double)g[sm].y;
        j[d] = a.inv_2dx[d] * (cr * di - ci * dr);
      }
      j[d] += a.vgC[d] * nrm;  // velocityGaugeCorrection (propagator.h:249-257)
    }
    const double* bv = a.bvec + p * nb * D;
    for (int b = 0; b < nb; ++b) {
      double s = 0.0;
      for (int d = 0; d < D; ++d) s += j[d] * bv[b * D + d];
      acc[b] += s;
    }
  }
  __shared__ double red[256];
  for (int b = 0; b < nb; ++b) {
    red[threadIdx.x] = acc[b];
    block_sync();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      block_sync();
    }
    if (threadIdx.x == 0) {
      if (gridDim.x == 1) {
        // one block holds the whole sum: finish here (what flux_finish_kernel computes for a
        // single partial, bit for bit) and save that launch
        a.out[b] = red[0] * a.scale;
        if (a.accum) a.accum[b] += a.out[b] * a.accum_weight;
      } else {
        a.partials[blockIdx.x * nb + b] = red[0];
      }
    }
    block_sync();
  }
}

__global__ void flux_finish_kernel(const double* partials, int64_t count, int nb, double scale, double* out,
                                   double* accum, double w) {
  // one family at a time: the block loads the per-block sums and folds them in a fixed tree
  __shared__ double red[256];
  for (int b = 0; b < nb; ++b) {
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < count; k += blockDim.x) acc += partials[k * nb + b];
    red[threadIdx.x] = acc;
    block_sync();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      block_sync();
    }
    if (threadIdx.x == 0) {
      out[b] = red[0] * scale;
      if (accum) accum[b] += out[b] * w;
    }
    block_sync();
  }
}

void launch_flux(const FluxArgs& a, hipStream_t s) {
  int64_t want = (a.npts + 255) / 256;
  const int blocks = (int)(want < 1 ? 1 : (want > 1024 ? 1024 : want));
  if (a.precision == QSF_FP64) hipLaunchKernelGGL(flux_kernel<double>, dim3(blocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(flux_kernel<float>, dim3(blocks), dim3(256), 0, s, a);
  if (blocks > 1)
    hipLaunchKernelGGL(flux_finish_kernel, dim3(1), dim3(256), 0, s, a.partials, (int64_t)blocks, a.nb, a.scale,
                       a.out, a.accum, a.accum_weight);
}

template <typename Real>
__global__ void scale_kernel(typename cvec<Real>::T* psi, int64_t m, double s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    auto v = psi[i];
    psi[i] = typename cvec<Real>::T{(Real)(v.x * s), (Real)(v.y * s)};
  }
}

void launch_scale(int precision, void* psi, int64_t m, double s, hipStream_t st) {
  const int blocks = (int)((m + 255) / 256 < 8192 ? (m + 255) / 256 : 8192);
  if (precision == QSF_FP64)
    hipLaunchKernelGGL(scale_kernel<double>, dim3(blocks), dim3(256), 0, st, reinterpret_cast<double2*>(psi), m, s);
  else
    hipLaunchKernelGGL(scale_kernel<float>, dim3(blocks), dim3(256), 0, st, reinterpret_cast<float2*>(psi), m, s);
}

template <typename A, typename B>
__global__ void convert_kernel(const A* src, B* dst, int64_t m) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    A v = src[i];
    dst[i] = B{(decltype(dst->x))v.x, (decltype(dst->x))v.y};
  }
}

void launch_copy_convert(int from_prec, const void* src, int to_prec, void* dst, int64_t m, hipStream_t s) {
  const int blocks = (int)((m + 255) / 256 < 8192 ? (m + 255) / 256 : 8192);
  if (from_prec == QSF_FP64 && to_prec == QSF_FP32_MIXED)
    hipLaunchKernelGGL((convert_kernel<double2, float2>), dim3(blocks), dim3(256), 0, s,
                       reinterpret_cast<const double2*>(src), reinterpret_cast<float2*>(dst), m);
  else if (from_prec == QSF_FP32_MIXED && to_prec == QSF_FP64)
    hipLaunchKernelGGL((convert_kernel<float2, double2>), dim3(blocks), dim3(256), 0, s,
                       reinterpret_cast<const float2*>(src), reinterpret_cast<double2*>(dst), m);
}

}  // namespace qsf

namespace qsf {

template <typename Real>
__global__ void slice_2d_kernel(SliceArgs a) {
  using T = typename cvec<Real>::T;
  const int64_t m = a.n0 * a.n1;
  const T* src = reinterpret_cast<const T*>(a.src);
  T* out = reinterpret_cast<T*>(a.slice);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = k / a.n1, i1 = k % a.n1;
    const int64_t i_f = a.dirFree == 0 ? i0 : i1, n_f = a.dirFree == 0 ? a.n0 : a.n1;
    const int64_t i_o = a.dirFree == 0 ? i1 : i0, n_o = a.dirFree == 0 ? a.n1 : a.n0;
    // edge rows/columns of the free axis (getBox multigrid.h:454-482)
    const bool edge = i_f < a.nCAP || i_f >= n_f - a.nCAP;
    T v{Real(0), Real(0)};
    if (edge) {
      const int64_t nd = i_f < n_f / 2 ? -i_f : i_f - n_f + 1;   // neg_dist_from_edge
      int64_t node = a.nCAP + nd;
      node = node < 0 ? 0 : node;
      double w = a.inv_tab[node];
      if (a.corr_tab) {
        const int64_t ndo = i_o < n_o / 2 ? -i_o : i_o - n_o + 1;
        int64_t nodo = a.nCAP + ndo;
        nodo = nodo < 0 ? 0 : nodo;
        w = w * a.corr_tab[nodo];                                  // maskSlice :547-548
      }
      const T x = src[k];
      v = T{(Real)((double)x.x * w), (Real)((double)x.y * w)};
    }
    out[k] = v;
  }
}

void launch_slice_2d(int precision, const SliceArgs& a, hipStream_t s) {
  const int64_t m = a.n0 * a.n1;
  const int blocks = (int)((m + 255) / 256 < 16384 ? (m + 255) / 256 : 16384);
  if (precision == QSF_FP64) hipLaunchKernelGGL(slice_2d_kernel<double>, dim3(blocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(slice_2d_kernel<float>, dim3(blocks), dim3(256), 0, s, a);
}

// getBox + maskSlice (multigrid.h:454-569) of a 3D receiver, every box at once: the freed
// axis' edge nodes of the source, times inv_mask of the freed-axis node and, for group-1
// receivers, mask_correction (:519-530) with its folds in the reference's order.
template <typename Real>
__global__ void slice_3d_kernel(SliceArgs a) {
  using T = typename cvec<Real>::T;
  const int64_t m = a.n0 * a.n1 * a.n2;
  const int64_t n[3] = {a.n0, a.n1, a.n2};
  const T* src = reinterpret_cast<const T*>(a.src);
  T* out = reinterpret_cast<T*>(a.slice);
  const int f = a.dirFree;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t idx[3] = {k / (a.n1 * a.n2), (k / a.n2) % a.n1, k % a.n2};
    const int64_t i_f = idx[f], n_f = n[f];
    T v{Real(0), Real(0)};
    if (i_f < a.nCAP || i_f >= n_f - a.nCAP) {
      int64_t node[3];
      for (int d = 0; d < 3; ++d) {
        const int64_t nd = idx[d] < n[d] / 2 ? -idx[d] : idx[d] - n[d] + 1;   // neg_dist_from_edge
        const int64_t c = a.nCAP + nd;
        node[d] = c < 0 ? 0 : c;
      }
      double w = a.inv_tab[node[f]];
      if (a.m1_tab) {
        double sum = 0.0;   // t0 + (t1 + (t2 + 0))
        for (int d = 2; d >= 0; --d) sum = (d == f ? 0.0 : a.m1_tab[node[d]]) + sum;
        double prod = 1.0;  // u0 * (u1 * u2)
        for (int d = 2; d >= 0; --d) {
          const double u = d == f ? 1.0 : a.m1_tab[node[d]];
          prod = d == 2 ? u : u * prod;
        }
        const double onesixth = 1.0 / 6.0, onehalf = 0.5;
        w = w * (onehalf * sum + onesixth * (1 - prod));
      }
      const T x = src[k];
      v = T{(Real)((double)x.x * w), (Real)((double)x.y * w)};
    }
    out[k] = v;
  }
}

void launch_slice_3d(int precision, const SliceArgs& a, hipStream_t s) {
  const int64_t m = a.n0 * a.n1 * a.n2;
  const int blocks = (int)((m + 255) / 256 < 16384 ? (m + 255) / 256 : 16384);
  if (precision == QSF_FP64) hipLaunchKernelGGL(slice_3d_kernel<double>, dim3(blocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(slice_3d_kernel<float>, dim3(blocks), dim3(256), 0, s, a);
}

template <typename T>
__global__ void add_kernel(T* dst, const T* src, int64_t m) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    T a = dst[k], b = src[k];
    dst[k] = T{a.x + b.x, a.y + b.y};
  }
}

// operation<AntiSymmetrize<...>> (wf/grid.h:611-723), in place, X representation. Every node
// has exactly one writer: the thread of its sorted pair / triple, or, for a node with repeated
// indices, its own thread.
// AntiSymmetrize<2_D> on a 2D grid (:700-716, psi_total is the unmodified state):
// psi(i,j) - psi(j,i).
template <typename T>
__global__ void antisym2_kernel(T* psi, int64_t n) {
  const int64_t m = n * n;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q / n, j = q % n;
    if (j > i) continue;
    const int64_t b = j * n + i;
    const T x = psi[q], y = psi[b];
    psi[q] = T{x.x - y.x, x.y - y.y};
    if (b != q) psi[b] = T{y.x - x.x, y.y - x.y};
  }
}
// AntiSymmetrize<2_D> on a 3D grid, over the last two axes (:686-699): (i,j,k) -= (i,k,j) for
// k < j, (i,k,j) = -(i,j,k), (i,j,j) = 0.
template <typename T>
__global__ void antisym2_partial_kernel(T* psi, int64_t n0, int64_t n) {
  const int64_t nn = n * n, m = n0 * nn;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q / nn, r = q % nn, j = r / n, k = r % n;
    if (k > j) continue;
    if (k == j) {
      psi[q] = T{0, 0};
      continue;
    }
    const int64_t b = (i * n + k) * n + j;
    const T x = psi[q], y = psi[b];
    const T v{x.x - y.x, x.y - y.y};
    psi[q] = v;
    psi[b] = T{-v.x, -v.y};
  }
}
// AntiSymmetrize<3_D>, one rank (:646-677): for i > j > k,
// (i,j,k) += (j,k,i) + (k,i,j) - (j,i,k) - (i,k,j) - (k,j,i) in that order, the cyclic
// permutations get the value, the transpositions its negative; nodes with two equal indices 0.
template <typename T>
__global__ void antisym3_kernel(T* psi, int64_t n) {
  const int64_t nn = n * n, m = n * nn;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q / nn, r = q % nn, j = r / n, k = r % n;
    if (i == j || j == k || i == k) {
      psi[q] = T{0, 0};
      continue;
    }
    if (!(i > j && j > k)) continue;
    const int64_t r2 = (j * n + k) * n + i, r3 = (k * n + i) * n + j;
    const int64_t s1 = (j * n + i) * n + k, s2 = (i * n + k) * n + j, s3 = (k * n + j) * n + i;
    const T a = psi[q], b = psi[r2], c = psi[r3], d = psi[s1], e = psi[s2], f = psi[s3];
    T t{b.x + c.x, b.y + c.y};
    t = T{t.x - d.x, t.y - d.y};
    t = T{t.x - e.x, t.y - e.y};
    t = T{t.x - f.x, t.y - f.y};
    const T v{a.x + t.x, a.y + t.y}, w{-v.x, -v.y};
    psi[q] = v;
    psi[r2] = v;
    psi[r3] = v;
    psi[s1] = w;
    psi[s2] = w;
    psi[s3] = w;
  }
}

// kind 2: AntiSymmetrize<2_D> (2D whole grid, 3D last two axes); kind 3: AntiSymmetrize<3_D>.
void launch_antisym(int precision, void* psi, int dim, const int64_t* n, int kind, hipStream_t s) {
  int64_t m = 1;
  for (int a = 0; a < dim; ++a) m *= n[a];
  const int blocks = (int)((m + 255) / 256 < 16384 ? (m + 255) / 256 : 16384);
  const bool f64 = precision == QSF_FP64;
  if (kind == 2 && dim == 2) {
    if (f64) hipLaunchKernelGGL(antisym2_kernel<double2>, dim3(blocks), dim3(256), 0, s, (double2*)psi, n[0]);
    else hipLaunchKernelGGL(antisym2_kernel<float2>, dim3(blocks), dim3(256), 0, s, (float2*)psi, n[0]);
  } else if (kind == 2 && dim == 3) {
    if (f64) hipLaunchKernelGGL(antisym2_partial_kernel<double2>, dim3(blocks), dim3(256), 0, s, (double2*)psi, n[0], n[1]);
    else hipLaunchKernelGGL(antisym2_partial_kernel<float2>, dim3(blocks), dim3(256), 0, s, (float2*)psi, n[0], n[1]);
  } else if (kind == 3 && dim == 3) {
    if (f64) hipLaunchKernelGGL(antisym3_kernel<double2>, dim3(blocks), dim3(256), 0, s, (double2*)psi, n[0]);
    else hipLaunchKernelGGL(antisym3_kernel<float2>, dim3(blocks), dim3(256), 0, s, (float2*)psi, n[0]);
  }
}

// croossOut (wf/grid.h:890-907) with multiIonizationFilter<2_D, 2_D> (:862-888): every node is
// scaled by 1 where exactly two coordinates lie beyond `from_center` nodes from the centre, else
// by a cos ramp of the integer distance picked by the filter's folds. `tab[k]` =
// 0.5 * (1 + cos(pi * k / ramp)), k = 0..ramp, is built on the host in FP64.
template <typename T>
__global__ void cross_out_kernel(T* psi, int dim, int64_t n0, int64_t n1, int64_t n2, int64_t from_boundary,
                                 int64_t ramp, const double* tab) {
  const int64_t n[3] = {n0, n1, n2};
  const int64_t m = n0 * n1 * n2;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
    int64_t idx[3] = {q / (n1 * n2), (q / n2) % n1, q % n2};
    int64_t nodes[3];
    int count = 0;
    for (int a = 0; a < dim; ++a) {
      const int64_t i = idx[3 - dim + a], na = n[3 - dim + a];
      const int64_t nd = i < na / 2 ? -i : i - na + 1;      // neg_dist_from_edge
      nodes[a] = nd + from_boundary;
      count += nodes[a] > 0 ? 1 : 0;
    }
    if (count == 2) continue;                                // from <= count <= to
    int64_t mv = nodes[dim - 1];
    const bool take_max = count == 0 && dim != 2;            // from != DIM
    for (int a = 0; a < dim; ++a)
      if (take_max ? nodes[a] > mv : nodes[a] < mv) mv = nodes[a];
    mv = mv < 0 ? -mv : mv;
    if (mv > ramp) mv = ramp;
    const double w = tab[mv];
    const T x = psi[q];
    psi[q] = T{(decltype(x.x))((double)x.x * w), (decltype(x.y))((double)x.y * w)};
  }
}

// from_boundary = n[0] / 2 - nodesFromCenter (n2[0] of the grid's first axis, :868)
void launch_cross_out(int precision, void* psi, int dim, const int64_t* n, int64_t from_boundary, int64_t ramp,
                      const double* tab, hipStream_t s) {
  const int64_t n0 = dim == 3 ? n[0] : 1, n1 = dim >= 2 ? n[dim - 2] : 1, n2 = n[dim - 1];
  const int64_t m = n0 * n1 * n2;
  const int blocks = (int)((m + 255) / 256 < 16384 ? (m + 255) / 256 : 16384);
  if (precision == QSF_FP64)
    hipLaunchKernelGGL(cross_out_kernel<double2>, dim3(blocks), dim3(256), 0, s, (double2*)psi, dim, n0, n1, n2,
                       from_boundary, ramp, tab);
  else
    hipLaunchKernelGGL(cross_out_kernel<float2>, dim3(blocks), dim3(256), 0, s, (float2*)psi, dim, n0, n1, n2,
                       from_boundary, ramp, tab);
}

// orthogonalizeWith (wf/grid.h:909-946): <phi|psi> = sum conj(phi) psi in FP64 (block partials
// in a fixed order, then one thread over the partials), then psi -= s * phi with s the overlap
// times dV. phi is complex FP64 in both precisions.
template <typename T>
__global__ void overlap_partials_kernel(const double2* phi, const T* psi, int64_t m, double* partials) {
  __shared__ double sr[256], si[256];
  double ar = 0.0, ai = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = phi[q];
    const double bx = psi[q].x, by = psi[q].y;
    ar += a.x * bx + a.y * by;
    ai += a.x * by - a.y * bx;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  block_sync();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    block_sync();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sr[0];
    partials[2 * blockIdx.x + 1] = si[0];
  }
}
__global__ void overlap_finish_kernel(const double* partials, int nb, double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double r = 0.0, i = 0.0;
  for (int b = 0; b < nb; ++b) {
    r += partials[2 * b];
    i += partials[2 * b + 1];
  }
  out[0] = r;
  out[1] = i;
}
template <typename T>
__global__ void sub_scaled_kernel(T* psi, const double2* phi, int64_t m, double sr, double si) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = phi[q];
    const double px = (double)psi[q].x - (sr * a.x - si * a.y);
    const double py = (double)psi[q].y - (sr * a.y + si * a.x);
    psi[q] = T{(decltype(psi[q].x))px, (decltype(psi[q].y))py};
  }
}

// partials: 2 * 1024 doubles; out: 2 doubles
void launch_overlap(int precision, const void* phi, const void* psi, int64_t m, double* partials, double* out,
                    hipStream_t s) {
  const int nb = (int)((m + 255) / 256 < 1024 ? (m + 255) / 256 : 1024);
  if (precision == QSF_FP64)
    hipLaunchKernelGGL(overlap_partials_kernel<double2>, dim3(nb), dim3(256), 0, s, (const double2*)phi,
                       (const double2*)psi, m, partials);
  else
    hipLaunchKernelGGL(overlap_partials_kernel<float2>, dim3(nb), dim3(256), 0, s, (const double2*)phi,
                       (const float2*)psi, m, partials);
  hipLaunchKernelGGL(overlap_finish_kernel, dim3(1), dim3(64), 0, s, (const double*)partials, nb, out);
}
void launch_sub_scaled(int precision, void* psi, const void* phi, int64_t m, double sr, double si, hipStream_t s) {
  const int nb = (int)((m + 255) / 256 < 16384 ? (m + 255) / 256 : 16384);
  if (precision == QSF_FP64)
    hipLaunchKernelGGL(sub_scaled_kernel<double2>, dim3(nb), dim3(256), 0, s, (double2*)psi, (const double2*)phi, m, sr, si);
  else
    hipLaunchKernelGGL(sub_scaled_kernel<float2>, dim3(nb), dim3(256), 0, s, (float2*)psi, (const double2*)phi, m, sr, si);
}

void launch_add(int precision, void* dst, const void* src, int64_t m, hipStream_t s) {
  const int blocks = (int)((m + 255) / 256 < 16384 ? (m + 255) / 256 : 16384);
  if (precision == QSF_FP64)
    hipLaunchKernelGGL(add_kernel<double2>, dim3(blocks), dim3(256), 0, s, reinterpret_cast<double2*>(dst),
                       reinterpret_cast<const double2*>(src), m);
  else
    hipLaunchKernelGGL(add_kernel<float2>, dim3(blocks), dim3(256), 0, s, reinterpret_cast<float2*>(dst),
                       reinterpret_cast<const float2*>(src), m);
}

}  // namespace qsf

namespace qsf {

template <typename Real>
__global__ void batch_ptables_kernel(const BatchTableArgs* args, const int64_t* step_ctr) {
  using T = typename cvec<Real>::T;
  const BatchTableArgs& a = args[blockIdx.y];
  const int64_t k = *step_ctr;
  if (k >= a.total_steps) return;
  const int ax = blockIdx.z;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n[ax]) return;
  const double C = a.coupling[3 * k + a.storage_axis[ax]];
  const double p = a.dp[ax] * (double)(i < a.pn2[ax] ? i : i - a.pn[ax]);
  const double v = a.delta * ((0.5 * p * p) + (p * C));
  double sn, cs;
  sincos(-v, &sn, &cs);
  const double nrm = 1.0 / (double)a.n[ax];
  reinterpret_cast<T*>(a.out[ax])[i] = T{(Real)(cs * nrm), (Real)(sn * nrm)};
}

void launch_batch_ptables(int precision, const BatchTableArgs* dev_args, int nsolves, int64_t nmax,
                          const int64_t* step_ctr, hipStream_t s) {
  dim3 grid((unsigned)((nmax + 255) / 256), (unsigned)nsolves, 2);
  (void)precision;  // FP64 in both precisions (see launch_axis_tables)
  hipLaunchKernelGGL(batch_ptables_kernel<double>, grid, dim3(256), 0, s, dev_args, step_ctr);
}

// One block per solve: norm from the pass partials and the flux families from the gathered
// boundary points, each a fixed-order sum (thread-strided, then a tree). The blocks are few (one
// per solve of a size group) and the point loop is a chain of dependent loads, so the block is
// as wide as it goes.
constexpr int kObsThreads = 1024;
template <typename Real>
__global__ void __launch_bounds__(kObsThreads)
    batch_obs_kernel(const BatchObsArgs* args, BatchFluxGeo geo, const int64_t* step_ctr) {
  using T = typename cvec<Real>::T;
  const BatchObsArgs& a = args[blockIdx.x];
  const int64_t k = *step_ctr;            // steps completed at the observed boundary
  if (k > a.total_steps || k < 1) return;
  __shared__ double red[kObsThreads];
  double* row = a.rec + k * (1 + geo.nb);
  double acc = 0.0;
  for (int64_t b = threadIdx.x; b < a.tiles; b += kObsThreads) acc += a.partials[3 * b];
  red[threadIdx.x] = acc;
  block_sync();
  for (int s = kObsThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    block_sync();
  }
  if (threadIdx.x == 0) row[0] = red[0] * a.dV;
  block_sync();
  const double* C = a.coupling + 3 * (k - 1);   // A after the step's P substep
  const T* g = reinterpret_cast<const T*>(a.gathered);
  const int D = geo.D, nb = geo.nb, stride = 1 + 2 * D;
  // one sweep over the points: each point's stencil is read once and its current added to every
  // family's sum (each family keeps its own per-thread order over the points)
  double fam[16];
  for (int f = 0; f < nb; ++f) fam[f] = 0.0;
  for (int64_t pidx = threadIdx.x; pidx < geo.npts; pidx += kObsThreads) {
    const int32_t* q = geo.pts + pidx * stride;
    const T c = g[q[0]];
    const double cr = c.x, ci = c.y, nrm = cr * cr + ci * ci;
    double jd[3];
    for (int d = 0; d < D; ++d) {
      jd[d] = 0.0;
      const int32_t sp = q[1 + d], sm = q[1 + D + d];
      if (sp >= 0 && sm >= 0) {
        const double dr = (double)g[sp].x - (double)g[sm].x, di = (double)g[sp].y - (double)g[sm].y;
        jd[d] = geo.inv_2dx[d] * (cr * di - ci * dr);
      }
      jd[d] += C[d] * nrm;
    }
    for (int f = 0; f < nb; ++f) {
      double j = 0.0;
      for (int d = 0; d < D; ++d) {
        const double bv = geo.bvec[(pidx * nb + f) * D + d];
        if (bv == 0.0) continue;
        j += jd[d] * bv;
      }
      fam[f] += j;
    }
  }
  for (int f = 0; f < nb; ++f) {
    red[threadIdx.x] = fam[f];
    block_sync();
    for (int t = kObsThreads / 2; t > 0; t >>= 1) {
      if ((int)threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
      block_sync();
    }
    if (threadIdx.x == 0) {
      row[1 + f] = red[0] * geo.scale;
      a.flux_int[f] += row[1 + f] * geo.dt;
    }
    block_sync();
  }
}

void launch_batch_obs(const BatchObsArgs* dev_args, int nsolves, const BatchFluxGeo& geo,
                      const int64_t* step_ctr, hipStream_t s) {
  if (geo.precision == QSF_FP64)
    hipLaunchKernelGGL(batch_obs_kernel<double>, dim3(nsolves), dim3(kObsThreads), 0, s, dev_args, geo, step_ctr);
  else
    hipLaunchKernelGGL(batch_obs_kernel<float>, dim3(nsolves), dim3(kObsThreads), 0, s, dev_args, geo, step_ctr);
}

__global__ void step_increment_kernel(int64_t* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c += 1;
}

void launch_step_increment(int64_t* step_ctr, hipStream_t s) {
  hipLaunchKernelGGL(step_increment_kernel, dim3(1), dim3(64), 0, s, step_ctr);
}

}  // namespace qsf
