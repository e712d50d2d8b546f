// transfer_pass.hip — the fused multigrid transfer of a 2D grid set (multigrid.h:454-657).
//
// One launch performs, for every line along the freed axis f of a receiving region:
//   slice  = source edge values of the line (first and last nCAP nodes along f, zero
//            elsewhere), each times inv_mask(node_f) * mask_correction(node_other) as one
//            FP64 factor (maskSlice multigrid.h:532-569, bit-exact host tables)
//   slice  = forward unnormalised FFT along f (transf_slice multigrid.h:232-270)
//   psi    = psi * (pending X factor of the receiver) + slice    (add multigrid.h:571-585)
//   then optionally writes the receiver's own edge nodes for the next group's transfer and
//   multiplies by the receiver's region mask (maskRegion grid.h:582-597).
#include "axis_pass_impl.hpp"

namespace qsf {

template <typename Real, int N, bool ROW>
__global__ void __launch_bounds__((Geo<Real, N, ROW, 0>::THREADS))
    transfer_kernel(TransferParams p) {
  using T = typename cvec<Real>::T;
  using PL = PlanOf<N, 0>;
  constexpr int E = PL::E, TT = PL::T;
  constexpr int W = Geo<Real, N, ROW, 0>::W;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);

  const int tid = threadIdx.x;
  const int t = ROW ? (tid % TT) : (tid / W);
  const int w = ROW ? (tid / TT) : (tid % W);
  // XCD-aware tile order (blocks go round-robin over the 8 XCDs): consecutive column tiles
  // share HBM lines, so they run on one XCD and its L2 serves the neighbours' partial lines
  int64_t b = blockIdx.x;
  const int64_t nb = gridDim.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  // 2D grid (n0, n1): f = axis; row mode transforms along axis 1 (o = row i0), column mode
  // along axis 0 (c = column i1)
  int64_t o, c;
  bool valid;
  if (ROW) {
    o = b * W + w;
    c = 0;
    valid = o < p.n0;
  } else {
    const int64_t cpt = (p.n1 + W - 1) / W;
    o = b / cpt;
    c = (b % cpt) * W + w;
    valid = c < p.n1;
  }
  const int64_t other = ROW ? o : c;        // index along the other axis
  const int64_t n_o = ROW ? p.n0 : p.n1;
  const int64_t base = ROW ? o * N : c;
  const int64_t step = ROW ? 1 : p.n1;
  const int i0 = ROW ? (int)o : 0, i1 = ROW ? 0 : (int)c;

  // slice weight of the other axis (mask_correction of group-1 receivers)
  double wo = 1.0;
  if (p.corr_tab) {
    const int64_t nd = other < n_o / 2 ? -other : other - n_o + 1;
    int64_t node = p.nCAP + nd;
    node = node < 0 ? 0 : node;
    wo = p.corr_tab[node];
  }
  const T* ein = reinterpret_cast<const T*>(p.edge_in);
  T v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int64_t i = t + TT * q;   // node along the freed axis
    const bool edge = i < p.nCAP || i >= N - p.nCAP;
    T x{Real(0), Real(0)};
    if (valid && edge) {
      const int64_t e = i < p.nCAP ? i : i - (N - 2 * p.nCAP);
      const T s = p.edge_in_layout == 0 ? ein[e * n_o + other] : ein[other * 2 * p.nCAP + e];
      const int64_t nd = i < N / 2 ? -i : i - N + 1;
      int64_t node = p.nCAP + nd;
      node = node < 0 ? 0 : node;
      const double wt = p.corr_tab ? p.inv_tab[node] * wo : p.inv_tab[node];
      x = T{(Real)((double)s.x * wt), (Real)((double)s.y * wt)};
    }
    v[q] = x;
  }

  const T* roots = reinterpret_cast<const T*>(p.roots);
  Twiddles<N, 0, T> tw;
  tw.load(roots, t);
  if constexpr (ROW) {
    RowAddr<T> a{w * row_stride<T, N>()};
    fft<N, 0, -1>(v, t, lds, a, roots, tw);
  } else {
    ColAddr<T, W> a{w};
    fft<N, 0, -1>(v, t, lds, a, roots, tw);
  }

  T* dst = reinterpret_cast<T*>(p.dst);
  XEval<N> xa;
  if (p.has_xa) xa.init(p.xa, ROW ? 1 : 0, i0, i1, 0);
  T* eout = reinterpret_cast<T*>(p.edge_out);
  int64_t mo = 0;
  if (p.mask_tab) {
    const int64_t nd = other < n_o / 2 ? -other : other - n_o + 1;
    int64_t node = p.nCAP + nd;
    node = node < 0 ? 0 : node;
    mo = node * p.mask_stride_other;
  }
  const bool other_edge = other < p.nCAP || other >= n_o - p.nCAP;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int64_t i = t + TT * q;
    if (!valid) continue;
    T d = dst[base + i * step];
    if (p.has_xa) d = capply(d, xa.at((int)i));
    T r = T{d.x + v[q].x, d.y + v[q].y};
    if (eout) {
      // this receiver's edge nodes along the axis the next group frees
      if (p.edge_out_layout == 0) {            // edge rows (axis 0) -> [2nCAP][n1]
        const int64_t r0 = ROW ? other : i, c1 = ROW ? i : other;
        if (r0 < p.nCAP || r0 >= p.n0 - p.nCAP) {
          const int64_t e = r0 < p.nCAP ? r0 : r0 - (p.n0 - 2 * p.nCAP);
          eout[e * p.n1 + c1] = r;
        }
      } else {                                  // edge columns (axis 1) -> [n0][2nCAP]
        const int64_t r0 = ROW ? other : i, c1 = ROW ? i : other;
        if (c1 < p.nCAP || c1 >= p.n1 - p.nCAP) {
          const int64_t e = c1 < p.nCAP ? c1 : c1 - (p.n1 - 2 * p.nCAP);
          eout[r0 * 2 * p.nCAP + e] = r;
        }
      }
    }
    if (p.mask_tab) {
      int64_t off = mo;
      if (p.mask_stride_axis) {
        const int64_t nd = i < N / 2 ? -i : i - N + 1;
        int64_t node = p.nCAP + nd;
        node = node < 0 ? 0 : node;
        off += node * p.mask_stride_axis;
      }
      const double mk = p.mask_tab[off];
      r = T{(Real)((double)r.x * mk), (Real)((double)r.y * mk)};
    }
    dst[base + i * step] = r;
  }
  (void)other_edge;
}

template <typename Real, int N, bool ROW>
static int launch_tr(const TransferParams& p, hipStream_t s) {
  using G = Geo<Real, N, ROW, 0>;
  int64_t tiles = ROW ? (p.n0 + G::W - 1) / G::W : (p.n1 + G::W - 1) / G::W;
  size_t lds = (size_t)G::lds_cap(G::W);
  static const bool attr_set = [&] {   // once per kernel, also under concurrent callers
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&transfer_kernel<Real, N, ROW>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return true;
  }();
  (void)attr_set;
  hipLaunchKernelGGL((transfer_kernel<Real, N, ROW>), dim3((unsigned)tiles), dim3(G::THREADS), lds, s, p);
  return (int)tiles;
}

template <typename Real, bool ROW>
static int tr_dispatch(int64_t n, const TransferParams& p, hipStream_t s) {
  switch (n) {
#define QSF_CASE(NN) case NN: return launch_tr<Real, NN, ROW>(p, s);
    QSF_SIZES(QSF_CASE)
#undef QSF_CASE
    default: return -1;
  }
}

int launch_transfer(int precision, const TransferParams& p, hipStream_t s) {
  const int64_t n = p.axis == 1 ? p.n1 : p.n0;
  const bool row = p.axis == 1;
  if (precision == 0) return row ? tr_dispatch<double, true>(n, p, s) : tr_dispatch<double, false>(n, p, s);
  return row ? tr_dispatch<float, true>(n, p, s) : tr_dispatch<float, false>(n, p, s);
}

}  // namespace qsf
