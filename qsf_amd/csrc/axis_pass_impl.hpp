#pragma once
// axis_pass_impl.hpp — the fused axis pass of the split-operator step (gfx950).
//
// One launch sweeps the whole grid once along one axis. A block owns W transforms of length
// N: W consecutive rows (row mode: the contiguous last axis) or W adjacent columns (column
// mode: a strided axis, so every global access is a W*sizeof(complex) contiguous segment).
// Per transform it can do, in registers: inverse FFT -> X phase (A half) -> observables of the
// step-boundary state -> X phase (B half) -> forward FFT -> P phase -> inverse FFT. This is
// the reference's chainEvolve sequence (propagator.h:311-332: fourier + evolve_ grid.h:394-418)
// with the X halves of consecutive substeps meeting inside one pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <utility>

#include "fft_core.hpp"
#include "pass_params.hpp"
#include "qsf_launch.hpp"

namespace qsf {

template <typename Real> struct Tr {
  using T = typename cvec<Real>::T;
};

// Stored state type of a pass: the arithmetic type's complex vector, or (S32, FP64 arithmetic
// over an FP32 state) float2 — read and widened at the load, rounded once at each store.
template <typename Real, bool S32> using StateOf = std::conditional_t<S32, float2, typename Tr<Real>::T>;
template <typename S, typename T> __device__ __forceinline__ S to_state(T v) {
  if constexpr (std::is_same_v<S, T>) return v;
  else return S{(float)v.x, (float)v.y};
}
template <typename T, typename S> __device__ __forceinline__ T from_state(S v) {
  if constexpr (std::is_same_v<S, T>) return v;
  else return T{(double)v.x, (double)v.y};
}

// Row-mode exchange layout. Stockham stage 0 writes element j*R + r from lane j (stride R);
// unpadded, 16 lanes of one write pass then hit two 4-byte-bank groups (FP64, R = 8: 128 B
// apart) — an 8-way conflict. One pad element per 256 B of line (16 complex FP64 or 32 complex
// FP32, one full sweep of the 64 banks) spreads those lanes over distinct banks.
// WAVE: the block's transforms are each held by threads of one wave (rows of at most 64
// threads), so their exchanges need no block barrier.
template <typename T, bool WAVE = false> struct RowAddr {
  static constexpr int SH = sizeof(T) == 16 ? 4 : 5;
  static constexpr bool kWave = WAVE;
  static constexpr bool kSplitOk = false;
  int base;
  __device__ __forceinline__ int operator()(int p) const { return base + p + (p >> SH); }
};
// elements of LDS one padded line occupies
template <typename T, int N> constexpr int row_stride() { return N + (N >> RowAddr<T>::SH); }

// Kernel variant V (0..5): ALT = V / 3 selects the elements-per-thread plan, V % 3 scales the
// tile width W (x1, x1/2, x2) within the LDS and block-size limits.
constexpr int kVariants = 6;
constexpr int var_alt(int v) { return v / 3; }
constexpr int var_wshift(int v) { return v % 3 == 0 ? 0 : (v % 3 == 1 ? -1 : 1); }

// Compile-time feature set of a kernel. A kernel's register allocation covers every path it
// contains, so passes that only transform and multiply by the P factor (the column passes, the
// middle passes, plain transforms) run a build without the X stage, the observation, the flux
// gather, the region mask and the multigrid edge writes: far fewer registers, more waves.
// Row passes that carry the X halves and an observation but no region mask, multigrid edges or
// kinetic-energy copy (the 3D step's last-axis pass) run a middle tier without those paths.
enum PassFeat : int { FEAT_FULL = 0, FEAT_LEAN = 1, FEAT_X = 2 };
constexpr uint32_t kLeanFlags = F_INV1 | F_FWD | F_P | F_INV2 | F_SCALE | F_NOSTORE;
constexpr uint32_t kXFlags = kLeanFlags | F_XA | F_XB | F_OBS;
inline int feat_of(const PassParams& p, bool row) {
  if ((p.flags & ~kLeanFlags) == 0 && !p.er_out) return FEAT_LEAN;
  if (row && (p.flags & ~kXFlags) == 0 && !p.er_out && !((p.flags & F_OBS) && p.obs.copy_out)) return FEAT_X;
  return FEAT_FULL;
}

// Block geometry per (precision, N, mode, variant).
template <typename Real, int N, bool ROW, int V = 0> struct Geo {
  using PL = PlanOf<N, var_alt(V)>;
  static constexpr int T = PL::T;
  static constexpr int CB = (int)sizeof(typename cvec<Real>::T);
  // LDS bytes of a block's transform exchange (padded layout; FP32 moves one real component
  // at a time)
  static constexpr int lds_cap(int w) {
    using T = typename cvec<Real>::T;
    return ROW ? w * row_stride<T, N>() * (int)sizeof(T) : (w * N + ((w * N) >> RowAddr<T>::SH)) * kExchangeBytes<T>;
  }
  // the data the tile holds, which the occupancy budgets below are stated for
  static constexpr int lds_data(int w) { return w * N * CB; }
  static constexpr int base() {
    if (ROW) {
      // at least 128 threads per block, at most 32 KiB of LDS: several blocks per CU overlap
      // their load, transform and store phases
      int w = 128 / T;
      if (w < 1) w = 1;
      while (w > 1 && lds_data(w) > 32768) w /= 2;
      return w;
    }
    // columns: W adjacent columns per block; 64 KiB LDS keeps two blocks per CU
    int w = 16;
    while (w > 1 && (lds_data(w) > 65536 || w * T > 1024)) w /= 2;
    return w;
  }
  // Blocks of 1024 threads need four waves per SIMD, i.e. at most 128 VGPRs: plans with 16 or
  // more elements per thread do not fit that without spilling, so their wide variant stops at
  // 512 threads.
  static constexpr int pick() {
    int w = base();
    const int most = PL::E >= 16 ? 512 : 1024;
    if (var_wshift(V) < 0 && w > 1) w /= 2;
    if (var_wshift(V) > 0 && 2 * w * T <= most && lds_cap(2 * w) <= 163840) w *= 2;
    return w;
  }
  static constexpr int W = pick();
  static constexpr int THREADS = W * T;
  // Rows held one per wave (or several per wave): their exchanges need no block barrier.
  static constexpr bool WAVE_LINES = ROW && T <= 64 && 64 % T == 0;
  // Flux gather of a row pass: line w's complex values in its own exchange row (rows exchange
  // whole complex values), so a line's LDS is only ever touched by its own threads.
  static constexpr int kAlign(int b) { return (b + 15) & ~15; }
  static constexpr int gather_stride = row_stride<typename cvec<Real>::T, N>();
  static constexpr int gather_off = 0;
  static constexpr int gather_end = W * gather_stride * CB;
  // per-wave observation sums (3 per wave) after everything else
  static constexpr int NWAVES = (THREADS + 63) / 64;
  static constexpr int red_off = kAlign(ROW && gather_end > lds_cap(W) ? gather_end : lds_cap(W));
  static constexpr int red_bytes = 3 * NWAVES * (int)sizeof(double);
};

// Column-mode exchange layout: element p of column w at p*W + w, padded like the rows (with
// narrow tiles, W = 2 for 2048-point FP64 columns, stage-0 writes of 16 lanes otherwise share
// one 32-byte bank window).
template <typename T, int W> struct ColAddr {
  static constexpr bool kWave = false;
  static constexpr bool kSplitOk = true;
  int w;
  __device__ __forceinline__ int operator()(int p) const {
    const int a = p * W + w;
    return a + (a >> RowAddr<T>::SH);
  }
};

__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

template <typename T>
__device__ __forceinline__ T ld(const void* tab, int64_t i) {
  return reinterpret_cast<const T*>(tab)[i];
}

// Per-thread evaluator of an X factor for the elements of one transform along `axis`
// (scalar members only: runtime-indexed arrays would live in scratch memory).
// Factors are complex FP64 in both precisions (tables and their products); the state meets them
// once per element through capply (one rounding to the state type).
template <int N> struct XEval {
  using T = double2;
  T cst;
  const T* ax;
  const T* pr0;
  const T* pr1;
  int o0, o1, npr, kind;
  __device__ void init(const XFactor& f, int axis, int i0, int i1, int i2) {
    auto id = [&](int d) { return d == 0 ? i0 : (d == 1 ? i1 : i2); };
    cst = T{1.0, 0.0};
    ax = reinterpret_cast<const T*>(axis == 0 ? f.axis[0] : (axis == 1 ? f.axis[1] : f.axis[2]));
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (d != axis && f.axis[d]) cst = cmul(cst, ld<T>(f.axis[d], id(d)));
    npr = 0;
    o0 = o1 = 0;
    pr0 = pr1 = nullptr;
    kind = f.pair_kind;
    if (kind != PAIR_NONE) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (!(f.pair_mask & (1 << k))) continue;
        const int p = k, q = (k + 1) % 3;   // pairs (0,1), (1,2), (2,0)
        if (p != axis && q != axis) {
          int64_t i = kind == PAIR_DIFF ? (int64_t)abs(id(p) - id(q)) : (int64_t)id(p) * f.pair_n + id(q);
          cst = cmul(cst, ld<T>(f.pair, i));
        } else {
          const int o = (p == axis) ? id(q) : id(p);
          const T* base = reinterpret_cast<const T*>(f.pair) + (kind == PAIR_FULL ? (int64_t)o * f.pair_n : 0);
          if (npr == 0) { pr0 = base; o0 = o; }
          else { pr1 = base; o1 = o; }
          ++npr;
        }
      }
    }
  }
  // product of the pair terms that involve the transform axis (1 if none)
  __device__ __forceinline__ T pair_at(int n) const {
    T r{1.0, 0.0};
    if (npr > 0) r = kind == PAIR_DIFF ? pr0[abs(n - o0)] : pr0[n];
    if (npr > 1) r = cmul(r, kind == PAIR_DIFF ? pr1[abs(n - o1)] : pr1[n]);
    return r;
  }
  // factor without the axis-dependent pair part
  __device__ __forceinline__ T base_at(int n) const { return ax ? cmul(cst, ax[n]) : cst; }
  __device__ __forceinline__ T at(int n) const { return npr ? cmul(base_at(n), pair_at(n)) : base_at(n); }
};

__device__ __forceinline__ double pot_value(const ObsParams& o, const double* c, int D) {
  // closed forms of potential.h (full or partial by free axes), evaluated in FP64
  auto Ne = [&](double x) { return o.NE / sqrt(x * x + o.Nsoft); };
  auto ee = [&](double x, double y) {
    if (o.pot_kind == 3) return o.EE / sqrt((x - y) * (x - y) + x * y + o.Esoft);
    return o.EE / sqrt((x - y) * (x - y) + o.Esoft);
  };
  const int fm = o.free_mask;
  auto b = [&](int d) { return !(fm & (1 << d)); };
  if (o.pot_kind == 0) return 0.0;
  if (o.pot_kind == 1) {  // Coulomb over the bounded coordinates
    double s = 0.0;
    for (int d = D - 1; d >= 0; --d)
      if (b(d)) s = c[d] * c[d] + s;
    bool any = false;
    for (int d = 0; d < D; ++d) any = any || b(d);
    return any ? o.NE / sqrt(s + o.Nsoft) : 0.0;
  }
  double v = 0.0;
  if (D == 1) return b(0) ? Ne(c[0]) : 0.0;
  if (D == 2) {
    if (b(0) && b(1)) v = ee(c[0], c[1]);
    if (b(0)) v += Ne(c[0]);
    if (b(1)) v += Ne(c[1]);
    return v;
  }
  if (b(0) && b(1)) v += ee(c[0], c[1]);
  if (b(1) && b(2)) v += ee(c[1], c[2]);
  if (b(2) && b(0)) v += ee(c[2], c[0]);
  for (int d = 0; d < 3; ++d)
    if (b(d)) v += Ne(c[d]);
  return v;
}

// Geometry of one thread's transform in tile b: (outer o, inner c), its global offsets and the
// grid coordinates of its line (the transform axis index is per element).
template <int N, bool ROW, int W>
struct TileGeo {
  int64_t o, c, base, step;
  bool valid;
  int i0, i1, i2;
  __device__ __forceinline__ void at(const PassParams& p, int64_t b, int w) {
    if (ROW) {
      o = b * W + w;
      c = 0;
      valid = o < p.outer;
    } else {
      const int64_t cpt = (p.inner + W - 1) / W;  // column tiles per outer index
      o = b / cpt;
      c = (b % cpt) * W + w;
      valid = (o < p.outer) && (c < p.inner);
    }
    base = ROW ? o * N : o * (int64_t)N * p.inner + c;
    step = ROW ? 1 : p.inner;
    i0 = i1 = i2 = 0;
    if (p.axis == 0) {
      i1 = (int)(c / p.n[2]);
      i2 = (int)(c % p.n[2]);
    } else if (p.axis == 1) {
      i0 = (int)o;
      i2 = (int)c;
    } else {
      i0 = (int)(o / p.n[1]);
      i1 = (int)(o % p.n[1]);
    }
    if (!valid) i0 = i1 = i2 = 0;   // idle threads of a partial tile read in-range table entries
  }
};

template <int W, bool ROW>
__device__ __forceinline__ int64_t tile_count(const PassParams& p) {
  return ROW ? (p.outer + W - 1) / W : p.outer * ((p.inner + W - 1) / W);
}

// Tile of the k-th position in launch order (k in [0, ntiles)). 3D rows (axis 2) read pair-table
// rows indexed by i1 and by i0 (the (1,2) and (2,0) pairs). Walking lines with i1 fastest cycles
// all n rows of the table through L2 every n lines (512 x 8 KiB at 512^3: re-fetched from HBM,
// +2 GiB per pass); walking (i0, i1) in 32 x 64 patches keeps the rows the in-flight tiles of
// one XCD use resident.
template <int W, bool ROW>
__device__ __forceinline__ int64_t tile_at(const PassParams& p, int64_t k, int64_t ntiles) {
  if (ROW && p.axis == 2 && p.n[2] > 1) {
    constexpr int64_t P0 = 32, P1 = 64;
    const int64_t n0 = p.n[0], n1 = p.n[1];
    if (n0 % P0 == 0 && n1 % P1 == 0 && P1 % W == 0 && ntiles * W == n0 * n1) {
      const int64_t tpp = P0 * (P1 / W);  // tiles per patch
      const int64_t pidx = k / tpp, r = k % tpp, pc = n1 / P1;
      const int64_t j0 = (pidx / pc) * P0 + r / (P1 / W);
      const int64_t j1 = (pidx % pc) * P1 + (r % (P1 / W)) * W;
      return (j0 * n1 + j1) / W;  // the tile whose first line is (j0, j1)
    }
  }
  return k;
}

// One tile of the pass on the state `v` already loaded: everything after the load, through the
// store. An observation leaves each wave's sums in the reduction area of LDS.
template <typename Real, int N, bool ROW, int V, int FEAT, bool S32>
__device__ __forceinline__ void pass_tile(const PassParams& p, const TileGeo<N, ROW, Geo<Real, N, ROW, V>::W>& g,
                                          typename Tr<Real>::T (&v)[PlanOf<N, var_alt(V)>::E], int tid, int t, int w,
                                          unsigned char* smem_raw, StateOf<Real, S32>* dst, const void* pf0,
                                          const void* pf1, const void* pf2, void* gather_out) {
  using T = typename Tr<Real>::T;
  using S = StateOf<Real, S32>;
  using G = Geo<Real, N, ROW, V>;
  constexpr int ALT = var_alt(V);
  using PL = PlanOf<N, ALT>;
  constexpr int E = PL::E, TT = PL::T;
  constexpr int W = Geo<Real, N, ROW, V>::W;
  T* lds = reinterpret_cast<T*>(smem_raw);
  T* glds = reinterpret_cast<T*>(smem_raw + G::gather_off);
  const T* roots = reinterpret_cast<const T*>(p.roots);
  const int64_t o = g.o, base = g.base, step = g.step;
  const bool valid = g.valid;
  const int i0 = g.i0, i1 = g.i1, i2 = g.i2;
  auto id = [&](int d) { return d == 0 ? i0 : (d == 1 ? i1 : i2); };

  // this thread's twiddles for every transform of the pass, and the flux gather's row range,
  // read with the state
  Twiddles<N, ALT, T, FEAT == FEAT_LEAN, !ROW> tw;   // FP32 column tiles read twiddles directly
  if (p.flags & (F_INV1 | F_FWD | F_INV2)) tw.load(roots, t);
  int gk0 = 0, gk1 = 0;
  if (FEAT != FEAT_LEAN && ROW && (p.flags & F_OBS) && p.obs.gather_rowptr && valid) {
    gk0 = p.obs.gather_rowptr[o];
    gk1 = p.obs.gather_rowptr[o + 1];
  }

  if (p.flags & F_SCALE) {
    // the scale (1/m of an inverse transform: 1/768 has no exact FP32 form) is applied in FP64
    // with one rounding; a once-rounded FP32 scale would bias the norm the same way every step
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = capply(v[q], double2{p.scale, 0.0});
  }

  auto run_fft = [&](auto dir_tag) {
    constexpr int DIR = decltype(dir_tag)::value;
    if constexpr (ROW) {
      RowAddr<T, Geo<Real, N, ROW, V>::WAVE_LINES> a{w * row_stride<T, N>()};
      fft<N, ALT, DIR>(v, t, lds, a, roots, tw);
    } else {
      ColAddr<T, W> a{w};
      fft<N, ALT, DIR>(v, t, lds, a, roots, tw);
    }
  };
  using FWD = std::integral_constant<int, -1>;
  using INV = std::integral_constant<int, +1>;

  if (p.flags & F_INV1) run_fft(INV{});

  // X stage: A factor -> observables / region mask -> B factor, one loop over the elements.
  // When both halves use the same pair-interaction table its lookup is shared.
  const uint32_t XF = FEAT != FEAT_LEAN ? p.flags & (F_XA | F_OBS | F_MASK | F_XB) : 0u;
  if (XF) {
    const bool doA = p.flags & F_XA, doB = p.flags & F_XB, obs = p.flags & F_OBS, msk = p.flags & F_MASK;
    XEval<N> xa, xb;
    if (doA) xa.init(p.xa, p.axis, i0, i1, i2);
    if (doB) xb.init(p.xb, p.axis, i0, i1, i2);
    const bool shared = doA && doB && xa.npr > 0 && p.xa.pair_kind == p.xb.pair_kind &&
                        p.xa.pair == p.xb.pair && p.xa.pair_mask == p.xb.pair_mask &&
                        p.xa.delta == p.xb.delta && p.xa.mode == p.xb.mode;
    // both halves are one and the same factor (the velocity gauge's X factor does not change
    // from step to step, and the host then hands the same tables twice): evaluate it once
    const bool same = shared && p.xa.axis[0] == p.xb.axis[0] && p.xa.axis[1] == p.xb.axis[1] &&
                      p.xa.axis[2] == p.xb.axis[2];
    const ObsParams& ob = p.obs;
    double nrm = 0.0, ep = 0.0, ek = 0.0, kc = 0.0;
    const double* ka = nullptr;
    int D = 1;
    if (obs) {
      D = p.n[2] > 1 ? 3 : (p.n[1] > 1 ? 2 : 1);
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (d != p.axis && ob.kin_w[d]) kc += ob.kin_w[d][id(d)];
      ka = p.axis == 0 ? ob.kin_w[0] : (p.axis == 1 ? ob.kin_w[1] : ob.kin_w[2]);
    }
    // region mask: offset of the other axes' nodes
    int64_t mo = 0, msa = 0, na = 1;
    if (msk) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (d == p.axis || p.mask_stride[d] == 0) continue;
        const int64_t h = p.n[d] / 2, i = id(d);
        const int64_t nd = i < h ? -i : i - p.n[d] + 1;
        int64_t node = p.nCAP + nd;
        node = node < 0 ? 0 : node;
        mo += node * p.mask_stride[d];
      }
      msa = p.axis == 0 ? p.mask_stride[0] : (p.axis == 1 ? p.mask_stride[1] : p.mask_stride[2]);
      na = p.axis == 0 ? p.n[0] : (p.axis == 1 ? p.n[1] : p.n[2]);
    }
    const bool gather = ROW && obs && ob.gather_rowptr;
    const bool gline = gather && gk1 > gk0;   // this line has border points to gather
    // The state between the halves is formed explicitly only where something acts on it or keeps
    // it (the region mask, the multigrid edges, the copy for the kinetic energy). Otherwise both
    // halves meet in one product per element, the observables weight |psi|^2 by |A factor|^2,
    // and the flux gather keeps the state before A and applies A at the points it reads.
    const bool compact = FEAT == FEAT_X || (!msk && !(ROW && p.er_out) && !(obs && ob.copy_out));
    if (compact) {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int ia = t + TT * q;
        double2 zab{1.0, 0.0};
        double wa = 1.0;   // |A factor|^2
        if (same) {
          const double2 z = cmul(xa.base_at(ia), xa.pair_at(ia));
          zab = cmul(z, z);
          if (obs) wa = cabs2(z);
        } else if (shared) {
          const double2 pp = xa.pair_at(ia), ba = xa.base_at(ia);
          zab = cmul(cmul(ba, xb.base_at(ia)), cmul(pp, pp));
          if (obs) wa = cabs2(ba) * cabs2(pp);
        } else if (doA) {
          const double2 za = xa.at(ia);
          zab = doB ? cmul(za, xb.at(ia)) : za;
          if (obs) wa = cabs2(za);
        } else if (doB) {
          zab = xb.at(ia);
        }
        if (obs && valid) {
          const double a2 = ((double)v[q].x * (double)v[q].x + (double)v[q].y * (double)v[q].y) * wa;
          nrm += a2;
          if (ob.want_epot) {
            double co[3];
            co[0] = ob.xmin[0] + (p.axis == 0 ? ia : i0) * ob.dx[0];
            co[1] = ob.xmin[1] + (p.axis == 1 ? ia : i1) * ob.dx[1];
            co[2] = ob.xmin[2] + (p.axis == 2 ? ia : i2) * ob.dx[2];
            ep += a2 * pot_value(ob, co, D);
          }
          if (ka || kc != 0.0) ek += a2 * (kc + (ka ? ka[ia] : 0.0));
        }
        if (gline) glds[w * G::gather_stride + ia] = v[q];
        if (doA || doB) v[q] = capply(v[q], zab);
      }
    } else if constexpr (FEAT == FEAT_FULL) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int ia = t + TT * q;
      double2 pp, za;
      if (shared) pp = xa.pair_at(ia);
      if (doA) {
        za = shared ? cmul(xa.base_at(ia), pp) : xa.at(ia);
        v[q] = capply(v[q], za);
      }
      if (obs && valid) {
        const double a2 = (double)v[q].x * (double)v[q].x + (double)v[q].y * (double)v[q].y;
        nrm += a2;
        if (ob.want_epot) {
          double co[3];
          co[0] = ob.xmin[0] + (p.axis == 0 ? ia : i0) * ob.dx[0];
          co[1] = ob.xmin[1] + (p.axis == 1 ? ia : i1) * ob.dx[1];
          co[2] = ob.xmin[2] + (p.axis == 2 ? ia : i2) * ob.dx[2];
          ep += a2 * pot_value(ob, co, D);
        }
        if (ka || kc != 0.0) ek += a2 * (kc + (ka ? ka[ia] : 0.0));
        if (ob.copy_out) reinterpret_cast<S*>(ob.copy_out)[base + (int64_t)ia * step] = to_state<S>(v[q]);
      }
      if (gline) glds[w * G::gather_stride + ia] = v[q];
      if (ROW && p.er_out && valid) {
        // multigrid source edges of the boundary state (2D main region, row i0 = o)
        const int64_t nc = p.nCAP, n0 = p.n[0], n1 = N;
        if (o < nc || o >= n0 - nc) {
          const int64_t e = o < nc ? o : o - (n0 - 2 * nc);
          reinterpret_cast<S*>(p.er_out)[e * n1 + ia] = to_state<S>(v[q]);
        }
        if (ia < nc || ia >= n1 - nc) {
          const int64_t e = ia < nc ? ia : ia - (n1 - 2 * nc);
          reinterpret_cast<S*>(p.ec_out)[o * 2 * nc + e] = to_state<S>(v[q]);
        }
      }
      if (msk) {
        // maskRegion (grid.h:582-597): one exact FP64 factor per point
        int64_t off = mo;
        if (msa) {
          const int64_t nd = ia < na / 2 ? -ia : ia - na + 1;
          int64_t node = p.nCAP + nd;
          node = node < 0 ? 0 : node;
          off += node * msa;
        }
        const double mk = p.mask_tab[off];
        v[q] = T{(Real)((double)v[q].x * mk), (Real)((double)v[q].y * mk)};
      }
      if (doB) v[q] = capply(v[q], same ? za : (shared ? cmul(xb.base_at(ia), pp) : xb.at(ia)));
    }
    }
    if (obs) {
      // wave sums of the requested observables (fixed order); the kernel adds the waves' sums
      // once the tile is done
      const bool want_ek = ob.kin_w[0] || ob.kin_w[1] || ob.kin_w[2];
      for (int off = 32; off > 0; off >>= 1) {
        nrm += __shfl_xor(nrm, off);
        if (ob.want_epot) ep += __shfl_xor(ep, off);
        if (want_ek) ek += __shfl_xor(ek, off);
      }
      if (gather) {
        exchange_sync<G::WAVE_LINES>();
        if (valid) {
          S* go = reinterpret_cast<S*>(gather_out);
          for (int k = gk0 + t; k < gk1; k += TT) {
            const int pos = ob.gather_pos[k];
            T val = glds[w * G::gather_stride + pos];
            if (compact && doA) val = capply(val, xa.at(pos));   // the same product as a full pass
            go[k] = to_state<S>(val);
          }
        }
        exchange_sync<G::WAVE_LINES>();   // the line's LDS is written again by the next exchange
      }
      double* red = reinterpret_cast<double*>(smem_raw + G::red_off);
      const int wave = tid >> 6;
      if ((tid & 63) == 0) { red[3 * wave] = nrm; red[3 * wave + 1] = ep; red[3 * wave + 2] = ek; }
    }
  }

  if (p.flags & F_FWD) run_fft(FWD{});

  if (p.flags & F_P) {
    // P phase: complex FP64 per-axis tables, one FP64 product per element
    double2 cst{1.0, 0.0};
    if (p.axis != 0 && pf0) cst = cmul(cst, ld<double2>(pf0, i0));
    if (p.axis != 1 && pf1) cst = cmul(cst, ld<double2>(pf1, i1));
    if (p.axis != 2 && pf2) cst = cmul(cst, ld<double2>(pf2, i2));
    const double2* pa = reinterpret_cast<const double2*>(p.axis == 0 ? pf0 : (p.axis == 1 ? pf1 : pf2));
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const double2 f = pa ? cmul(cst, pa[t + TT * q]) : cst;
      v[q] = capply(v[q], f);
    }
  }

  if (p.flags & F_INV2) run_fft(INV{});

  if (valid && !(p.flags & F_NOSTORE)) {
#pragma unroll
    for (int q = 0; q < E; ++q) dst[base + (int64_t)(t + TT * q) * step] = to_state<S>(v[q]);
  }
}

// Waves per SIMD the register allocation must leave room for. Left alone, the scheduler issues
// every table load of the unrolled element loop at once and a full kernel lands at 169-171
// VGPRs: two waves per SIMD, where 168 would give three. Plans with 8 elements per thread fit
// four waves in the lean tier (twiddles preloaded; the column tiles' LDS allows no more) and in
// full row passes of 64k threads per line (their line geometry is wave-uniform, in scalar
// registers; a few spilled
// bytes cost far less than the fourth wave returns: the 512^3 row pass 2.6 ms at three waves,
// 1.9 ms at four), three in full column passes; 16-element plans two (full) or three (lean);
// the 20- and 24-element plans of the 5*2^k and 3*2^k lengths need their registers and are left
// uncapped.
template <int E, int FEAT, bool UNIFORM_LINE>
constexpr int pass_waves() {
  if (E <= 8) return (FEAT == FEAT_LEAN || UNIFORM_LINE) ? 4 : 3;
  if (E <= 16) return FEAT == FEAT_LEAN ? 3 : 2;   // (the X tier counts as full)
  return 1;
}

template <typename Real, int N, bool ROW, int V, int FEAT, bool S32>
__global__ void __launch_bounds__((Geo<Real, N, ROW, V>::THREADS),
                                  (pass_waves<PlanOf<N, var_alt(V)>::E, FEAT, ROW && PlanOf<N, var_alt(V)>::T % 64 == 0>()))
    axis_pass_kernel(PassParams p) {
  using T = typename Tr<Real>::T;
  constexpr int ALT = var_alt(V);
  using PL = PlanOf<N, ALT>;
  constexpr int E = PL::E, TT = PL::T;
  constexpr int W = Geo<Real, N, ROW, V>::W;
  extern __shared__ __align__(16) unsigned char smem_raw[];

  const int tid = threadIdx.x;
  const int t = ROW ? (tid % TT) : (tid / W);
  // rows of 64k threads: a wave holds part of one line, so the line index and everything
  // derived from it (offsets, coordinates, table rows, constant factors) is wave-uniform and
  // lives in scalar registers
  const int w = ROW ? (TT % 64 == 0 ? __builtin_amdgcn_readfirstlane(tid / TT) : tid / TT) : (tid % W);

  using S = StateOf<Real, S32>;
  const S* src = reinterpret_cast<const S*>(p.src);
  S* dst = reinterpret_cast<S*>(p.dst);
  const void* pf0 = p.pf[0];
  const void* pf1 = p.pf[1];
  const void* pf2 = p.pf[2];
  double* partials = p.obs.partials;
  void* gather_out = p.obs.gather_out;
  if (p.batch) {
    // batched launch: this block's solve (whole block leaves together, before any barrier)
    const BatchEntry& be = p.batch[blockIdx.y];
    if (*p.step_ctr >= be.total_steps) return;
    src = reinterpret_cast<const S*>(be.psi);
    dst = reinterpret_cast<S*>(be.psi);
    pf0 = be.pf[0];
    pf1 = be.pf[1];
    pf2 = be.pf[2];
    partials = be.partials;
    gather_out = be.gather_out;
  }
  const int64_t ntiles = tile_count<W, ROW>(p);
  const int64_t nb = gridDim.x;
  // XCD-aware order: blocks go round-robin over the 8 XCDs, so block i takes position
  // (i % 8) * (nb / 8) + i / 8 and each XCD walks a contiguous run of tiles (its L2 sees the
  // neighbours' halves of shared lines)
  const int64_t k0 = (nb & 7) == 0 ? (int64_t)(blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3) : (int64_t)blockIdx.x;
  TileGeo<N, ROW, W> g;
  g.at(p, tile_at<W, ROW>(p, k0, ntiles), w);
  T v[E];
#pragma unroll
  for (int q = 0; q < E; ++q)
    v[q] = g.valid ? from_state<T>(src[g.base + (int64_t)(t + TT * q) * g.step]) : T{Real(0), Real(0)};
  pass_tile<Real, N, ROW, V, FEAT, S32>(p, g, v, tid, t, w, smem_raw, dst, pf0, pf1, pf2, gather_out);
  if (FEAT != FEAT_LEAN && (p.flags & F_OBS)) {
    // the block's sums: its waves' sums in wave order
    using G = Geo<Real, N, ROW, V>;
    block_sync();
    if (tid == 0 && partials) {
      const double* red = reinterpret_cast<const double*>(smem_raw + G::red_off);
      double a = 0.0, e2 = 0.0, e3 = 0.0;
      for (int k = 0; k < G::NWAVES; ++k) { a += red[3 * k]; e2 += red[3 * k + 1]; e3 += red[3 * k + 2]; }
      partials[3 * blockIdx.x] = a;
      partials[3 * blockIdx.x + 1] = e2;
      partials[3 * blockIdx.x + 2] = e3;
    }
  }
}

// Blocks of one launch: one per tile.
template <typename Real, int N, bool ROW, int V>
inline int64_t launch_blocks(const PassParams& p) {
  using G = Geo<Real, N, ROW, V>;
  return ROW ? (p.outer + G::W - 1) / G::W : p.outer * ((p.inner + G::W - 1) / G::W);
}

template <typename Real, int N, bool ROW, int V, int FEAT, bool S32>
inline int launch_feat(const PassParams& p, hipStream_t s) {
  using G = Geo<Real, N, ROW, V>;
  // LDS per launch: the transform exchange; for a flux gather the complex lines (inside the
  // exchange rows in FP64, after them in FP32); for an observation the waves' sums after both.
  // A pointwise pass (a region whose axes all stay in P) without an observation needs none.
  constexpr size_t most = FEAT != FEAT_LEAN ? (size_t)(G::red_off + G::red_bytes) : (size_t)G::lds_cap(G::W);
  static_assert(most <= 163840, "a block's LDS exceeds the 160 KiB of a CU");
  // once per kernel (a function-local static initialises once, also under concurrent callers)
  static const bool attr_set = [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&axis_pass_kernel<Real, N, ROW, V, FEAT, S32>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)most);
    return true;
  }();
  (void)attr_set;
  size_t lds = (p.flags & (F_INV1 | F_FWD | F_INV2)) ? (size_t)G::lds_cap(G::W) : 0;
  if (FEAT != FEAT_LEAN && (p.flags & F_OBS)) lds = most;
  const unsigned ny = p.batch ? (unsigned)p.nbatch : 1u;
  const int64_t blocks = launch_blocks<Real, N, ROW, V>(p);
  hipLaunchKernelGGL((axis_pass_kernel<Real, N, ROW, V, FEAT, S32>), dim3((unsigned)blocks, ny), dim3(G::THREADS), lds, s,
                     p);
  return (int)blocks;
}

template <typename Real, int N, bool ROW, int V, bool S32>
inline int launch_one(const PassParams& p, hipStream_t s) {
  const int f = feat_of(p, ROW);
  if (f == FEAT_LEAN) return launch_feat<Real, N, ROW, V, FEAT_LEAN, S32>(p, s);
  if constexpr (ROW) {
    if (f == FEAT_X) return launch_feat<Real, N, ROW, V, FEAT_X, S32>(p, s);
  }
  return launch_feat<Real, N, ROW, V, FEAT_FULL, S32>(p, s);
}

// Blocks of a launch, which is also the number of per-block observation sums it writes.
template <typename Real, int N, bool ROW, int V>
inline int64_t tiles_one(const PassParams& p) {
  return launch_blocks<Real, N, ROW, V>(p);
}

// Lengths with tuned variants; every other supported length has variant 0 only.
constexpr bool tuned_length(int64_t n) {
  return n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096 || n == 768 || n == 1536;
}
inline bool has_variants(int64_t n) { return tuned_length(n); }

// Launcher and tile count of one axis length (all its variants). Instantiated explicitly per
// (precision, mode, length group) in axis_pass_inst.hip, so the heavy kernels compile in
// parallel translation units; axis_pass.hip dispatches on the runtime length. S32: FP64
// arithmetic over an FP32 state (same tiles as the FP64 kernels).
template <typename Real, bool ROW, int N, bool S32 = false>
int launch_size(int v, const PassParams& p, hipStream_t s) {
  if constexpr (tuned_length(N)) {
    switch (v) {
      case 1: return launch_one<Real, N, ROW, 1, S32>(p, s);
      case 2: return launch_one<Real, N, ROW, 2, S32>(p, s);
      case 3: return launch_one<Real, N, ROW, 3, S32>(p, s);
      case 4: return launch_one<Real, N, ROW, 4, S32>(p, s);
      case 5: return launch_one<Real, N, ROW, 5, S32>(p, s);
      default: break;
    }
  }
  return launch_one<Real, N, ROW, 0, S32>(p, s);
}

template <typename Real, bool ROW, int N>
int64_t tiles_size(int v, const PassParams& p) {
  if constexpr (tuned_length(N)) {
    switch (v) {
      case 1: return tiles_one<Real, N, ROW, 1>(p);
      case 2: return tiles_one<Real, N, ROW, 2>(p);
      case 3: return tiles_one<Real, N, ROW, 3>(p);
      case 4: return tiles_one<Real, N, ROW, 4>(p);
      case 5: return tiles_one<Real, N, ROW, 5>(p);
      default: break;
    }
  }
  return tiles_one<Real, N, ROW, 0>(p);
}

#define QSF_SIZES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) \
  X(12) X(24) X(48) X(96) X(192) X(384) X(768) X(1536) X(3072)                         \
  X(10) X(20) X(40) X(80) X(160) X(320) X(640) X(1280) X(2560)

}  // namespace qsf
