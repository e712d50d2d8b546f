// fft_core.hpp — register-resident Stockham FFT building blocks for gfx950 (CDNA4).
//
// One length-N transform is spread over T = N/E threads; thread t holds the E elements
// x[t + T*q], q = 0..E-1. With that assignment every Stockham stage reads its inputs in the
// same pattern, and the last stage writes its outputs back into the same registers, so a
// transform costs (stages - 1) LDS exchanges and its global loads/stores are coalesced
// (consecutive threads touch consecutive elements of a row, or consecutive columns of a tile).
//
// The forward transform is e^{-2πi jk/N}, unnormalised, as FFTW_FORWARD in the reference
// (wf/grid.h:251-274); the inverse is e^{+2πi jk/N}, unnormalised here (the 1/N is folded into
// the phase tables by the caller).
#pragma once
#include <hip/hip_runtime.h>

namespace qsf {

// Block barrier for data exchanged through LDS. __syncthreads() is s_barrier between two
// workgroup fences, and the compiler has been seen to sink an LDS read of an exchange below it
// (into the branch of the next butterfly stage, which needs that element last); the read then
// races with other waves' writes of the next exchange and the transform comes out wrong, at
// random. An empty asm statement with a memory clobber on each side is a code-motion barrier for
// every memory access and emits no instruction.
__device__ __forceinline__ void block_sync() {
  asm volatile("" ::: "memory");
  __syncthreads();
  asm volatile("" ::: "memory");
}

// Exchange barrier of data that stays within one wave (a transform whose threads are all in one
// wave). A wave's LDS instructions execute in issue order, so no hardware barrier is needed: the
// wavefront-scope fences and the asm clobbers only keep the compiler from moving LDS accesses
// across this point.
__device__ __forceinline__ void wave_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  asm volatile("" ::: "memory");
}
template <bool WAVE> __device__ __forceinline__ void exchange_sync() {
  if constexpr (WAVE) wave_sync();
  else block_sync();
}

template <typename R> struct cvec;
template <> struct cvec<double> { using T = double2; };
template <> struct cvec<float> { using T = float2; };

template <typename T> __device__ __forceinline__ T cadd(T a, T b) { return T{a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ T csub(T a, T b) { return T{a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ T cmul(T a, T b) {
  return T{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ T cconj(T a) { return T{a.x, -a.y}; }
template <typename T, typename R> __device__ __forceinline__ T cscale(T a, R s) { return T{a.x * s, a.y * s}; }
// multiply by DIR*i
template <int DIR, typename T> __device__ __forceinline__ T mul_di(T a) {
  return DIR < 0 ? T{a.y, -a.x} : T{-a.y, a.x};
}

// Products by constants and tables that are the same at every step. In FP32 a once-rounded
// constant biases every transform the same way (fl(sqrt(1/2)) is 1.7e-8 short, so each
// rotation shrinks the amplitude: a norm drift linear in the step count). The FP32 forms carry
// the rounding remainder, c = c_hi + c_lo, and add v*c_lo inside the fused multiply-add of
// v*c_hi, so the correction (below half an ulp) enters before the single rounding instead of
// being rounded away afterwards.
__device__ __forceinline__ float2 cmul_comp(float2 v, float2 h, float2 l) {
  const float cr = v.x * l.x - v.y * l.y;   // small: v * c_lo
  const float ci = v.x * l.y + v.y * l.x;
  const float re = __builtin_fmaf(v.x, h.x, __builtin_fmaf(-v.y, h.y, cr));
  const float im = __builtin_fmaf(v.x, h.y, __builtin_fmaf(v.y, h.x, ci));
  return float2{re, im};
}
template <typename T> __device__ __forceinline__ T cmul_k(T v, double cr, double ci) {
  using R = decltype(v.x);
  if constexpr (sizeof(R) == 8) {
    return cmul(v, T{cr, ci});
  } else {
    const float hr = (float)cr, hi = (float)ci;
    const float lr = (float)(cr - (double)hr), li = (float)(ci - (double)hi);
    return cmul_comp(v, float2{hr, hi}, float2{lr, li});
  }
}
template <typename T> __device__ __forceinline__ T cscale_k(T v, double c) {
  using R = decltype(v.x);
  if constexpr (sizeof(R) == 8) {
    return T{v.x * c, v.y * c};
  } else {
    const float h = (float)c, l = (float)(c - (double)h);
    return T{__builtin_fmaf(v.x, h, v.x * l), __builtin_fmaf(v.y, h, v.y * l)};
  }
}
// state times an FP64 phase factor: FP64 product, one rounding to the state type
template <typename T> __device__ __forceinline__ T capply(T v, double2 z) {
  using R = decltype(v.x);
  if constexpr (sizeof(R) == 8) {
    return cmul(v, z);
  } else {
    const double2 r = cmul(double2{(double)v.x, (double)v.y}, z);
    return T{(float)r.x, (float)r.y};
  }
}

// ---------------------------------------------------------------- small DFTs (in registers)
template <int DIR, typename T> __device__ __forceinline__ void dft2(T& a, T& b) {
  T t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR, typename T> __device__ __forceinline__ void dft3(T& a, T& b, T& c) {
  using R = decltype(a.x);
  const R h = R(0.5);
  T t1 = cadd(b, c), t2 = csub(b, c);
  T m = T{a.x - h * t1.x, a.y - h * t1.y};
  T r = mul_di<DIR>(cscale_k(t2, 0.86602540378443864676372317075294));  // sqrt(3)/2
  a = cadd(a, t1);
  b = cadd(m, r);
  c = csub(m, r);
}

// DFT of length 5 with w = exp(DIR 2πi/5): conjugate pairs w^k, w^(5-k) share the real parts
// cos(2π/5), cos(4π/5) on the sums b+e, c+d and the imaginary parts on the differences
template <int DIR, typename T> __device__ __forceinline__ void dft5(T& a, T& b, T& c, T& d, T& e) {
  constexpr double C1 = 0.30901699437494742410229341718282;    // cos(2π/5)
  constexpr double C2 = -0.80901699437494742410229341718282;   // cos(4π/5)
  constexpr double S1 = 0.95105651629515357211643933337938;    // sin(2π/5)
  constexpr double S2 = 0.58778525229247312916870595463907;    // sin(4π/5)
  const T t1 = cadd(b, e), t2 = cadd(c, d), t3 = csub(b, e), t4 = csub(c, d);
  const T m1 = cadd(cadd(a, cscale_k(t1, C1)), cscale_k(t2, C2));
  const T m2 = cadd(cadd(a, cscale_k(t1, C2)), cscale_k(t2, C1));
  const T n1 = mul_di<DIR>(cadd(cscale_k(t3, S1), cscale_k(t4, S2)));
  const T n2 = mul_di<DIR>(csub(cscale_k(t3, S2), cscale_k(t4, S1)));
  a = cadd(cadd(a, t1), t2);
  b = cadd(m1, n1);
  e = csub(m1, n1);
  c = cadd(m2, n2);
  d = csub(m2, n2);
}

template <int DIR, typename T> __device__ __forceinline__ void dft4(T& a, T& b, T& c, T& d) {
  T s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = mul_di<DIR>(csub(b, d));
  a = cadd(s0, s2);
  c = csub(s0, s2);
  b = cadd(s1, s3);
  d = csub(s1, s3);
}

// w8^k for k = 1..3 with w8 = exp(DIR*2πi/8)
template <int DIR, int K, typename T> __device__ __forceinline__ T tw8(T v) {
  constexpr double c = 0.70710678118654752440084436210485;
  if constexpr (K == 0) return v;
  else if constexpr (K == 1) return cmul_k(v, c, DIR * c);
  else if constexpr (K == 2) return mul_di<DIR>(v);
  else return cmul_k(v, -c, DIR * c);
}

template <int DIR, typename T> __device__ __forceinline__ void dft8(T* x) {
  // even/odd split: A = DFT4(x0,x2,x4,x6), B = DFT4(x1,x3,x5,x7), X[k] = A[k] + w^k B[k]
  T a0 = x[0], a1 = x[2], a2 = x[4], a3 = x[6];
  T b0 = x[1], b1 = x[3], b2 = x[5], b3 = x[7];
  dft4<DIR>(a0, a1, a2, a3);
  dft4<DIR>(b0, b1, b2, b3);
  b1 = tw8<DIR, 1>(b1);
  b2 = tw8<DIR, 2>(b2);
  b3 = tw8<DIR, 3>(b3);
  x[0] = cadd(a0, b0); x[4] = csub(a0, b0);
  x[1] = cadd(a1, b1); x[5] = csub(a1, b1);
  x[2] = cadd(a2, b2); x[6] = csub(a2, b2);
  x[3] = cadd(a3, b3); x[7] = csub(a3, b3);
}

template <int DIR, typename T> __device__ __forceinline__ T tw16(T v, int m) {
  // exp(DIR*2πi*m/16) for the m values the radix-16 kernel uses (0..9)
  constexpr double c1 = 0.92387953251128675612818318939679;  // cos(π/8)
  constexpr double s1 = 0.38268343236508977172845998403040;  // sin(π/8)
  constexpr double c2 = 0.70710678118654752440084436210485;
  switch (m) {
    case 0: return v;
    case 1: return cmul_k(v, c1, DIR * s1);
    case 2: return cmul_k(v, c2, DIR * c2);
    case 3: return cmul_k(v, s1, DIR * c1);
    case 4: return mul_di<DIR>(v);
    case 6: return cmul_k(v, -c2, DIR * c2);
    case 9: return cmul_k(v, -c1, -DIR * s1);
    default: return v;
  }
}

template <int DIR, typename T> __device__ __forceinline__ void dft16(T* x) {
  // X[k1 + 4 k2] = sum_j0 w16^{j0 k1} w4^{j0 k2} sum_j1 x[j0 + 4 j1] w4^{j1 k1}
  T y[16];
#pragma unroll
  for (int j0 = 0; j0 < 4; ++j0) {
    T a = x[j0], b = x[j0 + 4], c = x[j0 + 8], d = x[j0 + 12];
    dft4<DIR>(a, b, c, d);
    y[j0 * 4 + 0] = a;
    y[j0 * 4 + 1] = tw16<DIR>(b, j0 * 1);
    y[j0 * 4 + 2] = tw16<DIR>(c, j0 * 2);
    y[j0 * 4 + 3] = tw16<DIR>(d, j0 * 3);
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    T a = y[0 * 4 + k1], b = y[1 * 4 + k1], c = y[2 * 4 + k1], d = y[3 * 4 + k1];
    dft4<DIR>(a, b, c, d);
    x[k1 + 0] = a;
    x[k1 + 4] = b;
    x[k1 + 8] = c;
    x[k1 + 12] = d;
  }
}

template <int R, int DIR, typename T> __device__ __forceinline__ void dft(T* v) {
  if constexpr (R == 2) dft2<DIR>(v[0], v[1]);
  else if constexpr (R == 3) dft3<DIR>(v[0], v[1], v[2]);
  else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 5) dft5<DIR>(v[0], v[1], v[2], v[3], v[4]);
  else if constexpr (R == 8) dft8<DIR>(v);
  else dft16<DIR>(v);
}

// ---------------------------------------------------------------- compile-time plans
struct Plan {
  int n, e, t, nst;
  int r[8];
};

// Elements per thread. `alt` selects the other register/LDS balance for the same length
// (fewer elements per thread: more threads, more LDS exchanges, fewer VGPRs; or the reverse).
__host__ __device__ constexpr int plan_e(int n, int alt = 0) {
  if (n % 5 == 0) return n <= 20 ? n : 20;   // 5·2^k: one radix-5 stage, then radix 4 / 2
  if (n % 3 == 0) {
    const int e = n <= 12 ? n : (n <= 24 ? n : 24);
    return (alt && e == 24 && n >= 48) ? 12 : e;
  }
  if (n <= 16) return n;
  const int e = (n <= 64 || n == 512) ? 8 : 16;
  if (!alt) return e;
  if (e == 16 && n >= 128) return 8;
  if (e == 8 && n >= 256) return 16;
  return e;
}

__host__ __device__ constexpr Plan make_plan(int n, int alt = 0) {
  Plan p{n, plan_e(n, alt), 0, 0, {0, 0, 0, 0, 0, 0, 0, 0}};
  p.t = n / p.e;
  int rem = n;
  if (rem % 3 == 0) { p.r[p.nst++] = 3; rem /= 3; }
  if (rem % 5 == 0) { p.r[p.nst++] = 5; rem /= 5; }
  const int cand[4] = {16, 8, 4, 2};
  while (rem > 1) {
    for (int i = 0; i < 4; ++i) {
      int r = cand[i];
      if (rem % r == 0 && p.e % r == 0) { p.r[p.nst++] = r; rem /= r; break; }
    }
  }
  return p;
}

template <int N, int ALT = 0> struct PlanOf {
  static constexpr Plan P = make_plan(N, ALT);
  static constexpr int E = P.e;
  static constexpr int T = P.t;
  static constexpr int NST = P.nst;
  template <int S> static constexpr int R = P.r[S];
  template <int S> static constexpr int NS() {
    int ns = 1;
    for (int i = 0; i < S; ++i) ns *= P.r[i];
    return ns;
  }
};

// ---------------------------------------------------------------- per-thread twiddles
// The twiddle roots a thread needs: for every stage after the first and every butterfly b of the
// thread, w^k and (radix > 4) w^4k with k = (t + T b) mod NS; the other powers follow by FP64
// products. When they fit in six 16-byte records (the 8- and 16-element plans) they are read once
// when the pass starts, next to the state loads, so no stage waits on a table load behind its
// exchange barrier; forward and inverse transforms of the pass share them (conjugated). Larger
// sets (the 24-element plans of the 3*2^k lengths), and kernels that keep PRE = false because
// their registers are spoken for, read them stage by stage.
// DIRECT32 (FP32 column tiles): every power w^(r k) is read as its own {hi, lo} record at the
// stage, with no FP64 products and nothing preloaded. The FP64 power products held the 16-element
// FP32 column kernels at 160 VGPRs (two waves per SIMD) and the direct reads bring them to 122;
// FP32 rows keep the preloaded form, whose per-stage reads would wait behind block barriers
// (measured: 2048-point rows 36 -> 46 us with direct reads, batched 2048-point columns
// 797 -> 593 us).
template <typename T> struct RootRec;
template <> struct RootRec<double2> { using type = double2; };   // the FP64 root
template <> struct RootRec<float2> { using type = float4; };     // FP32 {hi, lo} of the root

template <int N, int ALT, typename T, bool PRE = true, bool DIRECT32 = false> struct Twiddles {
  using PL = PlanOf<N, ALT>;
  using Rec = typename RootRec<T>::type;
  static constexpr int slots(int S) {
    return S == 0 ? 0 : (PL::E / PL::P.r[S]) * (PL::P.r[S] > 4 ? 2 : 1);
  }
  static constexpr int offset(int S) {
    int o = 0;
    for (int s = 0; s < S; ++s) o += slots(s);
    return o;
  }
  static constexpr int kTotal = offset(PL::NST);
  static constexpr bool kDirect = DIRECT32 && sizeof(T) == 8;
  static constexpr bool kPreload = PRE && kTotal > 0 && kTotal <= 6 && !kDirect;
  Rec r[kPreload ? kTotal : 1];

  __device__ __forceinline__ void load(const T* roots, int t) {
    if constexpr (kPreload) load_stage<1>(reinterpret_cast<const Rec*>(roots), t);
  }
  template <int S> __device__ __forceinline__ void load_stage(const Rec* tab, int t) {
    if constexpr (S < PL::NST) {
      constexpr int R = PL::template R<S>;
      constexpr int NS = PL::template NS<S>();
      constexpr int NB = PL::E / R;
      constexpr int TWS = N / (NS * R);
      constexpr int SL = R > 4 ? 2 : 1;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int k = (t + PL::T * b) % NS;
        r[offset(S) + b * SL] = tab[k * TWS];
        if constexpr (R > 4) r[offset(S) + b * SL + 1] = tab[4 * k * TWS];
      }
      load_stage<S + 1>(tab, t);
    }
  }
};

// ---------------------------------------------------------------- one Stockham transform
// FP32 transforms exchange real and imaginary parts in turn (half the LDS per block); FP64
// transforms exchange whole complex values.
template <typename T> constexpr bool kSplitExchange = sizeof(T) == 8;
template <typename T> constexpr int kExchangeBytes = kSplitExchange<T> ? (int)sizeof(T) / 2 : (int)sizeof(T);
// Row layouts exchange whole complex values in both precisions (a row tile's LDS is small, and
// each split exchange costs two more barriers); column tiles split FP32 exchanges.
template <typename T, typename Addr> constexpr bool kSplitFor = kSplitExchange<T> && Addr::kSplitOk;

// v[q] holds x[t + T*q] on entry and X[t + T*q] on exit. `lds` addresses element p of this
// thread's transform through addr(p), in units of one exchanged value (kExchangeBytes).
// roots[j] = exp(-2πi j/N) (N entries; in FP32 each entry is a pair {hi, lo}: roots[2j] the
// rounded root, roots[2j + 1] its rounding remainder, read as one 16-byte record).
template <int N, int ALT, int DIR, int S, typename T, typename Addr, bool PRE, bool D32>
__device__ __forceinline__ void fft_stage(T* v, int t, T* lds, const Addr& addr, const T* __restrict__ roots,
                                          const Twiddles<N, ALT, T, PRE, D32>& tw) {
  using PL = PlanOf<N, ALT>;
  using TW = Twiddles<N, ALT, T, PRE, D32>;
  constexpr int E = PL::E, TT = PL::T;
  constexpr int R = PL::template R<S>;
  constexpr int NS = PL::template NS<S>();
  constexpr int NB = E / R;             // butterflies per thread
  constexpr bool LAST = (S == PL::NST - 1);
  constexpr int TWS = N / (NS * R);     // root-table stride of this stage's twiddles
  constexpr int SL = R > 4 ? 2 : 1;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + TT * b;
    const int k = j % NS;
    T x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[b + NB * r];
    if constexpr (NS > 1 && TW::kDirect) {
      // w^(r*k) is entry r*k*TWS (< N) of the root table, read as its {hi, lo} record and
      // applied by the compensated product; no FP64 arithmetic and no power products
      const float4* tab = reinterpret_cast<const float4*>(roots);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const float4 q = tab[r * k * TWS];
        const float2 wh{q.x, DIR > 0 ? -q.y : q.y};
        const float2 wl{q.z, DIR > 0 ? -q.w : q.w};
        x[r] = cmul_comp(x[r], wh, wl);
      }
    }
    if constexpr (NS > 1 && !TW::kDirect) {
      // w^k and w^4k (preloaded or read now) in FP64; the other powers by at most three FP64
      // products. FP32 roots are {hi, lo} records taken to FP64, and every power is split back
      // into hi + lo for the compensated product.
      using Rec = typename TW::Rec;
      const Rec* tab = reinterpret_cast<const Rec*>(roots);
      auto as_root = [&](const Rec& q) {
        double2 d;
        if constexpr (sizeof(T) == 16) d = q;
        else d = double2{(double)q.x + (double)q.z, (double)q.y + (double)q.w};
        if constexpr (DIR > 0) d.y = -d.y;
        return d;
      };
      double2 w1, w4;
      if constexpr (TW::kPreload) {
        w1 = as_root(tw.r[TW::offset(S) + b * SL]);
        w4 = R > 4 ? as_root(tw.r[TW::offset(S) + b * SL + (SL - 1)]) : w1;
      } else {
        w1 = as_root(tab[k * TWS]);
        w4 = R > 4 ? as_root(tab[4 * k * TWS]) : w1;
      }
      const double2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 w;
        if (r == 1) w = w1;
        else if (r == 2) w = w2;
        else if (r == 3) w = w3;
        else if (r == 4) w = w4;
        else if (r < 8) w = cmul(w4, r == 5 ? w1 : (r == 6 ? w2 : w3));
        else {
          const double2 w8 = cmul(w4, w4);
          const int m = r & 3;
          const double2 bb = r < 12 ? w8 : cmul(w8, w4);
          w = m == 0 ? bb : cmul(bb, m == 1 ? w1 : (m == 2 ? w2 : w3));
        }
        if constexpr (sizeof(T) == 16) {
          x[r] = cmul(x[r], w);
        } else {
          const float2 wh{(float)w.x, (float)w.y};
          const float2 wl{(float)(w.x - (double)wh.x), (float)(w.y - (double)wh.y)};
          x[r] = cmul_comp(x[r], wh, wl);
        }
      }
    }
    dft<R, DIR>(x);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b + NB * r] = x[r];   // slot b + NB*r holds output r
  }
  if constexpr (!LAST && !kSplitFor<T, Addr>) {
    // complex exchange: output r of butterfly b goes to position (j/NS)*NS*R + j%NS + r*NS,
    // j = t + T*b; element t + T*q comes back
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + TT * b;
      const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[addr(base + r * NS)] = v[b + NB * r];
    }
    exchange_sync<Addr::kWave>();
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = lds[addr(t + TT * q)];
    exchange_sync<Addr::kWave>();
  }
  if constexpr (!LAST && kSplitFor<T, Addr>) {
    // Exchange in two halves, real parts then imaginary parts: a block needs half the LDS, so
    // twice as many blocks stay resident (measured: a gain for FP32 passes, a loss for FP64,
    // whose extra barriers cost more than the occupancy returns).
    using Re = decltype(v[0].x);
    Re* lr = reinterpret_cast<Re*>(lds);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + TT * b;
      const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) lr[addr(base + r * NS)] = v[b + NB * r].x;
    }
    exchange_sync<Addr::kWave>();
#pragma unroll
    for (int q = 0; q < E; ++q) v[q].x = lr[addr(t + TT * q)];   // the outputs' .y stay put
    exchange_sync<Addr::kWave>();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + TT * b;
      const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) lr[addr(base + r * NS)] = v[b + NB * r].y;
    }
    exchange_sync<Addr::kWave>();
#pragma unroll
    for (int q = 0; q < E; ++q) v[q].y = lr[addr(t + TT * q)];
    exchange_sync<Addr::kWave>();
  }
}

template <int N, int ALT, int DIR, int S, typename T, typename Addr, bool PRE, bool D32>
__device__ __forceinline__ void fft_stages(T* v, int t, T* lds, const Addr& addr, const T* __restrict__ roots,
                                           const Twiddles<N, ALT, T, PRE, D32>& tw) {
  if constexpr (S < PlanOf<N, ALT>::NST) {
    fft_stage<N, ALT, DIR, S>(v, t, lds, addr, roots, tw);
    fft_stages<N, ALT, DIR, S + 1>(v, t, lds, addr, roots, tw);
  }
}

// `roots` is the root table of length N (FP64 roots, or FP32 {hi, lo} records); `tw` holds this
// thread's twiddles when the plan preloads them (Twiddles::load at the start of the pass).
template <int N, int ALT, int DIR, typename T, typename Addr, bool PRE, bool D32>
__device__ __forceinline__ void fft(T* v, int t, T* lds, const Addr& addr, const T* __restrict__ roots,
                                    const Twiddles<N, ALT, T, PRE, D32>& tw) {
  fft_stages<N, ALT, DIR, 0>(v, t, lds, addr, roots, tw);
}

}  // namespace qsf
