// runtime.hip — host runtime of libqsf_hip.so and its C ABI (include/qsf_hip.h).
//
// Objects: a context (device + stream), grid sets (one grid, or the 2^D regions of a
// multigrid), grids (device state + tables), and propagators (the SplitPropagator loop of
// propagator.h:51-499 driving fused axis passes). All numerics run in HIP kernels; this file
// only sequences them and builds FP64 tables on the host where bit-exactness is required.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/qsf_hip.h"
#include "physics.hpp"
#include "qsf_launch.hpp"

namespace qsf {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define QSF_HIP(x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) throw Error(QSF_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define QSF_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) throw Error(code, msg); \
  } while (0)

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    if (p) QSF_HIP(hipFree(p));
    p = nullptr;
    QSF_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace qsf

struct qsf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double* scratch = nullptr;   // this device's address of the no-allocation scratch array
  std::string err;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int marks = 0;
  // live per-kernel timing (qsf_prof_*)
  bool prof = false;
  struct Launch { std::string key; hipEvent_t a, b; double bytes; };
  std::vector<Launch> launches;
  std::vector<hipEvent_t> spare;
  struct Entry { std::string key; int64_t n; double ms, bytes; };
  std::vector<Entry> entries;
  // kernel variant per (precision, pass class, axis length, grid size), chosen by the autotuner;
  // the class is tune_class() of the pass
  struct TuneKey {
    int prec, row;
    int64_t n, m;
    bool operator<(const TuneKey& o) const {
      return std::tie(prec, row, n, m) < std::tie(o.prec, o.row, o.n, o.m);
    }
  };
  std::map<TuneKey, std::pair<int, double>> tune;  // -> (variant, best ms)
  int variant_of(int prec, int row, int64_t n, int64_t m) const {
    auto it = tune.find({prec, row, n, m});
    return it == tune.end() ? 0 : it->second.first;
  }
  hipEvent_t event() {
    if (!spare.empty()) { hipEvent_t e = spare.back(); spare.pop_back(); return e; }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
};

// Tuning class of an axis pass: 1 = row mode (the contiguous last axis), 0 = column mode of an
// inner axis, 2 = column mode of the outermost axis. The outermost axis of a 3D grid strides
// n1*n2 elements, every element of a column in another page: wide tiles (more bytes per page
// visit) win there, narrow ones on the inner axes (measured at 512^3 FP64: 1.23 ms at 16
// columns vs 1.34 ms at 8 on axis 0, 0.95 ms at 4 vs 1.02 ms at 8 on axis 1).
int tune_class(const qsf::PassParams& p) { return p.inner == 1 ? 1 : (p.outer == 1 ? 2 : 0); }

// ---- scratch of qsf_average_noalloc ----
// One array per device, held until the process ends: [0, 8) the three sums of the result,
// [8, 8 + 256*3) the second reduction level, then 3 partial sums per block of the observation
// pass. Device code may name it; host code uses only the per-device address from
// hipGetSymbolAddress (the host-side name is a stand-in, not a device address).
namespace qsf {
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr int64_t kScratchResult = 0, kScratchLevel = 8, kScratchPartials = 8 + 256 * 3;
constexpr int64_t kScratchMaxTiles = ((int64_t)(kScratchBytes / sizeof(double)) - kScratchPartials) / 3;
constexpr int kMaxDevices = 64;
}  // namespace qsf
static __device__ __align__(256) double scratch[qsf::kScratchBytes / sizeof(double)];
static std::mutex g_scratch_mu;
static double* g_scratch_addr[qsf::kMaxDevices] = {};

namespace qsf {

// The address of `scratch` on the current device (taken once per device, with it current).
inline double* scratch_address(int device) {
  QSF_REQUIRE(device >= 0 && device < kMaxDevices, QSF_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    QSF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QSF_REQUIRE(p, QSF_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_addr[device] = static_cast<double*>(p);
  }
  return g_scratch_addr[device];
}

// Fold recorded launch events into per-kernel totals (synchronises the stream).
inline void drain_prof(qsf_ctx* c) {
  QSF_HIP(hipStreamSynchronize(c->stream));
  for (auto& l : c->launches) {
    float ms = 0;
    QSF_HIP(hipEventElapsedTime(&ms, l.a, l.b));
    auto it = std::find_if(c->entries.begin(), c->entries.end(), [&](const qsf_ctx::Entry& e) { return e.key == l.key; });
    if (it == c->entries.end()) c->entries.push_back({l.key, 1, ms, l.bytes});
    else { it->n += 1; it->ms += ms; }
    c->spare.push_back(l.a);
    c->spare.push_back(l.b);
  }
  c->launches.clear();
}

// ----------------------------------------------------------------------------------------
// Grid
// ----------------------------------------------------------------------------------------
struct Gather {        // sparse flux points of one border list (see build_gather)
  std::vector<qsf_border_desc> borders;
  DBuf rowptr, pos, rowidx, pts, bvec, values, partials, out;
  int64_t npts = 0, nvals = 0;
};

struct Grid {
  qsf_ctx* ctx;
  qsf_grid_desc d;
  int D, prec;
  size_t cb;
  int64_t n[3], n2[3], m;
  double xmin[3], dx[3], dp[3], dV, dVm, inv_m, eta = 0;
  int freeMask, group;
  bool bounded[3], isMain, many;
  // CoulombInteraction over two or more bounded coordinates (potential.h:40-44) does not
  // separate into per-axis factors: its X factor is applied by launch_radial_x after the
  // per-axis tables (which then carry the coupling and absorber only)
  bool radial = false;
  std::vector<int> fftAxes;
  int pperm[3];
  DBuf psi, aux, partials, red, redscratch, kinw, masktab, invtab, corrtab, m1tab, tabs;
  std::map<int64_t, std::unique_ptr<DBuf>> roots;
  std::map<std::pair<int64_t, int>, std::unique_ptr<DBuf>> pairs;
  std::unique_ptr<Gather> gather;
  // per-substep phase tables requested but not yet built: they point into `tabs`, so they
  // live and die with this grid; one launch builds them before the next pass
  std::vector<AxisTableArgs> pending_tables;
  int rep = QSF_REP_X;
  int slot = 0;
  static constexpr int kSlots = 48;
  static constexpr size_t kTabBytes = 16;   // phase tables are complex FP64 in both precisions

  Grid(qsf_ctx* c, const qsf_grid_desc& desc) : ctx(c), d(desc) {
    D = d.dim;
    QSF_REQUIRE(D >= 1 && D <= 3, QSF_ERR_ARG, "dim must be 1..3");
    prec = d.precision;
    QSF_REQUIRE(prec == QSF_FP64 || prec == QSF_FP32_MIXED || prec == QSF_FP32_STATE, QSF_ERR_ARG, "bad precision");
    cb = prec == QSF_FP64 ? 16 : 8;
    m = 1;
    for (int i = 0; i < 3; ++i) {
      n[i] = i < D ? d.n[i] : 1;
      dx[i] = i < D ? d.dx[i] : 1.0;
      n2[i] = i < D ? n[i] / 2 : 0;
      m *= n[i];
    }
    // CartesianGrid_::init (coords.h:55-93)
    inv_m = 1.0 / m;
    dV = 1.0;
    for (int i = 0; i < D; ++i) {
      dV *= dx[i];
      xmin[i] = -0.5 * (dx[i] * (n[i] - 1));
      dp[i] = 2.0 * phys::kPi / double(n[i]) / dx[i];
    }
    for (int i = D; i < 3; ++i) { xmin[i] = 0; dp[i] = 0; }
    dVm = dV * inv_m;
    if (d.absorber == QSF_ABS_CAP) eta = phys::ipow(3.0 / double(d.nCAP), 4);
    freeMask = d.region_free;
    group = 0;
    for (int i = 0; i < D; ++i) {
      bounded[i] = !(freeMask & (1 << i));
      group += !bounded[i];
    }
    isMain = group == 0;
    many = d.multigrid != 0;
    for (int i = 0; i < D; ++i)
      if (isMain || bounded[i]) fftAxes.push_back(i);
    for (int i = 0; i < 3; ++i) pperm[i] = i;
    if (isMain && D > 1) { pperm[0] = 1; pperm[1] = 0; }  // FFTW_MPI_TRANSPOSED_OUT P layout
    for (int i = 0; i < D; ++i)
      QSF_REQUIRE(axis_size_supported(n[i]), QSF_ERR_UNSUPPORTED,
                  "axis length " + std::to_string(n[i]) + " has no compiled FFT (2^k 8..4096, 3*2^k 12..3072, 5*2^k 10..2560)");
    radial = d.potential == QSF_POT_COULOMB && D - group >= 2;
    if (d.potential == QSF_POT_EBERLY || d.potential == QSF_POT_ECKHARDT_SACHA)
      for (int a = 0; a < D; ++a)
        for (int b = a + 1; b < D; ++b)
          if (bounded[a] && bounded[b])
            QSF_REQUIRE(n[a] == n[b] && dx[a] == dx[b], QSF_ERR_UNSUPPORTED,
                        "pair interactions need equal n and dx on both axes");
    if (many) QSF_REQUIRE(d.absorber == QSF_ABS_CAP || d.absorber == QSF_ABS_COS18, QSF_ERR_ARG,
                          "multigrid regions need an absorber (multigrid.h:19)");

    psi.ensure(m * cb);
    QSF_HIP(hipMemsetAsync(psi.p, 0, m * cb, ctx->stream));
    int64_t maxn = *std::max_element(n, n + 3);
    tabs.ensure((size_t)kSlots * 3 * maxn * kTabBytes);
    red.ensure(64 * sizeof(double));
    redscratch.ensure(256 * 16 * sizeof(double));
    build_static_tables();
  }

  hipStream_t st() const { return ctx->stream; }

  const void* root(int64_t len) {
    auto it = roots.find(len);
    if (it != roots.end()) return it->second->p;
    auto b = std::make_unique<DBuf>();
    // 16 bytes per root: the FP64 root (FP64 and FP32-state arithmetic), or the FP32 {hi, lo}
    b->ensure(len * 16);
    launch_roots(prec == QSF_FP32_MIXED ? QSF_FP32_MIXED : QSF_FP64, b->p, len, st());
    const void* p = b->p;
    roots[len] = std::move(b);
    return p;
  }

  // CAP / cos^(1/8) mask of per-axis nodes (absorbers.h:48-59), host FP64, exact fold order.
  double mask_nodes(int k, const int64_t* nodes) const {
    if (d.absorber == QSF_ABS_COS18) {
      double r = 1.0;
      for (int i = 0; i < k; ++i)
        // cos(π/2 · node/nCAP) as sin(π/2 · (nCAP − node)/nCAP): exactly 0 at the edge node
        r = r * (nodes[i] > 0 ? std::pow(std::sin(0.5 * phys::kPi * double(d.nCAP - nodes[i]) / double(d.nCAP)), 0.125)
                              : 1.0);
      return r;
    }
    bool allneg = true;
    for (int i = 0; i < k; ++i) allneg = allneg && nodes[i] < 0;
    if (allneg) return 1.0;
    double S = 0.0;
    for (int i = k - 1; i >= 0; --i) S = (nodes[i] > 0 ? double(phys::ipow<int64_t>(nodes[i], 4)) : 0.0) + S;
    return std::exp(-double(S) * eta);
  }

  void build_static_tables() {
    // <½p²> weights per axis with the reference's P coordinates (coords.h:132-138; storage axis
    // parameters under the transposed layout)
    int64_t maxn = *std::max_element(n, n + 3);
    kinw.ensure(3 * maxn * sizeof(double));
    std::vector<double> kw(3 * maxn, 0.0);
    for (int a = 0; a < D; ++a) {
      const int s = pperm[a];
      for (int64_t i = 0; i < n[a]; ++i) {
        const double p = dp[s] * (double)(i < n2[s] ? i : i - n[s]);
        kw[a * maxn + i] = 0.5 * p * p;
      }
    }
    QSF_HIP(hipMemcpy(kinw.p, kw.data(), kw.size() * sizeof(double), hipMemcpyHostToDevice));
    if (d.absorber != QSF_ABS_NONE && d.nCAP > 0) {
      // region mask over clamped nodes of the bounded axes (grid.h:582-597)
      const int64_t c = d.nCAP + 1;
      std::vector<double> mt((size_t)std::pow(c, D));
      for (size_t k = 0; k < mt.size(); ++k) {
        int64_t nodes[3] = {0, 0, 0};
        size_t r = k;
        for (int a = D - 1; a >= 0; --a) { nodes[a] = (int64_t)(r % c); r /= c; }
        int64_t passed[3];
        for (int a = 0; a < D; ++a) passed[a] = bounded[a] ? nodes[a] : d.nCAP - n2[a];
        mt[k] = mask_nodes(D, passed);
      }
      masktab.ensure(mt.size() * sizeof(double));
      QSF_HIP(hipMemcpy(masktab.p, mt.data(), mt.size() * sizeof(double), hipMemcpyHostToDevice));
      // single-node inv_mask and the 2D group-1 slice correction (multigrid.h:519-548)
      std::vector<double> inv(c), corr(c), m1(c);
      for (int64_t k = 0; k < c; ++k) {
        int64_t nd = k;
        m1[k] = mask_nodes(1, &nd);
        inv[k] = 1.0 - m1[k];
        corr[k] = 1 - 0.5 * (1.0 * inv[k]);
      }
      invtab.ensure(c * sizeof(double));
      corrtab.ensure(c * sizeof(double));
      m1tab.ensure(c * sizeof(double));
      QSF_HIP(hipMemcpy(invtab.p, inv.data(), c * sizeof(double), hipMemcpyHostToDevice));
      QSF_HIP(hipMemcpy(corrtab.p, corr.data(), c * sizeof(double), hipMemcpyHostToDevice));
      QSF_HIP(hipMemcpy(m1tab.p, m1.data(), c * sizeof(double), hipMemcpyHostToDevice));
    }
  }

  // ---- per-substep tables (ring of slots, each with D per-axis tables) ----
  void* next_slot() {
    int64_t maxn = *std::max_element(n, n + 3);
    void* p = static_cast<char*>(tabs.p) + (size_t)slot * 3 * maxn * kTabBytes;
    slot = (slot + 1) % kSlots;
    return p;
  }
  void* slot_axis(void* s, int a) const {
    int64_t maxn = *std::max_element(n, n + 3);
    return static_cast<char*>(s) + (size_t)a * maxn * kTabBytes;
  }

  bool reduced() const { return d.potential == QSF_POT_EBERLY || d.potential == QSF_POT_ECKHARDT_SACHA; }
  int pair_mask() const {
    if (!reduced()) return 0;
    int mk = 0;
    if (D >= 2 && bounded[0] && bounded[1]) mk |= 1;
    if (D == 3 && bounded[1] && bounded[2]) mk |= 2;
    if (D == 3 && bounded[2] && bounded[0]) mk |= 4;
    return mk;
  }
  const void* pair_table(double delta, int mode) {
    int64_t bits;
    std::memcpy(&bits, &delta, 8);
    auto key = std::make_pair(bits, mode);
    auto it = pairs.find(key);
    if (it != pairs.end()) return it->second->p;
    int a = 0;
    while (a < D && !bounded[a]) ++a;
    PairTableArgs pa{};
    pa.n = n[a];
    pa.kind = d.potential == QSF_POT_EBERLY ? PAIR_DIFF : PAIR_FULL;
    pa.model = d.potential;
    pa.mode = mode;
    pa.delta = delta;
    pa.EE = d.Echarge * d.Echarge;
    pa.Esoft = d.Esoft;
    pa.xmin = xmin[a];
    pa.dx = dx[a];
    auto b = std::make_unique<DBuf>();
    b->ensure((pa.kind == PAIR_DIFF ? pa.n : pa.n * pa.n) * kTabBytes);
    pa.out = b->p;
    launch_pair_table(prec, pa, st());
    const void* p = b->p;
    pairs[key] = std::move(b);
    return p;
  }

  // X factor of one X substep: per-axis tables built now + the static pair table.
  // X factors that do not couple (velocity gauge, no field term in X) depend on (mode, δ)
  // only: built once into tables of their own, so every step (and both X halves of a pass)
  // use the same tables and the pass evaluates the factor once per element
  std::map<std::pair<int64_t, int>, std::pair<std::unique_ptr<DBuf>, XFactor>> xstatic;

  XFactor x_factor(int mode, double delta, const double* C, bool coupled) {
    std::pair<int64_t, int> key{0, mode};
    if (!coupled) {
      std::memcpy(&key.first, &delta, 8);
      auto it = xstatic.find(key);
      if (it != xstatic.end()) return it->second.second;
    }
    void* s = nullptr;
    std::unique_ptr<DBuf> own;
    if (coupled) {
      s = next_slot();
    } else {
      own.reset(new DBuf());
      own->ensure((size_t)3 * *std::max_element(n, n + 3) * kTabBytes);
      s = own->p;
    }
    AxisTableArgs args[3];
    for (int a = 0; a < D; ++a) {
      AxisTableArgs& t = args[a];
      t = AxisTableArgs{};
      t.out = slot_axis(s, a);
      t.n = n[a];
      t.kind = 0;
      t.mode = mode;
      t.delta = delta;
      t.C = coupled ? C[a] : 0.0;
      t.xmin = xmin[a];
      t.dx = dx[a];
      t.use_ne = (d.potential != QSF_POT_NONE) && bounded[a] &&
                 (reduced() || (d.potential == QSF_POT_COULOMB && !radial));
      t.NE = d.Ncharge * d.Echarge;
      t.Nsoft = d.Nsoft;
      t.absorber = d.absorber;
      t.apply_abs = !many;
      t.nCAP = d.nCAP;
      t.eta = eta;
      t.tau = d.absorber_tau;
      t.norm = 1.0;
    }
    for (int a = 0; a < D; ++a) pending_tables.push_back(args[a]);
    XFactor f{};
    for (int a = 0; a < D; ++a) f.axis[a] = args[a].out;
    f.pair_mask = pair_mask();
    if (f.pair_mask) {
      int a = 0;
      while (a < D && !bounded[a]) ++a;
      f.pair_n = n[a];
      f.delta = delta;
      f.mode = mode;
      f.model = d.potential;
      f.EE = d.Echarge * d.Echarge;
      f.Esoft = d.Esoft;
      f.xmin = xmin[a];
      f.dx = dx[a];
      if (d.potential == QSF_POT_EBERLY) {
        // depends on |i-j| only: an n-entry table stays in L1/L2
        f.pair = pair_table(delta, mode);
        f.pair_kind = PAIR_DIFF;
      } else {
        // EckhardSacha ee(x,y) is not a difference kernel: n*n phase table (measured faster
        // than one FP64 sincos per element on MI355X)
        f.pair = pair_table(delta, mode);
        f.pair_kind = PAIR_FULL;
      }
    }
    if (!coupled) xstatic.emplace(key, std::make_pair(std::move(own), f));
    return f;
  }

  // P factor of one P substep; `norm` multiplies in the inverse-FFT normalisation.
  void p_factor(int mode, double delta, const double* C, bool coupled, bool normalise, const void** pf) {
    void* s = next_slot();
    AxisTableArgs args[3];
    for (int a = 0; a < D; ++a) {
      AxisTableArgs& t = args[a];
      t = AxisTableArgs{};
      const int sa = pperm[a];
      t.out = slot_axis(s, a);
      t.n = n[a];
      t.kind = 1;
      t.mode = mode;
      t.delta = delta;
      t.C = coupled ? C[sa] : 0.0;
      t.dp = dp[sa];
      t.pn = n[sa];
      t.pn2 = n2[sa];
      const bool fa = std::find(fftAxes.begin(), fftAxes.end(), a) != fftAxes.end();
      t.norm = (normalise && fa) ? 1.0 / double(n[a]) : 1.0;
    }
    for (int a = 0; a < D; ++a) pending_tables.push_back(args[a]);
    for (int a = 0; a < 3; ++a) pf[a] = a < D ? args[a].out : nullptr;
  }

  PassParams params(int axis, const void* src, void* dst, bool with_roots = true) {
    PassParams p{};
    p.src = src;
    p.dst = dst;
    for (int i = 0; i < 3; ++i) p.n[i] = n[i];
    p.axis = axis;
    p.inner = 1;
    for (int i = axis + 1; i < 3; ++i) p.inner *= n[i];
    p.outer = 1;
    for (int i = 0; i < axis; ++i) p.outer *= n[i];
    p.roots = with_roots ? root(n[axis]) : nullptr;   // root() builds the table on first use
    p.scale = 1.0;
    p.nCAP = d.nCAP;
    return p;
  }

  void obs_params(ObsParams& o, bool epot) {
    o.pot_kind = d.potential == QSF_POT_COULOMB ? 1 : d.potential == QSF_POT_EBERLY ? 2
               : d.potential == QSF_POT_ECKHARDT_SACHA ? 3 : 0;
    o.free_mask = freeMask;
    o.NE = d.Ncharge * d.Echarge;
    o.EE = d.Echarge * d.Echarge;
    o.Nsoft = d.Nsoft;
    o.Esoft = d.Esoft;
    for (int i = 0; i < 3; ++i) { o.xmin[i] = xmin[i]; o.dx[i] = dx[i]; }
    o.want_epot = epot;
  }

  void build_tables() {
    if (!pending_tables.empty()) launch_axis_tables(prec, pending_tables.data(), (int)pending_tables.size(), st());
    pending_tables.clear();
  }

  // `single_out`: where the three block sums go when the pass is one tile (they are then the
  // final sums and need no reduction launch)
  int run(PassParams& p, int force_variant = -1, double* single_out = nullptr) {
    build_tables();
    const int var = force_variant >= 0 ? force_variant : ctx->variant_of(prec, tune_class(p), n[p.axis], m);
    if (p.flags & F_OBS) {
      int64_t tiles = axis_pass_tiles(prec, p, var);
      partials.ensure(tiles * 3 * sizeof(double));
      if (tiles == 1 && single_out) p.obs.partials = single_out;
      if (!p.obs.partials) p.obs.partials = partials.as<double>();
    }
    hipEvent_t ea = nullptr, eb = nullptr;
    if (ctx->prof) {
      ea = ctx->event();
      eb = ctx->event();
      QSF_HIP(hipEventRecord(ea, st()));
    }
    int k = launch_axis_pass(prec, p, st(), var);
    QSF_REQUIRE(k >= 0, QSF_ERR_UNSUPPORTED, "no FFT kernel for this axis length");
    QSF_HIP(hipGetLastError());
    if (ctx->prof) {
      QSF_HIP(hipEventRecord(eb, st()));
      const bool row = p.inner == 1;
      std::string key = std::string("pass_") + (prec == QSF_FP64 ? "f64" : (prec == QSF_FP32_STATE ? "s32" : "f32")) +
                        (row ? "_row_n" : "_col_n") +
                        std::to_string(n[p.axis]) + "_v" + std::to_string(var);
      const double bytes = (double)m * cb * ((p.flags & F_NOSTORE) ? 1.0 : 2.0);
      ctx->launches.push_back({key, ea, eb, bytes});
      if (ctx->launches.size() >= 4096) drain_prof(ctx);
    }
    return k;
  }

  // ---- autotuner: time every compiled variant of the passes this grid's step uses ----
  void autotune(int reps) {
    aux.ensure(m * cb);
    QSF_HIP(hipMemcpyAsync(aux.p, psi.p, m * cb, hipMemcpyDeviceToDevice, st()));
    const double C[3] = {0.01, 0.01, 0.01};
    // the passes of this grid's fused step; passes that share a tuning key (same class and
    // length) are timed together per variant
    struct Cand { int axis; uint32_t flags; };
    std::vector<Cand> cands;
    if (D == 3 && fftAxes.size() == 3) {
      for (int a = D - 1; a >= 0; --a) cands.push_back({a, F_FWD | F_P | F_INV2 | (a == D - 1 ? F_XA | F_XB : 0u)});
    } else if (fftAxes.size() >= 2) {
      cands.push_back({D - 1, F_INV1 | F_XA | F_XB | F_FWD});
      cands.push_back({0, F_FWD | F_P | F_INV2});
      for (int a = 1; a <= D - 2; ++a) cands.push_back({a, F_FWD});
    } else if (fftAxes.size() == 1) {
      cands.push_back({fftAxes[0], F_XA | F_XB | F_FWD | F_P | F_INV2});
    }
    hipEvent_t e0, e1;
    QSF_HIP(hipEventCreate(&e0));
    QSF_HIP(hipEventCreate(&e1));
    std::map<std::pair<int, int64_t>, std::vector<Cand>> keyed;   // (row, n) -> passes
    for (const Cand& c : cands) keyed[{tune_class(params(c.axis, aux.p, aux.p)), n[c.axis]}].push_back(c);
    for (const auto& kv : keyed) {
      const int row = kv.first.first;
      const int64_t len = kv.first.second;
      const int nv = axis_variants(len);
      int best = 0;
      double best_ms = 1e300;
      for (int v = 0; v < nv; ++v) {
        double per = 0.0;
        for (const Cand& c : kv.second) {
          PassParams p = params(c.axis, aux.p, aux.p);
          p.flags = c.flags;
          if (c.flags & F_XA) p.xa = x_factor(QSF_MODE_RE, 0.01, C, true);
          if (c.flags & F_XB) p.xb = x_factor(QSF_MODE_RE, 0.01, C, true);
          if (c.flags & F_P) p_factor(QSF_MODE_RE, 0.01, C, true, true, p.pf);
          run(p, v);  // warm-up (and table build) outside the timed region
          QSF_HIP(hipEventRecord(e0, st()));
          for (int r = 0; r < reps; ++r) run(p, v);
          QSF_HIP(hipEventRecord(e1, st()));
          QSF_HIP(hipEventSynchronize(e1));
          float ms = 0;
          QSF_HIP(hipEventElapsedTime(&ms, e0, e1));
          per += ms / reps;
        }
        if (per < best_ms) { best_ms = per; best = v; }
      }
      ctx->tune[{prec, row, len, m}] = {best, best_ms};
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }

  // ---- standalone members ----
  void fourier(int to) {
    QSF_REQUIRE(to == QSF_REP_X || to == QSF_REP_P, QSF_ERR_ARG, "fourier: rep must be X or P");
    if (to == rep) return;
    if (to == QSF_REP_P) {
      for (int a : fftAxes) {
        PassParams p = params(a, psi.p, psi.p);
        p.flags = F_FWD;
        run(p);
      }
    } else {
      double s = 1.0;
      if (isMain) s = inv_m;  // normalizeAfterTwoFFT grid.h:363-367
      else {
        double pim = 1.0;  // partial_inv_m multigrid.h:148-150
        for (int a = 0; a < D; ++a)
          if (bounded[a]) pim *= 1.0 / double(n[a]);
        s = fftAxes.empty() ? 1.0 : pim;
      }
      bool first = true;
      for (int a : fftAxes) {
        PassParams p = params(a, psi.p, psi.p);
        p.flags = F_INV1 | (first ? F_SCALE : 0u);
        p.scale = s;
        first = false;
        run(p);
      }
    }
    rep = to;
  }

  void evolve(int mode, int R, double delta, const double* C, int couplesIn) {
    const double z[3] = {0, 0, 0};
    const double* c = C ? C : z;
    PassParams p = params(D - 1, psi.p, psi.p);
    if (R == QSF_REP_X) {
      p.xa = x_factor(mode, delta, c, couplesIn == QSF_REP_X);
      p.flags = F_XA;
    } else {
      p_factor(mode, delta, c, couplesIn == QSF_REP_P, false, p.pf);
      p.flags = F_P;
    }
    run(p);
    if (R == QSF_REP_X && radial) {
      RadialArgs r{};
      r.psi = psi.p;
      r.precision = prec;
      for (int a = 0; a < 3; ++a) {
        r.n[a] = n[a];
        r.xmin[a] = xmin[a];
        r.dx[a] = dx[a];
      }
      r.bounded_mask = 0;
      for (int a = 0; a < D; ++a)
        if (bounded[a]) r.bounded_mask |= 1 << a;
      r.NE = d.Ncharge * d.Echarge;
      r.Nsoft = d.Nsoft;
      r.delta = delta;
      r.mode = mode;
      launch_radial_x(r, st());
    }
  }

  // Partial sums [norm, epot, ekin] of the current state, reduced on the device.
  void observe(bool epot, bool ekin, double* host3) {
    PassParams p = params(D - 1, psi.p, psi.p);
    p.flags = F_OBS | F_NOSTORE;
    obs_params(p.obs, epot);
    if (ekin) {
      int64_t maxn = *std::max_element(n, n + 3);
      for (int a = 0; a < D; ++a) p.obs.kin_w[a] = kinw.as<double>() + a * maxn;
    }
    int tiles = run(p);
    launch_reduce_partials(partials.as<double>(), tiles, 3, red.as<double>(), redscratch.as<double>(), st());
    QSF_HIP(hipMemcpyAsync(host3, red.p, 3 * sizeof(double), hipMemcpyDeviceToHost, st()));
    QSF_HIP(hipStreamSynchronize(st()));
  }

  double average(int R, int op) {
    double r[3];
    observe(op == 1, op == 2, r);
    const double vol = R == QSF_REP_X ? dV : dVm;  // vol<R> coords.h:48-53
    return r[op] * vol;
  }

  // average<R, op> with the block partials and the reduction in `scr` (the per-device scratch):
  // the same observation pass (same kernel variant, so the same blocks) and the same fixed-order
  // reduction as average(), with no table build and no buffer of the grid touched or grown.
  double average_into(int R, int op, double* scr) {
    PassParams p = params(D - 1, psi.p, psi.p, false);
    p.flags = F_OBS | F_NOSTORE;
    obs_params(p.obs, op == 1);
    if (op == 2) {
      int64_t maxn = *std::max_element(n, n + 3);
      for (int a = 0; a < D; ++a) p.obs.kin_w[a] = kinw.as<double>() + a * maxn;
    }
    const int var = ctx->variant_of(prec, tune_class(p), n[p.axis], m);
    const int64_t tiles = axis_pass_tiles(prec, p, var);
    QSF_REQUIRE(tiles >= 1 && tiles <= kScratchMaxTiles, QSF_ERR_ARG,
                "the observation pass of this grid has " + std::to_string(tiles) + " blocks; the "
                "no-allocation scratch holds at most " + std::to_string(kScratchMaxTiles));
    double* res = scr + kScratchResult;
    p.obs.partials = scr + kScratchPartials;
    QSF_REQUIRE(launch_axis_pass(prec, p, st(), var) >= 0, QSF_ERR_UNSUPPORTED, "no FFT kernel for this axis length");
    QSF_HIP(hipGetLastError());
    launch_reduce_partials(p.obs.partials, tiles, 3, res, scr + kScratchLevel, st());
    QSF_HIP(hipGetLastError());
    double r[3];
    QSF_HIP(hipMemcpyAsync(r, res, 3 * sizeof(double), hipMemcpyDeviceToHost, st()));
    QSF_HIP(hipStreamSynchronize(st()));
    return r[op] * (R == QSF_REP_X ? dV : dVm);   // vol<R> coords.h:48-53
  }

  void normalize() {
    double r[3];
    observe(false, false, r);
    const double res = r[0] * dV;
    launch_scale(prec, psi.p, m, std::sqrt(1.0 / res), st());
  }

  void mask_region() {
    QSF_REQUIRE(masktab.p, QSF_ERR_STATE, "maskRegion needs an absorber");
    PassParams p = params(D - 1, psi.p, psi.p);
    p.flags = F_MASK;
    fill_mask(p);
    run(p);
  }
  void fill_mask(PassParams& p) {
    p.mask_tab = masktab.as<double>();
    const int64_t c = d.nCAP + 1;
    int64_t s = 1;
    for (int a = D - 1; a >= 0; --a) {
      p.mask_stride[a] = bounded[a] ? s : 0;
      s *= c;
    }
    p.nCAP = d.nCAP;
  }

  // ---- flux: sparse gather of boundary points (grid.h:467-573, borders.h) ----
  void build_gather(const qsf_border_desc* b, int nb) {
    QSF_REQUIRE(nb >= 1 && nb <= 16, QSF_ERR_ARG, "1..16 flux borders");
    if (gather && (int)gather->borders.size() == nb &&
        std::equal(b, b + nb, gather->borders.begin(), [](const qsf_border_desc& x, const qsf_border_desc& y) {
          return x.kind == y.kind && x.size == y.size;
        }))
      return;
    auto g = std::make_unique<Gather>();
    g->borders.assign(b, b + nb);
    // candidate planes: a component along axis a can be non-zero only where the centred index
    // of axis a is one of the border positions {at-1, -at}
    std::vector<int64_t> at;
    for (int k = 0; k < nb; ++k) {
      if (b[k].kind == QSF_BORDER_BOX) at.push_back(b[k].size);
      else for (long c : {phys::bord::kNS, phys::bord::kSD, phys::bord::kDT, phys::bord::kCAP}) at.push_back(c);
    }
    std::sort(at.begin(), at.end());
    at.erase(std::unique(at.begin(), at.end()), at.end());
    std::vector<int64_t> cand;
    int64_t stride[3];
    for (int a = 0; a < 3; ++a) {
      stride[a] = 1;
      for (int e = a + 1; e < 3; ++e) stride[a] *= n[e];
    }
    for (int a = 0; a < D; ++a)
      for (int64_t v : at)
        for (int64_t ci : {v - 1, -v}) {
          const int64_t ia = ci + n2[a];
          if (ia < 0 || ia >= n[a]) continue;
          const int64_t o1 = a == 0 ? 1 : 0;  // enumerate the other axes
          (void)o1;
          for (int64_t i0 = 0; i0 < (a == 0 ? 1 : n[0]); ++i0)
            for (int64_t i1 = 0; i1 < (a == 1 ? 1 : n[1]); ++i1)
              for (int64_t i2 = 0; i2 < (a == 2 ? 1 : n[2]); ++i2) {
                int64_t id[3] = {i0, i1, i2};
                id[a] = ia;
                cand.push_back(id[0] * stride[0] + id[1] * stride[1] + id[2] * stride[2]);
              }
        }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    struct P { int64_t c; bool need[3]; std::vector<double> bv; };
    std::vector<P> pts;
    for (int64_t c : cand) {
      int64_t id[3] = {c / stride[0], (c / stride[1]) % n[1], c % n[2]};
      P q;
      q.c = c;
      q.need[0] = q.need[1] = q.need[2] = false;
      q.bv.assign(nb * D, 0.0);
      bool any = false;
      for (int k = 0; k < nb; ++k) {
        auto v = phys::bord::border(b[k].kind, b[k].size, D, id[0] - n2[0], D > 1 ? id[1] - n2[1] : 0,
                                    D > 2 ? id[2] - n2[2] : 0);
        for (int a = 0; a < D; ++a) {
          q.bv[k * D + a] = v.c[a];
          if (v.c[a] != 0.0) { q.need[a] = true; any = true; }
        }
      }
      if (any) pts.push_back(std::move(q));
    }
    // needed values: the point and its ±stride neighbours (strict linear rule grid.h:470-472)
    std::vector<int64_t> need;
    for (auto& q : pts) {
      need.push_back(q.c);
      for (int a = 0; a < D; ++a)
        if (q.need[a] && q.c - stride[a] > 0 && q.c + stride[a] < m) {
          need.push_back(q.c + stride[a]);
          need.push_back(q.c - stride[a]);
        }
    }
    std::sort(need.begin(), need.end());
    need.erase(std::unique(need.begin(), need.end()), need.end());
    auto slot_of = [&](int64_t c) { return (int32_t)(std::lower_bound(need.begin(), need.end(), c) - need.begin()); };
    const int64_t rowlen = n[D - 1], rows = m / rowlen;
    std::vector<int32_t> rowptr(rows + 1, 0), pos(need.size()), rowidx(need.size());
    for (size_t k = 0; k < need.size(); ++k) {
      rowptr[need[k] / rowlen + 1]++;
      pos[k] = (int32_t)(need[k] % rowlen);
      rowidx[k] = (int32_t)(need[k] / rowlen);
    }
    for (int64_t r = 0; r < rows; ++r) rowptr[r + 1] += rowptr[r];
    std::vector<int32_t> pt(pts.size() * (1 + 2 * D));
    std::vector<double> bv(pts.size() * nb * D);
    for (size_t k = 0; k < pts.size(); ++k) {
      int32_t* q = &pt[k * (1 + 2 * D)];
      q[0] = slot_of(pts[k].c);
      for (int a = 0; a < D; ++a) {
        const bool ok = pts[k].need[a] && pts[k].c - stride[a] > 0 && pts[k].c + stride[a] < m;
        q[1 + a] = ok ? slot_of(pts[k].c + stride[a]) : -1;
        q[1 + D + a] = ok ? slot_of(pts[k].c - stride[a]) : -1;
      }
      std::copy(pts[k].bv.begin(), pts[k].bv.end(), bv.begin() + k * nb * D);
    }
    g->npts = (int64_t)pts.size();
    g->nvals = (int64_t)need.size();
    auto up = [&](DBuf& dst, const void* src, size_t bytes) {
      dst.ensure(std::max<size_t>(bytes, 16));
      if (bytes) QSF_HIP(hipMemcpy(dst.p, src, bytes, hipMemcpyHostToDevice));
    };
    up(g->rowptr, rowptr.data(), rowptr.size() * 4);
    up(g->pos, pos.data(), pos.size() * 4);
    up(g->rowidx, rowidx.data(), rowidx.size() * 4);
    up(g->pts, pt.data(), pt.size() * 4);
    up(g->bvec, bv.data(), bv.size() * 8);
    g->values.ensure(std::max<int64_t>(g->nvals, 1) * cb);
    g->partials.ensure(1024 * 16 * sizeof(double));
    g->out.ensure(16 * sizeof(double));
    gather = std::move(g);
  }

  // The stencil values of the X state ahead of the pass that observes it, with the pending X
  // factor (or none) applied at each point: what that pass would gather, read from HBM instead
  // of from the pass's registers (for passes whose input is the X state).
  void gather_ahead(const XFactor* xa) {
    GatherXArgs a{};
    a.psi = psi.p;
    a.out = gather->values.p;
    a.row = gather->rowidx.as<int32_t>();
    a.pos = gather->pos.as<int32_t>();
    a.nvals = gather->nvals;
    for (int k = 0; k < 3; ++k) a.n[k] = n[k];
    a.dim = D;
    a.has_xa = xa ? 1 : 0;
    if (xa) a.xa = *xa;
    launch_gather_x(prec, a, st());
  }

  void set_gather(ObsParams& o) {
    o.gather_rowptr = gather->rowptr.as<int32_t>();
    o.gather_pos = gather->pos.as<int32_t>();
    o.gather_out = gather->values.p;
  }

  void flux_from_gather(const double* vgC, double* dev_out, double* accum, double w) {
    FluxArgs f{};
    f.gathered = gather->values.p;
    f.precision = prec;
    f.pts = gather->pts.as<int32_t>();
    f.bvec = gather->bvec.as<double>();
    f.npts = gather->npts;
    f.D = D;
    f.nb = (int)gather->borders.size();
    for (int a = 0; a < 3; ++a) {
      f.inv_2dx[a] = a < D ? (1.0 / dx[a]) / 2.0 : 0.0;  // coords.h:77-78
      f.vgC[a] = vgC ? vgC[a] : 0.0;
    }
    f.scale = dV / dx[0];
    f.partials = gather->partials.as<double>();
    f.out = dev_out ? dev_out : gather->out.as<double>();
    f.accum = accum;
    f.accum_weight = w;
    launch_flux(f, st());
  }

  // flux of each border family into out (host); with accum, also accum[b] += flux * w on the
  // device (the propagator's FP64 flux time integral)
  void flux(const double* vgC, const qsf_border_desc* b, int nb, double* out, double* accum = nullptr,
            double w = 0.0) {
    QSF_REQUIRE(rep == QSF_REP_X, QSF_ERR_STATE, "flux needs the X representation");
    build_gather(b, nb);
    PassParams p = params(D - 1, psi.p, psi.p);
    p.flags = F_OBS | F_NOSTORE;
    obs_params(p.obs, false);
    set_gather(p.obs);
    run(p);
    flux_from_gather(vgC, nullptr, accum, w);
    QSF_HIP(hipMemcpyAsync(out, gather->out.p, nb * sizeof(double), hipMemcpyDeviceToHost, st()));
    QSF_HIP(hipStreamSynchronize(st()));
  }

  // ---- host transfer: FP64 interleaved in the reference layout ----
  void upload(const double* h, int64_t count, int r) {
    QSF_REQUIRE(count == m, QSF_ERR_ARG, "upload count must equal the grid size");
    std::vector<double> tmp;
    const double* srcp = h;
    if (r == QSF_REP_P && isMain && D > 1) {  // (n1, n0, n2) -> internal (n0, n1, n2)
      tmp.resize(2 * m);
      for (int64_t j = 0; j < n[1]; ++j)
        for (int64_t i = 0; i < n[0]; ++i)
          for (int64_t k = 0; k < n[2]; ++k) {
            const int64_t a = ((j * n[0] + i) * n[2] + k), b2 = ((i * n[1] + j) * n[2] + k);
            tmp[2 * b2] = h[2 * a];
            tmp[2 * b2 + 1] = h[2 * a + 1];
          }
      srcp = tmp.data();
    }
    if (prec == QSF_FP64) {
      QSF_HIP(hipMemcpyAsync(psi.p, srcp, m * 16, hipMemcpyHostToDevice, st()));
    } else {
      aux.ensure(m * 16);
      QSF_HIP(hipMemcpyAsync(aux.p, srcp, m * 16, hipMemcpyHostToDevice, st()));
      launch_copy_convert(QSF_FP64, aux.p, QSF_FP32_MIXED, psi.p, m, st());
    }
    QSF_HIP(hipStreamSynchronize(st()));
    rep = r;
  }

  void download(double* h, int64_t count) {
    QSF_REQUIRE(count == m, QSF_ERR_ARG, "download count must equal the grid size");
    std::vector<double> tmp;
    double* dstp = (rep == QSF_REP_P && isMain && D > 1) ? (tmp.resize(2 * m), tmp.data()) : h;
    if (prec == QSF_FP64) {
      QSF_HIP(hipMemcpyAsync(dstp, psi.p, m * 16, hipMemcpyDeviceToHost, st()));
    } else {
      aux.ensure(m * 16);
      launch_copy_convert(QSF_FP32_MIXED, psi.p, QSF_FP64, aux.p, m, st());
      QSF_HIP(hipMemcpyAsync(dstp, aux.p, m * 16, hipMemcpyDeviceToHost, st()));
    }
    QSF_HIP(hipStreamSynchronize(st()));
    if (dstp != h)
      for (int64_t i = 0; i < n[0]; ++i)
        for (int64_t j = 0; j < n[1]; ++j)
          for (int64_t k = 0; k < n[2]; ++k) {
            const int64_t a = ((i * n[1] + j) * n[2] + k), b2 = ((j * n[0] + i) * n[2] + k);
            h[2 * b2] = tmp[2 * a];
            h[2 * b2 + 1] = tmp[2 * a + 1];
          }
  }
};

// ----------------------------------------------------------------------------------------
// Grid sets and the multigrid transfer (multigrid.h:617-657, mpi_logic.h:103-263)
// ----------------------------------------------------------------------------------------
struct Source { int region; int fc; };
struct Prop;

struct Set {
  qsf_ctx* ctx;
  std::vector<std::unique_ptr<Grid>> g;
  // propagators created on this set; the last one created drives its grids (entry points that
  // read or change a grid settle it first). Owned by their callers; detached when the set goes.
  std::vector<Prop*> props;
  Prop* active = nullptr;
  ~Set();
  Set(const Set&) = delete;
  Set& operator=(const Set&) = delete;
  std::vector<std::vector<Source>> sources;
  int D, groupCount = 1;
  bool multi;
  DBuf slice;

  Set(qsf_ctx* c, const qsf_grid_desc& d) : ctx(c) {
    D = d.dim;
    multi = d.multigrid != 0;
    if (!multi) {
      g.emplace_back(new Grid(c, d));
      return;
    }
    QSF_REQUIRE(D == 2 || D == 3, QSF_ERR_UNSUPPORTED, "multigrid regions need 2 or 3 dimensions");
    QSF_REQUIRE(d.precision != QSF_FP32_STATE, QSF_ERR_UNSUPPORTED,
                "multigrid regions run in FP64 or FP32-mixed arithmetic");
    if (D == 3) {
      // the transfer boxes (multigrid.h:188, 297-309) tile axes 1 and 2 in steps of nCAP and
      // are counted on axis 0: a cube whose side nCAP divides
      QSF_REQUIRE(d.n[0] == d.n[1] && d.n[1] == d.n[2] && d.nCAP >= 1 && d.n[0] % d.nCAP == 0 &&
                      2 * d.nCAP <= d.n[0],
                  QSF_ERR_UNSUPPORTED, "3D multigrid needs a cubic grid whose side is a multiple of 2*nCAP or more");
      // partial() of the YZ region keeps z (potential.h:72-77), a free coordinate there
      QSF_REQUIRE(d.potential != QSF_POT_COULOMB, QSF_ERR_UNSUPPORTED,
                  "3D multigrid takes the reduced-dimension interactions");
    }
    // MultiRegionsReduced region order (mpi_logic.h:249-263)
    static const int free2[4] = {QSF_AXIS_NO, QSF_AXIS_X, QSF_AXIS_Y, QSF_AXIS_XY};
    static const int free3[8] = {QSF_AXIS_NO, QSF_AXIS_X,  QSF_AXIS_Y,  QSF_AXIS_Z,
                                 QSF_AXIS_XY, QSF_AXIS_XZ, QSF_AXIS_YZ, QSF_AXIS_XYZ};
    const int* fr = D == 2 ? free2 : free3;
    const int R = D == 2 ? 4 : 8;
    for (int r = 0; r < R; ++r) {
      qsf_grid_desc dd = d;
      dd.region_free = fr[r];
      g.emplace_back(new Grid(c, dd));
    }
    groupCount = D + 1;
    sources.resize(R);
    for (int r = 0; r < R; ++r)
      for (int i = 0; i < R; ++i)
        if (__builtin_popcount(fr[r]) - __builtin_popcount(fr[i]) == 1) {
          const int x = fr[r] ^ fr[i];
          if (x == QSF_AXIS_X || x == QSF_AXIS_Y || x == QSF_AXIS_Z) sources[r].push_back({i, x});
        }
  }

  // Edge buffers of the fused transfer: main region edge rows/columns, region X edge columns
  // and region Y edge rows (the sources of the XY region).
  DBuf er_main, ec_main, ec_x, er_y;

  // the pending phase tables of every region grid in one batched launch (24 per launch)
  void build_tables() {
    std::vector<AxisTableArgs> all;
    for (auto& gp : g) {
      all.insert(all.end(), gp->pending_tables.begin(), gp->pending_tables.end());
      gp->pending_tables.clear();
    }
    if (!all.empty()) launch_axis_tables(g[0]->prec, all.data(), (int)all.size(), ctx->stream);
  }
  void ensure_edges() {
    const Grid& m0 = *g[0];
    const size_t bytes_r = 2 * m0.d.nCAP * m0.n[1] * m0.cb, bytes_c = m0.n[0] * 2 * m0.d.nCAP * m0.cb;
    er_main.ensure(bytes_r);
    ec_main.ensure(bytes_c);
    ec_x.ensure(bytes_c);
    er_y.ensure(bytes_r);
  }

  // Fused transfer() + maskRegion() of the receiving regions, reading the main region's edges
  // written by its last observation pass. `xa[r]` is region r's pending X factor (or nullptr).
  void receive_fused(const XFactor* const* xa) {
    struct R { int region, axis; const DBuf* in; int in_layout; DBuf* out; int out_layout; bool mask; };
    const R plan[4] = {
        {1, 0, &er_main, 0, &ec_x, 1, true},    // X  <- NO along X
        {2, 1, &ec_main, 1, &er_y, 0, true},    // Y  <- NO along Y
        {3, 1, &ec_x, 1, nullptr, 0, false},    // XY <- X  along Y (source 0)
        {3, 0, &er_y, 0, nullptr, 0, false},    // XY <- Y  along X (source 1)
    };
    bool xa_used[4] = {false, false, false, false};
    for (const R& r : plan) {
      Grid& rg = *g[r.region];
      rg.build_tables();
      TransferParams tp{};
      tp.dst = rg.psi.p;
      tp.n0 = rg.n[0];
      tp.n1 = rg.n[1];
      tp.axis = r.axis;
      tp.roots = rg.root(rg.n[r.axis]);
      tp.nCAP = rg.d.nCAP;
      tp.edge_in = r.in->p;
      tp.edge_in_layout = r.in_layout;
      tp.edge_out = r.out ? r.out->p : nullptr;
      tp.edge_out_layout = r.out_layout;
      tp.inv_tab = rg.invtab.as<double>();
      tp.corr_tab = rg.group == 1 ? rg.corrtab.as<double>() : nullptr;
      if (xa[r.region] && !xa_used[r.region]) {
        tp.has_xa = 1;
        tp.xa = *xa[r.region];
        xa_used[r.region] = true;
      }
      if (r.mask) {
        const int64_t c = rg.d.nCAP + 1;
        const int64_t st1 = rg.bounded[1] ? 1 : 0, st0 = rg.bounded[0] ? c : 0;
        tp.mask_tab = rg.masktab.as<double>();
        tp.mask_stride_axis = r.axis == 0 ? st0 : st1;
        tp.mask_stride_other = r.axis == 0 ? st1 : st0;
      }
      const int k = launch_transfer(rg.prec, tp, ctx->stream);
      QSF_REQUIRE(k >= 0, QSF_ERR_UNSUPPORTED, "no transfer kernel for this axis length");
      QSF_HIP(hipGetLastError());
    }
  }

  // transfer() in group order; every source must be in X along the freed axis.
  void transfer() {
    for (int gI = 1; gI < groupCount; ++gI)
      for (int sI = 0; sI < gI; ++sI)
        for (size_t r = 0; r < g.size(); ++r) {
          Grid& rg = *g[r];
          if (rg.group != gI) continue;
          const Source& s = sources[r][sI];
          Grid& src = *g[s.region];
          slice.ensure(rg.m * rg.cb);
          SliceArgs a{};
          a.src = src.psi.p;
          a.slice = slice.p;
          a.n0 = rg.n[0];
          a.n1 = rg.n[1];
          a.n2 = rg.n[2];
          a.dirFree = s.fc == QSF_AXIS_X ? 0 : (s.fc == QSF_AXIS_Y ? 1 : 2);
          a.nCAP = rg.d.nCAP;
          a.inv_tab = rg.invtab.as<double>();
          if (D == 2) {
            a.corr_tab = rg.group == 1 ? rg.corrtab.as<double>() : nullptr;
            launch_slice_2d(rg.prec, a, ctx->stream);
          } else {
            a.m1_tab = rg.group == 1 ? rg.m1tab.as<double>() : nullptr;
            launch_slice_3d(rg.prec, a, ctx->stream);
          }
          PassParams p = rg.params(a.dirFree, slice.p, slice.p);
          p.flags = F_FWD;  // transf_slice: FFTW_FORWARD, unnormalised
          rg.run(p);
          launch_add(rg.prec, rg.psi.p, slice.p, rg.m, ctx->stream);
        }
  }
  void post_compute() {
    if (!multi) return;
    transfer();
    for (auto& gr : g)
      if (gr->group != groupCount - 1) gr->mask_region();
  }
};

// ----------------------------------------------------------------------------------------
// Propagator
// ----------------------------------------------------------------------------------------
struct SubOp {
  int rep;
  double delta;
  double C[3];
};

struct Chain {
  std::vector<int> reps;
  std::vector<double> mults;
};

// MultiProductSplit<Split3Base<X,P,X>,order>::Chain<c> (splitting.h:18-27, 43-58, 63-87)
static Chain chain_of(int c) {
  Chain ch;
  const double base[3] = {0.5, 1.0, 0.5};
  ch.reps.push_back(QSF_REP_X);
  for (int k = 0; k < c; ++k) { ch.reps.push_back(QSF_REP_P); ch.reps.push_back(QSF_REP_X); }
  ch.mults.assign(ch.reps.size(), 0.0);
  int i = 0;
  for (int k = 0; k < c; ++k) {
    const double v = 1.0 / c;
    for (int j = 0; j < 3; ++j) ch.mults[i + j] += base[j] * v;
    i += 2;
  }
  return ch;
}

struct PendingX {
  bool on = false;
  XFactor f{};
};

struct Prop {
  Set* set;            // nullptr once the set was destroyed (the propagator is then inert)
  qsf_ctx* ctx;        // outlives every set and propagator created on it
  qsf_prop_desc pd;
  phys::Coupling coup;
  std::vector<Chain> chains;
  int64_t max_steps = 0, step = 0;
  double dt, timer = 0, timer_copy = 0, state_accuracy = 0;
  double dif_energy = 1e300, last_change = 0;
  int exec = 0;
  int width;
  // host rows; rows produced by fused passes are completed from the device on sync
  std::vector<double> rows;
  struct DevRow { int row; bool ekin; };
  std::vector<DevRow> pending_rows;
  DBuf devrec, fluxint;
  int64_t devrec_cap = 0, devrec_used = 0;
  // fused pipeline state of each grid
  std::vector<bool> deferred;     // state is "after a P pass": pending inverse (main) + pending X
  std::vector<PendingX> pend;
  bool pend_obs = false;
  bool pend_post = false;   // multigrid transfer + masks of the last step boundary not yet run
  double pend_obs_C[3] = {0, 0, 0};
  int pend_row = -1;
  bool pend_ekin = false;

  Prop(Set* s, const qsf_prop_desc& p, const qsf_field_desc* f, int nf)
      : set(s), ctx(s->ctx), pd(p), coup(p.gauge, f, nf) {
    QSF_REQUIRE(p.order >= 1 && p.order <= 8, QSF_ERR_ARG, "order must be 1..8");
    QSF_REQUIRE(p.nborders >= 0 && p.nborders <= 16, QSF_ERR_ARG, "0..16 borders");
    dt = p.dt;
    for (int c = 1; c <= p.order; ++c) chains.push_back(chain_of(c));
    if (p.mode == QSF_MODE_RE) {
      max_steps = (int64_t)(coup.max_pulse_duration() / dt + 1);  // propagator.h:153
      state_accuracy = 0;
    } else {
      max_steps = p.max_steps;
      state_accuracy = p.state_accuracy;
    }
    width = 7 + p.nborders;
    deferred.assign(s->g.size(), false);
    pend.resize(s->g.size());
    fluxint.ensure(16 * sizeof(double));
    QSF_HIP(hipMemsetAsync(fluxint.p, 0, 16 * sizeof(double), s->ctx->stream));
    // observables that change the state or steer the loop every step run unfused, and so does
    // a non-separable potential (its X factor is a pass of its own)
    if (p.mode == QSF_MODE_IM || (p.outputs & (QSF_OUT_NORMALIZE | QSF_OUT_EDIFF)) || radial_set() ||
        (s->multi && s->D == 3))
      exec = 1;
    if (p.nborders > 0) main().build_gather(p.borders, p.nborders);
    s->props.push_back(this);
    s->active = this;
  }

  Grid& main() { return *set->g[0]; }
  hipStream_t st() { return ctx->stream; }
  bool radial_set() const {
    for (const auto& g : set->g)
      if (g->radial) return true;
    return false;
  }
  int couples() const { return coup.rep(); }

  // ---- host side of makeStep: chain bookkeeping, precalc side effects, substep list ----
  std::vector<SubOp> schedule() {
    std::vector<SubOp> ops;
    const size_t K = chains.size();
    if (K > 1) timer_copy = timer;  // chainBackup (propagator.h:396-409)
    for (size_t c = 1; c <= K; ++c) {
      if (c > 1) timer = timer_copy;  // chainRestore (:371-380)
      const Chain& ch = chains[c - 1];
      for (size_t SI = 0; SI < ch.reps.size(); ++SI) {
        const int rep = ch.reps[SI];
        if (couples() == rep) coup.precalc(timer);  // chainEvolve (:311-332)
        if (c == K) {
          // only the last chain's result survives chainComplete (:381-395)
          SubOp o{rep, dt * ch.mults[SI], {0, 0, 0}};
          if (couples() == rep) coup.values(o.C);
          ops.push_back(o);
        }
        if (couples() == rep) timer += dt * (ch.mults[SI] * 1.0);
      }
    }
    ++step;
    return ops;
  }

  // makeFakeStep (propagator.h:347-366): every chain's precalc side effects and timer
  // increments with no evolution and, unlike makeStep, no chainBackup/Restore of the timer.
  void fake_step() {
    for (const Chain& ch : chains)
      for (size_t SI = 0; SI < ch.reps.size(); ++SI)
        if (couples() == ch.reps[SI]) {
          coup.precalc(timer);
          timer += dt * (ch.mults[SI] * 1.0);
        }
    ++step;
  }

  // ---- unfused execution: one standalone call per reference member ----
  void exec_simple(const std::vector<SubOp>& ops) {
    for (size_t SI = 0; SI < ops.size(); ++SI) {
      for (auto& gr : set->g) {
        if (SI > 0) gr->fourier(ops[SI].rep);
        gr->evolve(pd.mode, ops[SI].rep, ops[SI].delta, ops[SI].C, couples());
      }
    }
  }

  // ---- fused execution ----
  // Each P substep is one row pass on the last axis (inverse of the previous P, the X factors,
  // forward FFT) and column passes for the other axes; the trailing X substep of the step
  // stays pending and is applied by the next pass that reads the grid.
  // The main region's first pass of a step observes the previous boundary state and, for a
  // multigrid, writes the transfer edges and applies its region mask; the receiving regions
  // then run their fused transfers before their own substeps.
  void run_receives() {
    std::vector<const XFactor*> xa(set->g.size(), nullptr);
    for (size_t r = 1; r < set->g.size(); ++r)
      if (deferred[r] && pend[r].on) xa[r] = &pend[r].f;
    set->receive_fused(xa.data());
    for (size_t r = 1; r < set->g.size(); ++r) {
      deferred[r] = false;
      pend[r].on = false;
    }
    pend_post = false;
  }

  void exec_fused(const std::vector<SubOp>& ops) {
    // every phase table of this step first (each grid's own ring of slots), so the tables of
    // all region grids build in one batched launch instead of one launch per pass
    struct Sub { XFactor fb; const void* pf[3]; };
    const size_t R = set->g.size();
    std::vector<std::vector<Sub>> subs(R);
    std::vector<XFactor> last(R);
    for (size_t r = 0; r < R; ++r) {
      Grid& g = *set->g[r];
      for (size_t SI = 0; SI < ops.size(); ++SI) {
        const SubOp& o = ops[SI];
        if (o.rep != QSF_REP_P) continue;
        const SubOp& xb = ops[SI - 1];
        Sub sb;
        sb.fb = g.x_factor(pd.mode, xb.delta, xb.C, couples() == QSF_REP_X);
        g.p_factor(pd.mode, o.delta, o.C, couples() == QSF_REP_P, true, sb.pf);
        subs[r].push_back(sb);
      }
      const SubOp& xl = ops.back();
      last[r] = g.x_factor(pd.mode, xl.delta, xl.C, couples() == QSF_REP_X);
    }
    set->build_tables();
    for (size_t r = 0; r < R; ++r) {
      Grid& g = *set->g[r];
      if (r == 1 && pend_post) run_receives();
      const bool obs_here = (r == 0) && pend_obs;
      size_t k = 0;
      for (size_t SI = 0; SI < ops.size(); ++SI) {
        if (ops[SI].rep != QSF_REP_P) continue;
        const bool first = (SI == 1);
        const Sub& sb = subs[r][k++];
        run_p_substep(g, r, sb.fb, sb.pf, first && obs_here, first && r == 0 && pend_post);
      }
      pend[r].on = true;
      pend[r].f = last[r];
      deferred[r] = true;
    }
    if (pend_obs) pend_obs = false;
    if (set->multi) pend_post = true;
  }

  bool fused_multi() const { return set->multi && set->D == 2; }

  // grids whose fused P substep runs as one forward-factor-inverse pass per axis (3D main
  // grids); 2D grids keep the row pass [inverse, X, forward] + column pass [forward, P,
  // inverse], which is two sweeps either way and carries the multigrid transfer bookkeeping
  static bool per_axis_kinetic(const Grid& g) { return g.D == 3 && g.fftAxes.size() == 3; }

  // multigrid post-compute work of the main region folded into its pass
  void add_post(Grid& g, PassParams& p) {
    set->ensure_edges();
    p.er_out = set->er_main.p;
    p.ec_out = set->ec_main.p;
    p.flags |= F_MASK;
    g.fill_mask(p);
  }

  void obs_into(Grid& g, PassParams& p) {
    p.flags |= F_OBS;
    g.obs_params(p.obs, (pd.outputs & QSF_OUT_EPOT) != 0);
    if (pd.nborders > 0 && (pd.outputs & QSF_OUT_FLUX)) g.set_gather(p.obs);
    if (pend_ekin) {
      g.aux.ensure(g.m * g.cb);
      p.obs.copy_out = g.aux.p;
    }
  }

  double* obs_row() { return devrec.as<double>() + (size_t)pend_row * (3 + 16); }

  void after_obs(Grid& g, int tiles) {
    // reduce [norm, epot] of the boundary state into this row's device record
    double* rec = obs_row();
    if (tiles > 1) launch_reduce_partials(g.partials.as<double>(), tiles, 3, rec, g.redscratch.as<double>(), st());
    if (pd.nborders > 0 && (pd.outputs & QSF_OUT_FLUX))
      g.flux_from_gather(couples() == QSF_REP_P ? pend_obs_C : nullptr, rec + 3, fluxint.as<double>(), dt);
    if (pend_ekin) {
      // kinetic energy: forward transform of the copied boundary state, then <½p²>
      Grid& gg = g;
      for (int a : gg.fftAxes) {
        PassParams q = gg.params(a, gg.aux.p, gg.aux.p);
        q.flags = F_FWD;
        gg.run(q);
      }
      PassParams q = gg.params(gg.D - 1, gg.aux.p, gg.aux.p);
      q.flags = F_OBS | F_NOSTORE;
      gg.obs_params(q.obs, false);
      int64_t maxn = *std::max_element(gg.n, gg.n + 3);
      for (int a = 0; a < gg.D; ++a) q.obs.kin_w[a] = gg.kinw.as<double>() + a * maxn;
      int t2 = gg.run(q);
      // the kinetic column alone, straight into the record row (its norm and epot stay)
      launch_reduce_partials(gg.partials.as<double>() + 2, t2, 3, rec + 2, gg.redscratch.as<double>(), st(), 1);
    }
  }

  void run_p_substep(Grid& g, size_t r, const XFactor& fb, const void* const* pf, bool observe, bool post) {
    const int D = g.D;
    const auto& ax = g.fftAxes;
    const bool defer = deferred[r];
    if (ax.empty()) {  // no transformed axis: pointwise pass
      PassParams p = g.params(D - 1, g.psi.p, g.psi.p);
      p.flags = F_XB | F_P;
      p.xb = fb;
      if (defer && pend[r].on) { p.flags |= F_XA; p.xa = pend[r].f; }
      for (int a = 0; a < 3; ++a) p.pf[a] = pf[a];
      g.run(p);
      deferred[r] = true;
      pend[r].on = false;
      return;
    }
    if (ax.size() == 1) {  // one transformed axis: the whole substep is one pass
      PassParams p = g.params(ax[0], g.psi.p, g.psi.p);
      p.flags = F_XB | F_FWD | F_P | F_INV2;
      p.xb = fb;
      if (defer && pend[r].on) { p.flags |= F_XA; p.xa = pend[r].f; }
      for (int a = 0; a < 3; ++a) p.pf[a] = pf[a];
      int tiles = 0;
      if (observe) obs_into(g, p);
      tiles = g.run(p, -1, observe ? obs_row() : nullptr);
      if (observe) after_obs(g, tiles);
      deferred[r] = true;
      pend[r].on = false;
      return;
    }
    if (per_axis_kinetic(g)) {
      // 3D main grid: the kinetic factor is a product of commuting one-axis factors,
      // exp(-iT dt) = prod_d F_d^-1 f_d(p_d) F_d, so each axis pass transforms forward, applies
      // its own axis' factor and transforms back. Three sweeps per P substep (the last-axis pass
      // also carries the X halves and the observation); the state stays in X between passes.
      for (int a = D - 1; a >= 0; --a) {
        PassParams p = g.params(a, g.psi.p, g.psi.p);
        p.flags = F_FWD | F_P | F_INV2;
        p.pf[a] = pf[a];
        if (a == D - 1) {
          p.flags |= F_XB;
          p.xb = fb;
          if (pend[r].on) { p.flags |= F_XA; p.xa = pend[r].f; }
          if (observe) {
            obs_into(g, p);
            if (p.obs.gather_rowptr) {
              // the state is in X before this pass: its flux stencil is read ahead, so the pass
              // does no gather (a row pass's gather costs it more than this read)
              g.gather_ahead(pend[r].on ? &pend[r].f : nullptr);
              p.obs.gather_rowptr = nullptr;
              p.obs.gather_pos = nullptr;
              p.obs.gather_out = nullptr;
            }
          }
          int tiles = g.run(p, -1, observe ? obs_row() : nullptr);
          if (observe) after_obs(g, tiles);
        } else {
          g.run(p);
        }
      }
      deferred[r] = true;
      pend[r].on = false;
      return;
    }
    // main grid, D >= 2: row pass on the last axis
    {
      PassParams p = g.params(D - 1, g.psi.p, g.psi.p);
      p.flags = F_XB | F_FWD;
      p.xb = fb;
      if (defer) {
        p.flags |= F_INV1;
        if (pend[r].on) { p.flags |= F_XA; p.xa = pend[r].f; }
        if (post) add_post(g, p);
      }
      if (observe) obs_into(g, p);
      int tiles = g.run(p, -1, observe ? obs_row() : nullptr);
      if (observe) after_obs(g, tiles);
    }
    for (int a = D - 2; a >= 1; --a) {  // middle axes forward
      PassParams p = g.params(a, g.psi.p, g.psi.p);
      p.flags = F_FWD;
      g.run(p);
    }
    {
      PassParams p = g.params(0, g.psi.p, g.psi.p);
      p.flags = F_FWD | F_P | F_INV2;
      for (int a = 0; a < 3; ++a) p.pf[a] = pf[a];
      g.run(p);
    }
    for (int a = 1; a <= D - 2; ++a) {  // middle axes inverse
      PassParams p = g.params(a, g.psi.p, g.psi.p);
      p.flags = F_INV1;
      g.run(p);
    }
    deferred[r] = true;
    pend[r].on = false;
  }

  // Apply whatever the pipeline left pending so every grid holds its step-boundary state.
  void flush() {
    const bool post = pend_post;
    for (size_t r = 0; r < set->g.size(); ++r) {
      if (r == 1 && post) {
        run_receives();
        break;  // the receives applied every region's pending X factor
      }
      if (!deferred[r]) continue;
      Grid& g = *set->g[r];
      const bool observe = (r == 0) && pend_obs;
      const auto& ax = g.fftAxes;
      PassParams p = g.params(ax.size() == 1 ? ax[0] : g.D - 1, g.psi.p, g.psi.p);
      p.flags = 0;
      if (ax.size() >= 2 && !per_axis_kinetic(g)) p.flags |= F_INV1;   // pending inverse of the row axis
      if (pend[r].on) { p.flags |= F_XA; p.xa = pend[r].f; }
      if (r == 0 && post) add_post(g, p);
      if (observe) obs_into(g, p);
      if (p.flags) {
        int tiles = g.run(p, -1, observe ? obs_row() : nullptr);
        if (observe) after_obs(g, tiles);
      }
      deferred[r] = false;
      pend[r].on = false;
    }
    pend_obs = false;
    pend_post = false;
  }

  // ---- computeEach (propagator.h:275-300), canonical output order of qsf_types.h ----
  void request_row(bool fused) {
    const double nan = std::nan("");
    std::vector<double> row(width, nan);
    row[0] = (double)step;
    row[1] = timer;
    const int o = pd.outputs;
    if (!fused) {
      Grid& g = main();
      if (o & QSF_OUT_NORMALIZE) g.normalize();
      double r3[3] = {nan, nan, nan};
      if (o & (QSF_OUT_NORM | QSF_OUT_EPOT)) g.observe((o & QSF_OUT_EPOT) != 0, false, r3);
      if (o & QSF_OUT_NORM) row[2] = r3[0] * g.dV;
      double epot = nan, ekin = nan;
      if (o & QSF_OUT_EPOT) row[3] = epot = r3[1] * g.dV;
      if ((o & QSF_OUT_FLUX) && pd.nborders > 0) {
        double C[3];
        coup.values(C);
        // the flux time integral accumulates on the device, as on the fused path
        g.flux(couples() == QSF_REP_P ? C : nullptr, pd.borders, pd.nborders, &row[7], fluxint.as<double>(), dt);
      }
      if (o & QSF_OUT_EKIN) {
        g.fourier(QSF_REP_P);
        double r[3];
        g.observe(false, true, r);
        row[4] = ekin = r[2] * g.dVm;
        g.fourier(QSF_REP_X);
      }
      if (o & QSF_OUT_ETOT) row[5] = epot + ekin;
      if (o & QSF_OUT_EDIFF) {
        const double curr = epot + ekin;
        row[6] = curr - last_change;
        dif_energy = curr - last_change;
        last_change = curr;
      }
      rows.insert(rows.end(), row.begin(), row.end());
      return;
    }
    // fused: the values come from the next pass over the main grid
    const int ri = (int)(rows.size() / width);
    rows.insert(rows.end(), row.begin(), row.end());
    if (devrec_used >= devrec_cap) {
      sync_rows();
      devrec_used = 0;
    }
    if (!devrec.p) {
      devrec_cap = 4096;
      devrec.ensure(devrec_cap * (3 + 16) * sizeof(double));
    }
    pend_obs = true;
    pend_row = (int)devrec_used++;
    pend_ekin = (o & (QSF_OUT_EKIN | QSF_OUT_ETOT)) != 0;
    coup.values(pend_obs_C);
    pending_rows.push_back({ri, pend_ekin});
  }

  void sync_rows() {
    if (pending_rows.empty()) return;
    if (pend_obs) flush_obs_only();
    std::vector<double> dev((size_t)devrec_used * (3 + 16));
    QSF_HIP(hipMemcpyAsync(dev.data(), devrec.p, dev.size() * sizeof(double), hipMemcpyDeviceToHost, st()));
    QSF_HIP(hipStreamSynchronize(st()));
    Grid& g = main();
    const int o = pd.outputs;
    for (size_t k = 0; k < pending_rows.size(); ++k) {
      double* row = &rows[(size_t)pending_rows[k].row * width];
      const double* d = &dev[k * (3 + 16)];
      if (o & QSF_OUT_NORM) row[2] = d[0] * g.dV;
      if (o & QSF_OUT_EPOT) row[3] = d[1] * g.dV;
      if (o & QSF_OUT_EKIN) row[4] = d[2] * g.dVm;
      if (o & QSF_OUT_ETOT) row[5] = d[1] * g.dV + d[2] * g.dVm;
      if (o & QSF_OUT_FLUX)
        for (int b = 0; b < pd.nborders; ++b) row[7 + b] = d[3 + b];
    }
    pending_rows.clear();
    devrec_used = 0;
  }

  void flush_obs_only() { flush(); }

  // the fused pipeline left work for the next pass (the grids are not at the step boundary)
  bool pending() const {
    if (!fused()) return false;
    for (size_t r = 0; r < deferred.size(); ++r)
      if (deferred[r] || pend[r].on) return true;
    return pend_obs || pend_post;
  }

  bool fused() const { return exec == 0; }
  ~Prop() {
    if (!set) return;
    set->props.erase(std::remove(set->props.begin(), set->props.end(), this), set->props.end());
    if (set->active == this) set->active = nullptr;
  }
  Prop(const Prop&) = delete;
  Prop& operator=(const Prop&) = delete;

  // A caller that reads or changes a grid between steps (a worker at WHEN::DURING,
  // propagator.h:446-452) must see the step-boundary state: apply what the fused pipeline
  // left pending (inverse transform, trailing X half, observation, multigrid transfer).
  void settle() {
    if (fused()) flush();
  }

  void compute_each() {
    if (pd.outputs == 0) return;
    if (fused() && step > 0) request_row(true);
    else {
      if (fused()) flush();
      request_row(false);
    }
  }

  void make_step() {
    auto ops = schedule();
    if (fused()) exec_fused(ops);
    else exec_simple(ops);
  }

  void post_compute() {
    if (!set->multi) return;
    if (fused() && fused_multi()) return;  // runs inside the next pass (pend_post)
    if (fused()) flush();
    set->post_compute();
  }

  bool should_continue() const {
    if (pd.mode == QSF_MODE_RE) return step <= max_steps;
    return std::fabs(dif_energy) > state_accuracy && step != max_steps;
  }
  void start() { compute_each(); }
  void loop_body() {
    make_step();
    compute_each();
    post_compute();
  }
};

Set::~Set() {
  for (Prop* p : props) p->set = nullptr;
}

// ----------------------------------------------------------------------------------------
// Batches: many single-region 2D solves (Strang, velocity gauge) advanced together. Solves are
// grouped by grid description; one step is a fixed sequence of batched launches driven by a
// device step counter, so it is captured once into a hipGraph and replayed.
// ----------------------------------------------------------------------------------------
struct Batch {
  qsf_ctx* ctx;
  qsf_prop_desc pd;
  int prec = QSF_FP64;
  struct Solve {
    std::unique_ptr<Grid> g;
    int64_t max_steps = 0, total = 0;
    std::vector<double> times;      // timer after each step (host schedule)
    std::vector<double> coup;       // [total][3]
    DBuf coupling, rec, fluxint, partials, ptab, gathered;
    std::vector<double> start_row, end_row;
    int group = 0;
  };
  std::vector<Solve> solves;
  struct Group {
    std::vector<int> members;
    Grid* lead = nullptr;
    XFactor xh{};                   // static X half factor (no X coupling in the velocity gauge)
    DBuf entries, oargs;
    int64_t tiles_row = 0;
    BatchFluxGeo geo{};
  };
  std::vector<Group> groups;
  DBuf targs, ctr;
  int64_t launched = 0;             // steps issued so far (host mirror of the device counter)
  int64_t tmax = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // Size groups share nothing within a step but the phase tables and the step counter, so their
  // passes and observations run on nlanes streams (lane 0 is the context stream;
  // QSF_BATCH_STREAMS = 1..4, default 4, one per hardware queue of the process): forked after the
  // table launch, joined before the counter advances. Groups go to lanes largest first, each to
  // the lane with the fewest grid points so far, so the small groups fill the tails of the large
  // ones.
  static constexpr int kLanes = 4;
  int nlanes = 0;                    // 0 until the lanes are made
  hipStream_t lane_st[kLanes] = {};
  hipEvent_t lane_ev[kLanes] = {};   // lane 0: the fork; lanes 1..: their last launch has finished
  std::vector<int> lane_of;          // per group

  hipStream_t st() const { return ctx->stream; }
  hipStream_t lane(int gi) const { return lane_of[gi] == 0 ? st() : lane_st[lane_of[gi]]; }

  void ensure_lanes() {
    if (nlanes) return;
    int want = kLanes;
    if (const char* e = std::getenv("QSF_BATCH_STREAMS")) want = std::max(1, std::min(kLanes, std::atoi(e)));
    for (int l = 0; l < want; ++l) {
      if (l > 0) QSF_HIP(hipStreamCreateWithFlags(&lane_st[l], hipStreamNonBlocking));
      QSF_HIP(hipEventCreateWithFlags(&lane_ev[l], hipEventDisableTiming));
    }
    nlanes = want;
    std::vector<int> order(groups.size());
    std::vector<double> work(groups.size());
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const Grid& L = *groups[gi].lead;
      work[gi] = (double)groups[gi].members.size() * (double)L.n[0] * (double)(L.D > 1 ? L.n[1] : 1);
      order[gi] = (int)gi;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    lane_of.assign(groups.size(), 0);
    double load[kLanes] = {};
    for (int gi : order) {
      int best = 0;
      for (int l = 1; l < nlanes; ++l)
        if (load[l] < load[best]) best = l;
      lane_of[gi] = best;
      load[best] += work[gi];
    }
  }

  Batch(qsf_ctx* c, const qsf_grid_desc* descs, const qsf_field_desc* fields, int ns, const qsf_prop_desc& p)
      : ctx(c), pd(p) {
    QSF_REQUIRE(ns >= 1, QSF_ERR_ARG, "empty batch");
    QSF_REQUIRE(p.mode == QSF_MODE_RE && p.order == 1 && p.gauge != QSF_GAUGE_LENGTH, QSF_ERR_UNSUPPORTED,
                "batched solves run real-time Strang steps in the velocity gauge");
    QSF_REQUIRE(p.nborders >= 0 && p.nborders <= 16, QSF_ERR_ARG, "0..16 borders");
    prec = descs[0].precision;
    solves.resize(ns);
    for (int s = 0; s < ns; ++s) {
      const qsf_grid_desc& d = descs[s];
      QSF_REQUIRE((d.dim == 1 || d.dim == 2) && !d.multigrid && d.precision == prec, QSF_ERR_UNSUPPORTED,
                  "a batch holds single-region 1D and 2D grids of one precision");
      Solve& so = solves[s];
      so.g.reset(new Grid(c, d));
      QSF_REQUIRE(!so.g->radial, QSF_ERR_UNSUPPORTED,
                  "batched solves take separable potentials (CoulombInteraction in 2D runs as a single solve)");
      // host schedule: SplitPropagator RE loop of Strang steps (propagator.h:153, 311-346, 443-451)
      phys::Coupling coup(QSF_GAUGE_VELOCITY, &fields[s], 1);
      so.max_steps = (int64_t)(coup.max_pulse_duration() / p.dt + 1);
      so.total = so.max_steps + 2;
      so.coup.assign(so.total * 3, 0.0);
      so.times.assign(so.total + 1, 0.0);
      double timer = 0.0;
      for (int64_t k = 0; k < so.total; ++k) {
        coup.precalc(timer);            // the P substep's precalc (X substeps do not couple)
        coup.values(&so.coup[3 * k]);
        timer += p.dt * (1.0 * 1.0);
        so.times[k + 1] = timer;
      }
      tmax = std::max(tmax, so.total);
      so.coupling.ensure(so.coup.size() * sizeof(double));
      QSF_HIP(hipMemcpy(so.coupling.p, so.coup.data(), so.coup.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // groups of identical grid descriptions
    for (int s = 0; s < ns; ++s) {
      int gi = -1;
      for (size_t k = 0; k < groups.size(); ++k)
        if (std::memcmp(&solves[groups[k].members[0]].g->d, &descs[s], sizeof(qsf_grid_desc)) == 0) gi = (int)k;
      if (gi < 0) { groups.emplace_back(); gi = (int)groups.size() - 1; }
      groups[gi].members.push_back(s);
      solves[s].group = gi;
    }
    const double zero[3] = {0, 0, 0};
    std::vector<BatchTableArgs> ta(ns);
    for (auto& G : groups) {
      G.lead = solves[G.members[0]].g.get();
      Grid& L = *G.lead;
      G.xh = L.x_factor(QSF_MODE_RE, 0.5 * p.dt, zero, false);   // built once, never recycled
      L.build_tables();
      if (p.nborders > 0) L.build_gather(p.borders, p.nborders);
      // the pass that observes: the row pass of a 2D solve, the one pass of a 1D solve
      const int ra = L.D - 1;
      PassParams rp = L.params(ra, L.psi.p, L.psi.p);
      G.tiles_row = axis_pass_tiles(prec, rp, ctx->variant_of(prec, tune_class(rp), L.n[ra], L.m));
      std::vector<BatchEntry> be(G.members.size());
      std::vector<BatchObsArgs> oa(G.members.size());
      for (size_t k = 0; k < G.members.size(); ++k) {
        Solve& so = solves[G.members[k]];
        Grid& g = *so.g;
        so.ptab.ensure(2 * std::max(g.n[0], g.n[1]) * Grid::kTabBytes);
        so.partials.ensure(std::max<int64_t>(G.tiles_row, 1) * 3 * sizeof(double));
        so.rec.ensure((so.total + 1) * (1 + 16) * sizeof(double));
        so.fluxint.ensure(16 * sizeof(double));
        QSF_HIP(hipMemset(so.fluxint.p, 0, 16 * sizeof(double)));
        if (p.nborders > 0) so.gathered.ensure(std::max<int64_t>(L.gather->nvals, 1) * g.cb);
        char* pt = static_cast<char*>(so.ptab.p);
        const int64_t half = std::max(g.n[0], g.n[1]) * Grid::kTabBytes;
        be[k] = BatchEntry{g.psi.p, {pt, g.D > 1 ? pt + half : nullptr, nullptr}, so.partials.as<double>(),
                           p.nborders > 0 ? so.gathered.p : nullptr, so.total};
        BatchTableArgs& t = ta[G.members[k]];
        t = BatchTableArgs{};   // n[a] = 0 for the axes a 1D solve lacks: no table there
        for (int a = 0; a < g.D; ++a) {
          const int sa = g.pperm[a];
          t.out[a] = a == 0 ? (void*)pt : (void*)(pt + half);
          t.n[a] = g.n[a];
          t.pn[a] = g.n[sa];
          t.pn2[a] = g.n2[sa];
          t.dp[a] = g.dp[sa];
          t.storage_axis[a] = sa;
        }
        t.delta = p.dt;
        t.coupling = so.coupling.as<double>();
        t.total_steps = so.total;
        oa[k] = BatchObsArgs{so.partials.as<double>(), G.tiles_row, so.gathered.p, so.coupling.as<double>(),
                             so.total, so.rec.as<double>(), so.fluxint.as<double>(), g.dV};
      }
      G.entries.ensure(be.size() * sizeof(BatchEntry));
      QSF_HIP(hipMemcpy(G.entries.p, be.data(), be.size() * sizeof(BatchEntry), hipMemcpyHostToDevice));
      G.oargs.ensure(oa.size() * sizeof(BatchObsArgs));
      QSF_HIP(hipMemcpy(G.oargs.p, oa.data(), oa.size() * sizeof(BatchObsArgs), hipMemcpyHostToDevice));
      G.geo = BatchFluxGeo{};
      G.geo.D = L.D;
      G.geo.precision = prec;
      G.geo.dt = p.dt;
      G.geo.nb = p.nborders;
      G.geo.scale = L.dV / L.dx[0];
      for (int a = 0; a < 3; ++a) G.geo.inv_2dx[a] = a < L.D ? (1.0 / L.dx[a]) / 2.0 : 0.0;
      if (p.nborders > 0) {
        G.geo.pts = L.gather->pts.as<int32_t>();
        G.geo.bvec = L.gather->bvec.as<double>();
        G.geo.npts = L.gather->npts;
      }
    }
    targs.ensure(ta.size() * sizeof(BatchTableArgs));
    QSF_HIP(hipMemcpy(targs.p, ta.data(), ta.size() * sizeof(BatchTableArgs), hipMemcpyHostToDevice));
    ctr.ensure(sizeof(int64_t));
    const int64_t z = 0;
    QSF_HIP(hipMemcpy(ctr.p, &z, sizeof(int64_t), hipMemcpyHostToDevice));
  }

  ~Batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int l = 0; l < kLanes; ++l) {
      if (lane_st[l]) {
        (void)hipStreamSynchronize(lane_st[l]);
        (void)hipStreamDestroy(lane_st[l]);
      }
      if (lane_ev[l]) (void)hipEventDestroy(lane_ev[l]);
    }
  }

  int64_t nmax() const {
    int64_t v = 0;
    for (auto& s : solves) v = std::max({v, s.g->n[0], s.g->n[1]});
    return v;
  }

  // observables of the starting states (computeEach<AT_START>)
  void start() {
    for (auto& so : solves) so.start_row = observe_now(so, 0);
  }

  std::vector<double> observe_now(Solve& so, int64_t k) {
    Grid& g = *so.g;
    std::vector<double> row(2 + 1 + pd.nborders, std::nan(""));
    row[0] = (double)k;
    row[1] = so.times[k];
    double r3[3];
    g.observe(false, false, r3);
    row[2] = r3[0] * g.dV;
    if (pd.nborders > 0) {
      const double* C = k > 0 ? &so.coup[3 * (k - 1)] : nullptr;
      g.flux(C, pd.borders, pd.nborders, &row[3]);
    }
    return row;
  }

  // one step of every active solve; the first step has no pending inverse / X / observation
  void enqueue_step(bool first) {
    launch_batch_ptables(prec, targs.as<BatchTableArgs>(), (int)solves.size(), nmax(), ctr.as<int64_t>(), st());
    bool used[kLanes] = {};
    for (size_t gi = 0; gi < groups.size(); ++gi) used[lane_of[gi]] = true;
    QSF_HIP(hipEventRecord(lane_ev[0], st()));
    for (int l = 1; l < kLanes; ++l)
      if (used[l]) QSF_HIP(hipStreamWaitEvent(lane_st[l], lane_ev[0], 0));
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      Group& G = groups[gi];
      Grid& L = *G.lead;
      const hipStream_t ls = lane((int)gi);
      // 1D: one pass [X_a, OBS, X_b, FWD, P, INV] (the state stays in X between steps);
      // 2D: row pass [INV, X_a, OBS, X_b, FWD] and column pass [FWD, P, INV]
      const int ra = L.D - 1;
      PassParams rp = L.params(ra, L.psi.p, L.psi.p);
      rp.batch = G.entries.as<BatchEntry>();
      rp.step_ctr = ctr.as<int64_t>();
      rp.nbatch = (int32_t)G.members.size();
      rp.flags = F_XB | F_FWD | (L.D == 1 ? F_P | F_INV2 : 0u);
      rp.xb = G.xh;
      if (!first) {
        rp.flags |= F_XA | F_OBS | (L.D == 1 ? 0u : F_INV1);
        rp.xa = G.xh;
        L.obs_params(rp.obs, false);
        if (pd.nborders > 0) L.set_gather(rp.obs);
      }
      int k = launch_axis_pass(prec, rp, ls, ctx->variant_of(prec, tune_class(rp), L.n[ra], L.m));
      QSF_REQUIRE(k >= 0, QSF_ERR_UNSUPPORTED, "no FFT kernel for this axis length");
      // reads only what the row pass just wrote (block partials, gathered border points)
      if (!first) launch_batch_obs(G.oargs.as<BatchObsArgs>(), (int)G.members.size(), G.geo, ctr.as<int64_t>(), ls);
      if (L.D == 1) continue;
      PassParams cp = L.params(0, L.psi.p, L.psi.p);
      cp.batch = G.entries.as<BatchEntry>();
      cp.step_ctr = ctr.as<int64_t>();
      cp.nbatch = (int32_t)G.members.size();
      cp.flags = F_FWD | F_P | F_INV2;
      k = launch_axis_pass(prec, cp, ls, ctx->variant_of(prec, tune_class(cp), L.n[0], L.m));
      QSF_REQUIRE(k >= 0, QSF_ERR_UNSUPPORTED, "no FFT kernel for this axis length");
    }
    for (int l = 1; l < kLanes; ++l)
      if (used[l]) {
        QSF_HIP(hipEventRecord(lane_ev[l], lane_st[l]));
        QSF_HIP(hipStreamWaitEvent(st(), lane_ev[l], 0));
      }
    launch_step_increment(ctr.as<int64_t>(), st());
    QSF_HIP(hipGetLastError());
  }

  void run(int64_t k, bool use_graph) {
    ensure_lanes();
    while (k > 0 && launched < tmax) {
      if (launched == 0) {
        enqueue_step(true);
      } else if (use_graph) {
        if (!gexec) {
          QSF_HIP(hipStreamBeginCapture(st(), hipStreamCaptureModeThreadLocal));
          enqueue_step(false);
          QSF_HIP(hipStreamEndCapture(st(), &graph));
          QSF_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        QSF_HIP(hipGraphLaunch(gexec, st()));
      } else {
        enqueue_step(false);
      }
      ++launched;
      --k;
    }
  }

  bool done() const { return launched >= tmax; }

  // After the last step: apply each solve's pending inverse + X half and observe the end state
  void finish() {
    for (auto& so : solves) {
      Grid& g = *so.g;
      PassParams p = g.params(g.D - 1, g.psi.p, g.psi.p);
      p.flags = F_XA | (g.D == 1 ? 0u : F_INV1);
      p.xa = groups[so.group].xh;
      g.run(p);
      so.end_row = observe_now(so, so.total);
    }
  }

  std::vector<double> records(int s) {
    Solve& so = solves[s];
    const int nb = pd.nborders, w = 3 + nb;
    std::vector<double> dev((so.total + 1) * (1 + 16));
    QSF_HIP(hipStreamSynchronize(st()));
    QSF_HIP(hipMemcpy(dev.data(), so.rec.p, dev.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> out;
    const int64_t upto = std::min<int64_t>(launched - 1, so.total - 1);
    if (!so.start_row.empty()) out.insert(out.end(), so.start_row.begin(), so.start_row.end());
    for (int64_t k = 1; k <= upto; ++k) {
      std::vector<double> row(w);
      row[0] = (double)k;
      row[1] = so.times[k];
      for (int j = 0; j < 1 + nb; ++j) row[2 + j] = dev[k * (1 + nb) + j];
      out.insert(out.end(), row.begin(), row.end());
    }
    if (!so.end_row.empty()) out.insert(out.end(), so.end_row.begin(), so.end_row.end());
    return out;
  }
};

}  // namespace qsf

using namespace qsf;

struct qsf_grid { Grid* g; Set* set = nullptr; };
struct qsf_set {
  Set s;
  std::vector<std::unique_ptr<qsf_grid>> handles;   // one per region grid, owned by the set
  qsf_set(qsf_ctx* c, const qsf_grid_desc& d) : s(c, d) {
    for (auto& g : s.g) handles.emplace_back(new qsf_grid{g.get(), &s});
  }
};
struct qsf_prop { Prop p; qsf_prop(Set* s, const qsf_prop_desc& d, const qsf_field_desc* f, int nf) : p(s, d, f, nf) {} };

struct qsf_batch { Batch b; qsf_batch(qsf_ctx* c, const qsf_grid_desc* d, const qsf_field_desc* f, int n,
                                         const qsf_prop_desc& p) : b(c, d, f, n, p) {} };

namespace {
thread_local std::string g_noctx_err;

template <typename F> int guard(qsf_ctx* ctx, F&& f) {
  try {
    f();
    return QSF_OK;
  } catch (const Error& e) {
    (ctx ? ctx->err : g_noctx_err) = e.what();
    return e.code;
  } catch (const std::exception& e) {
    (ctx ? ctx->err : g_noctx_err) = e.what();
    return QSF_ERR_ARG;
  }
}
void settle(const Set* s) {
  if (s->active) s->active->settle();
}
void settle(qsf_grid* g) {
  if (g->set) settle(g->set);
}
}  // namespace

extern "C" {

int qsf_ctx_create(int device, qsf_ctx** out) {
  if (!out) return QSF_ERR_ARG;
  auto* c = new qsf_ctx();
  c->device = device;
  int rc = guard(nullptr, [&] {
    int count = 0;
    QSF_HIP(hipGetDeviceCount(&count));
    QSF_REQUIRE(device >= 0 && device < count, QSF_ERR_ARG, "no such device");
    QSF_HIP(hipSetDevice(device));
    c->scratch = scratch_address(device);
    QSF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    QSF_HIP(hipEventCreate(&c->ev[0]));
    QSF_HIP(hipEventCreate(&c->ev[1]));
  });
  if (rc != QSF_OK) {
    delete c;
    *out = nullptr;
    return rc;
  }
  *out = c;
  return QSF_OK;
}

int qsf_prof_begin(qsf_ctx* c) {
  if (!c) return QSF_ERR_ARG;
  return guard(c, [&] {
    drain_prof(c);
    c->entries.clear();
    c->prof = true;
  });
}

int qsf_prof_end(qsf_ctx* c, int* n) {
  if (!c || !n) return QSF_ERR_ARG;
  return guard(c, [&] {
    drain_prof(c);
    c->prof = false;
    *n = (int)c->entries.size();
  });
}

int qsf_prof_entry(qsf_ctx* c, int i, char* name, int len, int64_t* k, double* ms, double* bytes) {
  if (!c || i < 0 || i >= (int)c->entries.size()) return QSF_ERR_ARG;
  const auto& e = c->entries[i];
  if (name && len > 0) {
    std::strncpy(name, e.key.c_str(), len - 1);
    name[len - 1] = 0;
  }
  if (k) *k = e.n;
  if (ms) *ms = e.ms;
  if (bytes) *bytes = e.bytes;
  return QSF_OK;
}

int qsf_autotune(qsf_grid* g, int reps) {
  if (!g || reps < 1) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] { g->g->autotune(reps); });
}

int qsf_tune_count(qsf_ctx* c, int* n) {
  if (!c || !n) return QSF_ERR_ARG;
  *n = (int)c->tune.size();
  return QSF_OK;
}

int qsf_tune_entry(qsf_ctx* c, int i, int* prec, int* row, int64_t* n, int64_t* m, int* variant, double* ms) {
  if (!c || i < 0 || i >= (int)c->tune.size()) return QSF_ERR_ARG;
  auto it = c->tune.begin();
  std::advance(it, i);
  if (prec) *prec = it->first.prec;
  if (row) *row = it->first.row;
  if (n) *n = it->first.n;
  if (m) *m = it->first.m;
  if (variant) *variant = it->second.first;
  if (ms) *ms = it->second.second;
  return QSF_OK;
}

int qsf_tune_set(qsf_ctx* c, int prec, int row, int64_t n, int64_t m, int variant) {
  if (!c || row < 0 || row > 2 || variant < 0 || variant >= qsf::axis_variants(n)) return QSF_ERR_ARG;
  c->tune[{prec, row, n, m}] = {variant, 0.0};
  return QSF_OK;
}

int qsf_ctx_destroy(qsf_ctx* c) {
  if (!c) return QSF_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& l : c->launches) { (void)hipEventDestroy(l.a); (void)hipEventDestroy(l.b); }
  for (auto& e : c->spare) (void)hipEventDestroy(e);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return QSF_OK;
}

const char* qsf_last_error(const qsf_ctx* c) { return c ? c->err.c_str() : g_noctx_err.c_str(); }

int qsf_ctx_synchronize(qsf_ctx* c) {
  if (!c) return QSF_ERR_ARG;
  return guard(c, [&] { QSF_HIP(hipStreamSynchronize(c->stream)); });
}

int qsf_ctx_query(qsf_ctx* c, int* idle) {
  if (!c || !idle) return QSF_ERR_ARG;
  return guard(c, [&] {
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipErrorNotReady) {
      (void)hipGetLastError();   // not an error: work is still running
      *idle = 0;
      return;
    }
    QSF_HIP(e);
    *idle = 1;
  });
}

int qsf_timer_mark(qsf_ctx* c) {
  if (!c) return QSF_ERR_ARG;
  return guard(c, [&] {
    QSF_HIP(hipEventRecord(c->ev[c->marks & 1], c->stream));
    ++c->marks;
  });
}

int qsf_timer_elapsed_ms(qsf_ctx* c, double* ms) {
  if (!c || !ms) return QSF_ERR_ARG;
  return guard(c, [&] {
    QSF_REQUIRE(c->marks >= 2, QSF_ERR_STATE, "need two marks");
    hipEvent_t a = c->ev[(c->marks - 2) & 1], b = c->ev[(c->marks - 1) & 1];
    QSF_HIP(hipEventSynchronize(b));
    float f = 0;
    QSF_HIP(hipEventElapsedTime(&f, a, b));
    *ms = f;
  });
}

int qsf_set_create(qsf_ctx* c, const qsf_grid_desc* d, qsf_set** out) {
  if (!c || !d || !out) return QSF_ERR_ARG;
  *out = nullptr;
  return guard(c, [&] {
    QSF_HIP(hipSetDevice(c->device));
    *out = new qsf_set(c, *d);
  });
}

int qsf_set_destroy(qsf_set* s) {
  if (!s) return QSF_ERR_ARG;
  (void)hipStreamSynchronize(s->s.ctx->stream);
  delete s;   // detaches its propagators; their own destroy stays valid
  return QSF_OK;
}

int qsf_set_count(const qsf_set* s, int* n) {
  if (!s || !n) return QSF_ERR_ARG;
  *n = (int)s->s.g.size();
  return QSF_OK;
}

int qsf_set_grid(qsf_set* s, int r, qsf_grid** out) {
  if (!s || !out || r < 0 || r >= (int)s->s.g.size()) return QSF_ERR_ARG;
  *out = s->handles[r].get();
  return QSF_OK;
}

int qsf_set_post_compute(qsf_set* s) {
  if (!s) return QSF_ERR_ARG;
  return guard(s->s.ctx, [&] {
    settle(&s->s);
    s->s.post_compute();
  });
}

int qsf_grid_size(const qsf_grid* g, int64_t* m) {
  if (!g || !m) return QSF_ERR_ARG;
  *m = g->g->m;
  return QSF_OK;
}
int qsf_grid_rep(const qsf_grid* g, int* r) {
  if (!g || !r) return QSF_ERR_ARG;
  *r = g->g->rep;
  return QSF_OK;
}
int qsf_grid_upload(qsf_grid* g, const double* h, int64_t count, int rep) {
  if (!g || !h) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    g->g->upload(h, count, rep);
  });
}
int qsf_grid_download(qsf_grid* g, double* h, int64_t count) {
  if (!g || !h) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    g->g->download(h, count);
  });
}
int qsf_fourier(qsf_grid* g, int to) {
  if (!g) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    g->g->fourier(to);
  });
}
int qsf_evolve(qsf_grid* g, int mode, int rep, double delta, const double* C, int couples_in) {
  if (!g) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    QSF_REQUIRE(rep == g->g->rep, QSF_ERR_STATE, "evolve: grid is in the other representation");
    QSF_REQUIRE(mode == QSF_MODE_IM || mode == QSF_MODE_RE, QSF_ERR_ARG, "bad mode");
    g->g->evolve(mode, rep, delta, C, couples_in);
  });
}
int qsf_average(qsf_grid* g, int rep, int op, double* out) {
  if (!g || !out) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    QSF_REQUIRE(rep == g->g->rep, QSF_ERR_STATE, "average: grid is in the other representation");
    QSF_REQUIRE(op >= 0 && op <= 2, QSF_ERR_ARG, "op must be 0..2");
    *out = g->g->average(rep, op);
  });
}
int qsf_average_noalloc(qsf_grid* g, int rep, int op, double* out) {
  if (!g || !out) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    Grid& G = *g->g;
    QSF_REQUIRE(G.ctx->scratch, QSF_ERR_STATE, "no scratch address on this context's device");
    // settling would run the pending passes through the grid's own buffers: refused instead
    QSF_REQUIRE(!(g->set && g->set->active && g->set->active->pending()), QSF_ERR_STATE,
                "a propagator step is pending on this grid set; settle it first (any other grid entry "
                "point, or qsf_prop_finish)");
    QSF_REQUIRE(rep == G.rep, QSF_ERR_STATE, "average: grid is in the other representation");
    QSF_REQUIRE(op >= 0 && op <= 2, QSF_ERR_ARG, "op must be 0..2");
    *out = G.average_into(rep, op, G.ctx->scratch);
  });
}
int qsf_normalize(qsf_grid* g) {
  if (!g) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    g->g->normalize();
  });
}
int qsf_flux(qsf_grid* g, const double* vgC, const qsf_border_desc* b, int nb, double* out) {
  if (!g || !b || !out) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    g->g->flux(vgC, b, nb, out);
  });
}
extern "C++" {
namespace qsf {
void launch_antisym(int precision, void* psi, int dim, const int64_t* n, int kind, hipStream_t s);
void launch_cross_out(int precision, void* psi, int dim, const int64_t* n, int64_t from_boundary, int64_t ramp,
                      const double* tab, hipStream_t s);
void launch_overlap(int precision, const void* phi, const void* psi, int64_t m, double* partials, double* out,
                    hipStream_t s);
void launch_sub_scaled(int precision, void* psi, const void* phi, int64_t m, double sr, double si, hipStream_t s);
}
}
int qsf_orthogonalize(qsf_grid* g, const double* phi, int64_t count, double* overlap) {
  if (!g || !phi) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    Grid& G = *g->g;
    QSF_REQUIRE(G.rep == QSF_REP_X, QSF_ERR_STATE, "orthogonalizeWith acts on the X representation");
    QSF_REQUIRE(count == G.m, QSF_ERR_ARG, "the stored state must have the grid's node count");
    const size_t phi_bytes = (size_t)G.m * 16;
    G.aux.ensure(phi_bytes + (2 * 1024 + 2) * sizeof(double));
    char* base = static_cast<char*>(G.aux.p);
    double* partials = reinterpret_cast<double*>(base + phi_bytes);
    double* res = partials + 2 * 1024;
    QSF_HIP(hipMemcpyAsync(base, phi, phi_bytes, hipMemcpyHostToDevice, G.st()));
    launch_overlap(G.prec, base, G.psi.p, G.m, partials, res, G.st());
    QSF_HIP(hipGetLastError());
    double h[2];
    QSF_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, G.st()));
    QSF_HIP(hipStreamSynchronize(G.st()));
    const double sr = h[0] * G.dV, si = h[1] * G.dV;   // scalarProduct *= vol<X>
    launch_sub_scaled(G.prec, G.psi.p, base, G.m, sr, si, G.st());
    QSF_HIP(hipGetLastError());
    QSF_HIP(hipStreamSynchronize(G.st()));
    if (overlap) { overlap[0] = sr; overlap[1] = si; }
  });
}
int qsf_cross_out(qsf_grid* g, double au_distance) {
  if (!g) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    Grid& G = *g->g;
    QSF_REQUIRE(G.rep == QSF_REP_X, QSF_ERR_STATE, "croossOut acts on the X representation");
    const int64_t ramp = 30;
    // multiIonizationFilter's nodesFromCenter receives au_distance + 2 * ramp as a node count
    // (grid.h:897-899; the au_distance / dx computed there is unused)
    const int64_t from_center = (int64_t)(au_distance + 2 * ramp);
    std::vector<double> tab(ramp + 1);
    for (int64_t k = 0; k <= ramp; ++k) tab[k] = 0.5 * (1. + std::cos(M_PI * (double)k / (double)ramp));
    G.aux.ensure(tab.size() * sizeof(double));
    QSF_HIP(hipMemcpyAsync(G.aux.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, G.st()));
    launch_cross_out(G.prec, G.psi.p, G.D, G.n, G.n[0] / 2 - from_center, ramp, G.aux.as<double>(), G.st());
    QSF_HIP(hipGetLastError());
    QSF_HIP(hipStreamSynchronize(G.st()));   // `tab` is host memory read by the copy
  });
}
int qsf_antisymmetrize(qsf_grid* g, int particles) {
  if (!g || (particles != 2 && particles != 3)) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    Grid& G = *g->g;
    if (G.D < particles) return;  // the reference does nothing for DIM < particles (grid.h:613, 684)
    QSF_REQUIRE(G.rep == QSF_REP_X, QSF_ERR_STATE, "antisymmetrize acts on the X representation");
    if (particles == 2 && G.D == 2) QSF_REQUIRE(G.n[0] == G.n[1], QSF_ERR_ARG, "needs n0 == n1");
    if (G.D == 3) QSF_REQUIRE(G.n[1] == G.n[2] && (particles == 2 || G.n[0] == G.n[1]), QSF_ERR_ARG,
                              "needs equal axis lengths");
    launch_antisym(G.prec, G.psi.p, G.D, G.n, particles, G.st());
    QSF_HIP(hipGetLastError());
  });
}
int qsf_mask_region(qsf_grid* g) {
  if (!g) return QSF_ERR_ARG;
  return guard(g->g->ctx, [&] {
    settle(g);
    g->g->mask_region();
  });
}

int qsf_prop_create(qsf_set* s, const qsf_prop_desc* d, const qsf_field_desc* f, int nf, qsf_prop** out) {
  if (!s || !d || !out || (nf > 0 && !f)) return QSF_ERR_ARG;
  *out = nullptr;
  return guard(s->s.ctx, [&] { *out = new qsf_prop(&s->s, *d, f, nf); });
}
int qsf_prop_destroy(qsf_prop* p) {
  if (!p) return QSF_ERR_ARG;
  (void)hipStreamSynchronize(p->p.st());
  delete p;
  return QSF_OK;
}

namespace {
// a propagator whose set was destroyed first only accepts qsf_prop_destroy and the scalar queries
void require_set(const qsf_prop* p) {
  QSF_REQUIRE(p->p.set, QSF_ERR_STATE, "the propagator's grid set was destroyed");
}
}  // namespace
int qsf_prop_max_steps(const qsf_prop* p, int64_t* o) {
  if (!p || !o) return QSF_ERR_ARG;
  *o = p->p.max_steps;
  return QSF_OK;
}
int qsf_prop_start(qsf_prop* p) {
  if (!p) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p); p->p.start(); });
}
int qsf_prop_fast_forward(qsf_prop* p, int64_t at_step) {
  if (!p || at_step < 0) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p);
    QSF_REQUIRE(p->p.step == 0 && p->p.rows.empty(), QSF_ERR_STATE,
                "fast_forward replaces start: call it on a fresh propagator");
    while (p->p.step < at_step) p->p.fake_step();
  });
}
int qsf_prop_should_continue(qsf_prop* p, int* o) {
  if (!p || !o) return QSF_ERR_ARG;
  *o = p->p.should_continue();
  return QSF_OK;
}
int qsf_prop_step(qsf_prop* p) {
  if (!p) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p); p->p.loop_body(); });
}
int qsf_prop_run_steps(qsf_prop* p, int64_t k, int64_t* done) {
  if (!p) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p);
    int64_t i = 0;
    while (i < k && p->p.should_continue()) {
      p->p.loop_body();
      ++i;
    }
    if (done) *done = i;
  });
}
int qsf_prop_finish(qsf_prop* p) {
  if (!p) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p);
    p->p.loop_body();
    p->p.flush();
    p->p.sync_rows();
  });
}
int qsf_prop_step_count(const qsf_prop* p, int64_t* o) {
  if (!p || !o) return QSF_ERR_ARG;
  *o = p->p.step;
  return QSF_OK;
}
int qsf_prop_time(const qsf_prop* p, double* o) {
  if (!p || !o) return QSF_ERR_ARG;
  *o = p->p.timer;
  return QSF_OK;
}
int qsf_prop_coupling(const qsf_prop* p, double* o) {
  if (!p || !o) return QSF_ERR_ARG;
  p->p.coup.values(o);
  return QSF_OK;
}
int qsf_prop_record_width(const qsf_prop* p, int* o) {
  if (!p || !o) return QSF_ERR_ARG;
  *o = p->p.width;
  return QSF_OK;
}
int qsf_prop_nrecords(qsf_prop* p, int* o) {
  if (!p || !o) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p);
    p->p.flush();
    p->p.sync_rows();
    *o = (int)(p->p.rows.size() / p->p.width);
  });
}
int qsf_prop_records(qsf_prop* p, double* out, int max_rows) {
  if (!p || !out) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p);
    p->p.flush();
    p->p.sync_rows();
    const size_t n = std::min<size_t>((size_t)max_rows * p->p.width, p->p.rows.size());
    std::memcpy(out, p->p.rows.data(), n * sizeof(double));
  });
}
int qsf_prop_flux_integral(qsf_prop* p, double* out) {
  if (!p || !out) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p);
    p->p.flush();
    QSF_HIP(hipMemcpyAsync(out, p->p.fluxint.p, p->p.pd.nborders * sizeof(double), hipMemcpyDeviceToHost,
                           p->p.st()));
    QSF_HIP(hipStreamSynchronize(p->p.st()));
  });
}
int qsf_batch_create(qsf_ctx* c, const qsf_grid_desc* grids, const qsf_field_desc* fields, int n,
                     const qsf_prop_desc* p, qsf_batch** out) {
  if (!c || !grids || !fields || !p || !out || n < 1) return QSF_ERR_ARG;
  *out = nullptr;
  return guard(c, [&] { *out = new qsf_batch(c, grids, fields, n, *p); });
}
int qsf_batch_destroy(qsf_batch* b) {
  if (!b) return QSF_ERR_ARG;
  (void)hipStreamSynchronize(b->b.st());
  delete b;
  return QSF_OK;
}
int qsf_batch_upload(qsf_batch* b, int s, const double* psi) {
  if (!b || !psi || s < 0 || s >= (int)b->b.solves.size()) return QSF_ERR_ARG;
  return guard(b->b.ctx, [&] {
    QSF_REQUIRE(b->b.launched == 0, QSF_ERR_STATE, "upload before the first step");
    Grid& g = *b->b.solves[s].g;
    g.upload(psi, g.m, QSF_REP_X);
  });
}
int qsf_batch_download(qsf_batch* b, int s, double* psi) {
  if (!b || !psi || s < 0 || s >= (int)b->b.solves.size()) return QSF_ERR_ARG;
  return guard(b->b.ctx, [&] {
    QSF_REQUIRE(b->b.launched == 0 || !b->b.solves[s].end_row.empty(), QSF_ERR_STATE,
                "download after qsf_batch_finish (the state is mid-pipeline)");
    Grid& g = *b->b.solves[s].g;
    g.download(psi, g.m);
  });
}
int qsf_batch_start(qsf_batch* b) {
  if (!b) return QSF_ERR_ARG;
  return guard(b->b.ctx, [&] { b->b.start(); });
}
int qsf_batch_run(qsf_batch* b, int64_t k, int use_graph, int* done) {
  if (!b) return QSF_ERR_ARG;
  return guard(b->b.ctx, [&] {
    b->b.run(k, use_graph != 0);
    if (done) *done = b->b.done();
  });
}
int qsf_batch_finish(qsf_batch* b) {
  if (!b) return QSF_ERR_ARG;
  return guard(b->b.ctx, [&] {
    QSF_REQUIRE(b->b.done(), QSF_ERR_STATE, "every solve must have taken its steps");
    b->b.finish();
  });
}
int qsf_batch_info(qsf_batch* b, int s, int64_t* max_steps, int64_t* grid_size) {
  if (!b || s < 0 || s >= (int)b->b.solves.size()) return QSF_ERR_ARG;
  if (max_steps) *max_steps = b->b.solves[s].max_steps;
  if (grid_size) *grid_size = b->b.solves[s].g->m;
  return QSF_OK;
}
int qsf_batch_records(qsf_batch* b, int s, double* out, int max_rows, int* rows) {
  if (!b || s < 0 || s >= (int)b->b.solves.size()) return QSF_ERR_ARG;
  return guard(b->b.ctx, [&] {
    auto r = b->b.records(s);
    const int w = 3 + b->b.pd.nborders;
    const int n = (int)(r.size() / w);
    if (rows) *rows = n;
    if (out) std::memcpy(out, r.data(), std::min(n, max_rows) * w * sizeof(double));
  });
}
int qsf_batch_flux_integral(qsf_batch* b, int s, double* out) {
  if (!b || !out || s < 0 || s >= (int)b->b.solves.size()) return QSF_ERR_ARG;
  return guard(b->b.ctx, [&] {
    QSF_HIP(hipStreamSynchronize(b->b.st()));
    QSF_HIP(hipMemcpy(out, b->b.solves[s].fluxint.p, b->b.pd.nborders * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int qsf_prop_set_exec(qsf_prop* p, int mode) {
  if (!p || (mode != 0 && mode != 1)) return QSF_ERR_ARG;
  return guard(p->p.ctx, [&] {
    require_set(p);
    QSF_REQUIRE(mode == 1 || !(p->p.pd.mode == QSF_MODE_IM || (p->p.pd.outputs & (QSF_OUT_NORMALIZE | QSF_OUT_EDIFF))),
                QSF_ERR_UNSUPPORTED, "imaginary time, Normalize and ENERGY_DIFFERENCE run unfused");
    QSF_REQUIRE(mode == 1 || !p->p.radial_set(), QSF_ERR_UNSUPPORTED,
                "CoulombInteraction over two or more bounded coordinates runs unfused");
    QSF_REQUIRE(mode == 1 || !(p->p.set->multi && p->p.set->D == 3), QSF_ERR_UNSUPPORTED,
                "3D multigrid sets run unfused");
    p->p.flush();
    p->p.exec = mode;
  });
}

}  // extern "C"
