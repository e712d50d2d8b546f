// device_grid.h — the reference's wavefunction vocabulary over the MI355X C ABI.
//
// A program written against QSF composes its wavefunction from header types:
//   Schrodinger::Spin0<GridType, Potential, Coupling> wf(grid, potential, coupling);
// (hamiltonian.h:8-40). The same composition here builds plain descriptors (qsf_types.h) and
// places the grid (or the 2^D multigrid regions) on the device through libqsf_hip.so, so the
// propagator in propagator.h drives the GPU with the reference's type names unchanged.
//
// Supported instantiations are the ones the device implements: CartesianGrid /
// MultiCartesianGrid (coords.h:14-148) under CAP (absorbers.h:16-60) or the cos^(1/8) mask;
// CoulombInteraction and ReducedDimInteraction<Eberly|EckhardSacha> (potential.h:21-179);
// DipoleCoupling<VelocityGauge|LengthGauge, Field<AXIS, Envelope<Pulse>>...>
// (coupling.h:43-118, field.h:9-53, field_p.h:6-194). Errors throw qsf_error with the
// library's message (the reference aborts through FAIL(), logging.h:12,114).
#pragma once

#include <complex>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../../include/qsf_hip.h"

namespace QSF {

using ind = std::int64_t;
using uind = std::uint64_t;
using cxd = std::complex<double>;

// enums.h:15-62 (values are this library's; only the names are the reference's)
enum class MODE : int { IM = QSF_MODE_IM, RE = QSF_MODE_RE };
enum class REP : int { NONE = QSF_REP_NONE, X = QSF_REP_X, P = QSF_REP_P, BOTH = QSF_REP_BOTH };
enum class AXIS : int {
  NO = QSF_AXIS_NO, X = QSF_AXIS_X, Y = QSF_AXIS_Y, Z = QSF_AXIS_Z,
  XY = QSF_AXIS_XY, XZ = QSF_AXIS_XZ, YZ = QSF_AXIS_YZ, XYZ = QSF_AXIS_XYZ
};
enum class WHEN { AT_START, DURING, AT_END };

struct qsf_error : std::runtime_error {
  int code;
  qsf_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// One device context (HIP stream) per process and device, created on first use.
class Device {
 public:
  static qsf_ctx* get(int device = 0) {
    static std::vector<std::unique_ptr<Device>> devs;
    for (auto& d : devs)
      if (d->id_ == device) return d->ctx_;
    devs.emplace_back(new Device(device));
    return devs.back()->ctx_;
  }
  static void check(qsf_ctx* ctx, int rc, const char* what) {
    if (rc != QSF_OK) throw qsf_error(rc, std::string(what) + ": " + (ctx ? qsf_last_error(ctx) : "no context"));
  }
  ~Device() { qsf_ctx_destroy(ctx_); }

 private:
  explicit Device(int id) : id_(id) { check(nullptr, qsf_ctx_create(id, &ctx_), "qsf_ctx_create"); }
  int id_;
  qsf_ctx* ctx_ = nullptr;
};

// ---------------------------------------------------------------- grids and absorbers
// CartesianGrid<dims> (coords.h:113-122: one dx and n for every axis); MultiCartesianGrid
// holds the 2^D regions of MultiRegionsReduced (mpi_logic.h:246-263).
template <int dims, bool many = false> struct CartesianGrid_ {
  static constexpr int DIM = dims;
  static constexpr bool multigrid = many;
  double dx;
  ind n;
  CartesianGrid_(double dx_, ind n_) : dx(dx_), n(n_) {}
  void fill(qsf_grid_desc& d) const {
    d.dim = dims;
    for (int i = 0; i < 3; ++i) {
      d.n[i] = i < dims ? n : 1;
      d.dx[i] = i < dims ? dx : 1.0;
    }
    d.multigrid = many ? 1 : 0;
  }
};
template <int dims> using CartesianGrid = CartesianGrid_<dims, false>;
template <int dims> using MultiCartesianGrid = CartesianGrid_<dims, true>;

// CAP<Grid>(base, nCAP) (absorbers.h:27-37)
template <class Grid> struct CAP : Grid {
  ind nCAP;
  CAP(Grid base, ind nCAP_) : Grid(base), nCAP(nCAP_) {}
  void fill(qsf_grid_desc& d) const {
    Grid::fill(d);
    d.absorber = QSF_ABS_CAP;
    d.nCAP = nCAP;
  }
};

// cos^(1/8) edge mask over the outer `width` nodes; the whole mask is applied once per `tau`
// of propagation time (no counterpart in the reference).
template <class Grid> struct Cos18Absorber : Grid {
  ind width;
  double tau;
  Cos18Absorber(Grid base, ind width_, double tau_) : Grid(base), width(width_), tau(tau_) {}
  void fill(qsf_grid_desc& d) const {
    Grid::fill(d);
    d.absorber = QSF_ABS_COS18;
    d.nCAP = width;
    d.absorber_tau = tau;
  }
};

// ---------------------------------------------------------------- potentials (potential.h)
struct InteractionBase {
  double Ncharge, Echarge, Nsoft, Esoft;
};
struct CoulombInteraction : InteractionBase {
  explicit CoulombInteraction(InteractionBase b) : InteractionBase(b) {}
  void fill(qsf_grid_desc& d) const { set(d, QSF_POT_COULOMB); }

 protected:
  void set(qsf_grid_desc& d, int kind) const {
    d.potential = kind;
    d.Ncharge = Ncharge;
    d.Echarge = Echarge;
    d.Nsoft = Nsoft;
    d.Esoft = Esoft;
  }
};
enum class ReducedModel { Eberly, EckhardSacha };
template <ReducedModel model> struct ReducedDimInteraction : CoulombInteraction {
  explicit ReducedDimInteraction(InteractionBase b) : CoulombInteraction(b) {}
  void fill(qsf_grid_desc& d) const {
    set(d, model == ReducedModel::Eberly ? QSF_POT_EBERLY : QSF_POT_ECKHARDT_SACHA);
  }
};
struct NoInteraction {
  void fill(qsf_grid_desc& d) const { d.potential = QSF_POT_NONE; }
};

// ---------------------------------------------------------------- fields (field.h, field_p.h)
// _FieldConfig (field_p.h:6-12) in the reference's member order.
struct _FieldConfig {
  double field, omega, ncycles, FWHM_percent, phase_in_pi_units, delay_in_cycles;
};
struct SinPulse : _FieldConfig {
  static constexpr int kind = QSF_PULSE_SIN;
  SinPulse(_FieldConfig c) : _FieldConfig(c) {}
};
struct ConstantPulse : _FieldConfig {
  static constexpr int kind = QSF_PULSE_CONST;
  ConstantPulse(_FieldConfig c) : _FieldConfig(c) {}
};
template <class Pulse> struct Sin2Envelope : Pulse {
  static constexpr int env = QSF_ENV_SIN2;
  Sin2Envelope(_FieldConfig c) : Pulse(c) {}
};
template <class Pulse> struct GaussianEnvelope : Pulse {
  static constexpr int env = QSF_ENV_GAUSS;
  GaussianEnvelope(_FieldConfig c) : Pulse(c) {}
};
template <class P, class = void> struct env_of { static constexpr int value = QSF_ENV_NONE; };
template <class P> struct env_of<P, std::void_t<decltype(P::env)>> { static constexpr int value = P::env; };

// Field<AXIS, Pulses...> (field.h:9-53): every pulse acts on the same axis mask.
template <AXIS Ax, class... Pulses> struct Field {
  std::tuple<Pulses...> pulses;
  Field(Pulses... p) : pulses(std::move(p)...) {}
  void append(std::vector<qsf_field_desc>& out) const {
    std::apply([&](const auto&... p) { (out.push_back(desc(p)), ...); }, pulses);
  }

 private:
  template <class P> static qsf_field_desc desc(const P& p) {
    qsf_field_desc f{};
    f.axis = static_cast<int>(Ax);
    f.pulse = P::kind;
    f.envelope = env_of<P>::value;
    f.field = p.field;
    f.omega = p.omega;
    f.ncycles = p.ncycles;
    f.fwhm_percent = p.FWHM_percent;
    f.phase_in_pi = p.phase_in_pi_units;
    f.delay_in_cycles = p.delay_in_cycles;
    return f;
  }
};

// ---------------------------------------------------------------- couplings (coupling.h)
struct VelocityGauge { static constexpr int gauge = QSF_GAUGE_VELOCITY; };  // p·A
struct LengthGauge { static constexpr int gauge = QSF_GAUGE_LENGTH; };      // x·F
template <class Gauge, class... Fields> struct DipoleCoupling {
  static constexpr int gauge = sizeof...(Fields) ? Gauge::gauge : QSF_GAUGE_NONE;
  static constexpr REP couplesInRep = Gauge::gauge == QSF_GAUGE_LENGTH ? REP::X : REP::P;
  std::tuple<Fields...> fields;
  DipoleCoupling(Fields... f) : fields(std::move(f)...) {}
  std::vector<qsf_field_desc> descs() const {
    std::vector<qsf_field_desc> out;
    std::apply([&](const auto&... f) { (f.append(out), ...); }, fields);
    return out;
  }
};

namespace Schrodinger {

// Spin0<GType, V_Op, C_Op> (hamiltonian.h:8-97) with its grid on the device. Copying shares
// the device grid (the reference's wf copy is shallow too, propagator.h:115).
template <class GType, class V_Op, class C_Op> struct Spin0 {
  static constexpr int DIM = GType::DIM;
  static constexpr REP couplesInRep = C_Op::couplesInRep;
  GType grid;
  V_Op _potential;
  C_Op _coupling;

  Spin0(GType g, V_Op potential, C_Op coupling, int precision = QSF_FP64, int device = 0)
      : grid(g), _potential(potential), _coupling(coupling) {
    qsf_grid_desc d{};
    d.precision = precision;
    grid.fill(d);
    _potential.fill(d);
    state_ = std::make_shared<State>();
    state_->ctx = Device::get(device);
    check(qsf_set_create(state_->ctx, &d, &state_->set), "qsf_set_create");
    int count = 0;
    check(qsf_set_count(state_->set, &count), "qsf_set_count");
    state_->regions = count;
  }

  qsf_ctx* ctx() const { return state_->ctx; }
  qsf_set* set() const { return state_->set; }
  int regions() const { return state_->regions; }
  qsf_grid* region(int r = 0) const {
    qsf_grid* g = nullptr;
    check(qsf_set_grid(state_->set, r, &g), "qsf_set_grid");
    return g;
  }

  // localSize() (grid.h:359) of one region
  ind localSize(int r = 0) const {
    int64_t m = 0;
    check(qsf_grid_size(region(r), &m), "qsf_grid_size");
    return m;
  }
  // bulk operator[] (grid.h:296): the state of region r in the reference's layout
  std::vector<cxd> psi(int r = 0) const {
    std::vector<cxd> out(localSize(r));
    check(qsf_grid_download(region(r), reinterpret_cast<double*>(out.data()), (int64_t)out.size()),
          "qsf_grid_download");
    return out;
  }
  void set_psi(const std::vector<cxd>& v, int r = 0, REP rep = REP::X) {
    check(qsf_grid_upload(region(r), reinterpret_cast<const double*>(v.data()), (int64_t)v.size(),
                          static_cast<int>(rep)),
          "qsf_grid_upload");
  }
  // fourier<R>() (grid.h:369-378)
  template <REP R> void fourier(int r = 0) { check(qsf_fourier(region(r), static_cast<int>(R)), "qsf_fourier"); }
  // evolve<M,R>(delta) (grid.h:394-423) with explicit coupling values per axis
  template <MODE M, REP R> void evolve(double delta, const double coupling[3], int r = 0) {
    check(qsf_evolve(region(r), static_cast<int>(M), static_cast<int>(R), delta, coupling,
                     static_cast<int>(couplesInRep)),
          "qsf_evolve");
  }
  // average<R, Op>() (grid.h:425-441): op 0 Identity, 1 PotentialEnergy, 2 KineticEnergy
  double average(REP rep, int op, int r = 0) {
    double v = 0;
    check(qsf_average(region(r), static_cast<int>(rep), op, &v), "qsf_average");
    return v;
  }
  void postCompute() { check(qsf_set_post_compute(state_->set), "qsf_set_post_compute"); }

  // save() of a state in X rep as a complex `.psi` file (wf/grid.h:1025-1071): header
  // u8 dim, u8 rep, u8 complex, i64 downscale, i64 n[dim], f64 dx[dim], u8 bounded[dim], then the
  // row-major complex128 values. `path` gets "_repX.psi" appended; returns the file name.
  std::string save(const std::string& path, int r = 0) const {
    const std::string name = path + "_repX.psi";
    std::FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) throw qsf_error(QSF_ERR_ARG, "cannot open " + name);
    const unsigned char head[3] = {(unsigned char)DIM, (unsigned char)QSF_REP_X, 1};
    const int64_t downscale = 1;
    std::fwrite(head, 1, 3, f);
    std::fwrite(&downscale, 8, 1, f);
    for (int i = 0; i < DIM; ++i) { const int64_t n = grid.n; std::fwrite(&n, 8, 1, f); }
    for (int i = 0; i < DIM; ++i) { const double dx = grid.dx; std::fwrite(&dx, 8, 1, f); }
    for (int i = 0; i < DIM; ++i) { const unsigned char b = 1; std::fwrite(&b, 1, 1, f); }
    const std::vector<cxd> v = psi(r);
    std::fwrite(v.data(), sizeof(cxd), v.size(), f);
    std::fclose(f);
    return name;
  }

  void check(int rc, const char* what) const { Device::check(state_ ? state_->ctx : nullptr, rc, what); }

 private:
  struct State {
    qsf_ctx* ctx = nullptr;
    qsf_set* set = nullptr;
    int regions = 0;
    ~State() {
      if (set) qsf_set_destroy(set);
    }
  };
  std::shared_ptr<State> state_;
};

}  // namespace Schrodinger
}  // namespace QSF
