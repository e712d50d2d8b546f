// propagator.h — SplitPropagator, the split tables and the output list of the reference over
// the MI355X C ABI.
//
//   SplitPropagator<MODE::RE, MultiProductSplit<Split3Base<REP::X, REP::P, REP::X>, 2>, WF>
//       p({.dt = 0.02}, std::move(wf));
//   p.run(BufferedOutputs<AVG<Identity>, ZOA_FLUX_2D>{}, worker);
//
// follows propagator.h:51-499: max_steps = (ind)(maxPulseDuration/dt + 1) for RE (:153), the
// run loop of :420-462 (AT_START observables, makeStep + computeEach + postCompute while
// shouldContinue, one more step, AT_END), the worker called at the same three points. The step
// itself runs on the device as fused axis passes (qsf_prop_step).
#pragma once

#include <cstring>
#include <vector>

#include "device_grid.h"

namespace QSF {

// ---------------------------------------------------------------- splitting (splitting.h)
template <REP first, REP middle, REP last> struct Split3Base {
  static_assert(first == REP::X && middle == REP::P && last == REP::X,
                "the device step implements the X,P,X split (Split3Base<X,P,X>)");
};
// MultiProductSplit<Split3Base<X,P,X>, Order> (splitting.h:63-87); Order 1 is Strang.
template <class Base, int Order> struct MultiProductSplit {
  static constexpr int order = Order;
};
using StrangSplit = MultiProductSplit<Split3Base<REP::X, REP::P, REP::X>, 1>;

// ---------------------------------------------------------------- outputs (computations.h)
struct Identity {};
struct PotentialEnergy {};
struct KineticEnergy {};
struct Normalize {};
template <class Op> struct AVG {};
template <class... Ops> struct SUM {};
template <class Op> struct CHANGE {};
template <class... Ops> struct OPERATION {};
using ENERGY_TOTAL = SUM<AVG<PotentialEnergy>, AVG<KineticEnergy>>;        // propagator.h:26
using ENERGY_DIFFERENCE = CHANGE<SUM<AVG<PotentialEnergy>, AVG<KineticEnergy>>>;  // :27

// flux borders (fluxes/borders.h:38-257)
template <int size> struct BOX { static constexpr int kind = QSF_BORDER_BOX, n = size; };
#define QSF_REGION_BORDER(NAME, KIND) \
  struct NAME { static constexpr int kind = KIND, n = 0; };
QSF_REGION_BORDER(S2CAP, QSF_BORDER_S2CAP)
QSF_REGION_BORDER(D2CAP, QSF_BORDER_D2CAP)
QSF_REGION_BORDER(N2S, QSF_BORDER_N2S)
QSF_REGION_BORDER(N2D, QSF_BORDER_N2D)
QSF_REGION_BORDER(N2D_SYM, QSF_BORDER_N2D_SYM)
QSF_REGION_BORDER(N2D_ASYM, QSF_BORDER_N2D_ASYM)
QSF_REGION_BORDER(S2D, QSF_BORDER_S2D)
QSF_REGION_BORDER(S2D_SYM, QSF_BORDER_S2D_SYM)
QSF_REGION_BORDER(S2D_ASYM, QSF_BORDER_S2D_ASYM)
QSF_REGION_BORDER(N2T, QSF_BORDER_N2T)
QSF_REGION_BORDER(S2T, QSF_BORDER_S2T)
QSF_REGION_BORDER(D2T, QSF_BORDER_D2T)
QSF_REGION_BORDER(T2CAP, QSF_BORDER_T2CAP)
#undef QSF_REGION_BORDER
template <class... Borders> struct FLUX {};
using ZOA_FLUX_2D = FLUX<N2S, N2D_SYM, N2D_ASYM, S2D_SYM, S2D_ASYM, S2CAP, D2CAP>;  // computations.h:70
using ZOA_FLUX_2D_SIMPLE = FLUX<N2S, N2D, S2D, S2CAP, D2CAP>;                      // computations.h:69
using ZOA_FLUX_3D = FLUX<N2S, N2D, N2T, S2D, S2T, S2CAP, D2CAP, T2CAP>;

namespace detail {
template <class T> struct out_of { static void add(qsf_prop_desc&) { static_assert(sizeof(T) == 0, "output not supported on the device"); } };
template <> struct out_of<AVG<Identity>> { static void add(qsf_prop_desc& d) { d.outputs |= QSF_OUT_NORM; } };
template <> struct out_of<AVG<PotentialEnergy>> { static void add(qsf_prop_desc& d) { d.outputs |= QSF_OUT_EPOT; } };
template <> struct out_of<AVG<KineticEnergy>> { static void add(qsf_prop_desc& d) { d.outputs |= QSF_OUT_EKIN; } };
template <> struct out_of<ENERGY_TOTAL> { static void add(qsf_prop_desc& d) { d.outputs |= QSF_OUT_ETOT; } };
template <> struct out_of<ENERGY_DIFFERENCE> { static void add(qsf_prop_desc& d) { d.outputs |= QSF_OUT_EDIFF; } };
template <> struct out_of<OPERATION<Normalize>> { static void add(qsf_prop_desc& d) { d.outputs |= QSF_OUT_NORMALIZE; } };
template <class... B> struct out_of<FLUX<B...>> {
  static void add(qsf_prop_desc& d) {
    d.outputs |= QSF_OUT_FLUX;
    ((d.borders[d.nborders].kind = B::kind, d.borders[d.nborders].size = B::n, ++d.nborders), ...);
  }
};
}  // namespace detail

// BufferedOutputs<Ts...> (buffer.h:13): which observables computeEach evaluates, and how often
// a record row is kept (comp_interval). The rows are read back with SplitPropagator::records().
template <class... Ts> struct BufferedOutputs {
  int comp_interval = 1;
  void fill(qsf_prop_desc& d) const {
    (detail::out_of<Ts>::add(d), ...);
    d.comp_interval = comp_interval;
  }
};

// PropagatorBase (propagator.h:30-35)
struct PropagatorBase {
  double dt;
  ind max_steps = 0;
  double state_accuracy = 0;
};

// SplitPropagator<M, SpType, HamWF> (propagator.h:51-499)
template <MODE M, class SpType, class HamWF> struct SplitPropagator : PropagatorBase {
  HamWF wf;

  SplitPropagator(PropagatorBase pb, HamWF&& w) : PropagatorBase(pb), wf(std::move(w)) {}
  ~SplitPropagator() {
    if (h_) qsf_prop_destroy(h_);
  }
  SplitPropagator(const SplitPropagator&) = delete;
  SplitPropagator& operator=(const SplitPropagator&) = delete;

  // run(outputs, worker) (propagator.h:420-462); worker(WHEN, step, pass, wf)
  template <class OUTS, class Worker> void run(OUTS&& outputs, Worker&& worker) {
    create(outputs);
    wf.check(qsf_prop_start(h_), "qsf_prop_start");
    worker(WHEN::AT_START, step(), 0, wf);
    for (;;) {
      int more = 0;
      wf.check(qsf_prop_should_continue(h_, &more), "qsf_prop_should_continue");
      if (!more) break;
      wf.check(qsf_prop_step(h_), "qsf_prop_step");
      worker(WHEN::DURING, step(), 0, wf);
    }
    wf.check(qsf_prop_finish(h_), "qsf_prop_finish");
    worker(WHEN::AT_END, step(), 0, wf);
  }
  template <class OUTS> void run(OUTS&& outputs) {
    run(outputs, [](WHEN, ind, uind, HamWF&) {});
  }

  ind step() const {
    int64_t s = 0;
    if (h_) wf.check(qsf_prop_step_count(h_, &s), "qsf_prop_step_count");
    return s;
  }
  double time() const {
    double t = 0;
    if (h_) wf.check(qsf_prop_time(h_, &t), "qsf_prop_time");
    return t;
  }
  // one row per computeEach: step, time, norm, epot, ekin, etot, ediff, flux[borders]
  std::vector<double> records(int* width = nullptr) {
    int w = 0, n = 0;
    wf.check(qsf_prop_record_width(h_, &w), "qsf_prop_record_width");
    wf.check(qsf_prop_nrecords(h_, &n), "qsf_prop_nrecords");
    std::vector<double> out((size_t)w * n);
    if (n) wf.check(qsf_prop_records(h_, out.data(), n), "qsf_prop_records");
    if (width) *width = w;
    return out;
  }
  ind maxSteps() const {
    int64_t s = 0;
    wf.check(qsf_prop_max_steps(h_, &s), "qsf_prop_max_steps");
    return s;
  }

 private:
  template <class OUTS> void create(const OUTS& outputs) {
    if (h_) {
      qsf_prop_destroy(h_);
      h_ = nullptr;
    }
    qsf_prop_desc d;
    std::memset(&d, 0, sizeof d);
    d.mode = static_cast<int>(M);
    d.order = SpType::order;
    d.dt = dt;
    d.max_steps = max_steps;
    d.state_accuracy = state_accuracy;
    d.gauge = decltype(wf._coupling)::gauge;
    outputs.fill(d);
    auto f = wf._coupling.descs();
    wf.check(qsf_prop_create(wf.set(), &d, f.data(), (int)f.size(), &h_), "qsf_prop_create");
  }
  qsf_prop* h_ = nullptr;
};

}  // namespace QSF
