// he2d.cpp — a reference-style program on the device: the 2D two-electron reduced model
// (Eberly interaction, CAP, velocity gauge, MultiProductSplit order 2, ZOA fluxes) written
// with the reference's type names, propagated on the MI355X through libqsf_hip.so.
//
// usage: he2d <n> <steps> [multigrid 0|1] [save_path] — prints one CSV row per computeEach
//        (and, with save_path, writes the final state as <save_path>_repX.psi):
//        step,time,norm,epot,ekin,etot,ediff,flux...
// The initial state is a seeded Gaussian wave packet so a caller can rebuild it.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../qsf/qsf.h"

using namespace QSF;

template <class Grid> static int run(Grid grid, long n, long steps, const char* save_path) {
  using Pot = ReducedDimInteraction<ReducedModel::Eberly>;
  using Pulse = Field<AXIS::XY, Sin2Envelope<SinPulse>>;
  using Coupling = DipoleCoupling<VelocityGauge, Pulse>;
  using WF = Schrodinger::Spin0<CAP<Grid>, Pot, Coupling>;
  const double dx = 0.3;
  WF wf(CAP<Grid>(grid, n / 8), Pot({2.0, -1.0, 0.5, 0.33}),
        Coupling(Pulse(Sin2Envelope<SinPulse>({0.12, 0.057, 3, 0, 0, 0}))));
  std::vector<cxd> psi(wf.localSize());
  const double xmin = -0.5 * dx * (n - 1);
  for (long i = 0; i < n; ++i)
    for (long j = 0; j < n; ++j) {
      const double x = xmin + i * dx, y = xmin + j * dx;
      psi[i * n + j] = std::exp(-0.5 * ((x - 1) * (x - 1) + (y + 0.5) * (y + 0.5))) * std::polar(1.0, 0.3 * x);
    }
  wf.set_psi(psi);
  SplitPropagator<MODE::RE, MultiProductSplit<Split3Base<REP::X, REP::P, REP::X>, 2>, WF> p({0.02}, std::move(wf));
  long done = 0;
  struct Stop {};
  try {
    p.run(BufferedOutputs<AVG<Identity>, AVG<PotentialEnergy>, ZOA_FLUX_2D>{},
          [&](WHEN when, ind, uind, WF&) {
            if (when == WHEN::DURING && ++done >= steps) throw Stop{};
          });
  } catch (Stop&) {
  }
  if (save_path) p.wf.save(save_path);
  int w = 0;
  auto rec = p.records(&w);
  for (size_t r = 0; r < rec.size() / w; ++r) {
    for (int c = 0; c < w; ++c) std::printf(c ? ",%.17g" : "%.17g", rec[r * w + c]);
    std::printf("\n");
  }
  return 0;
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 64;
  const long steps = argc > 2 ? std::atol(argv[2]) : 10;
  const int multigrid = argc > 3 ? std::atoi(argv[3]) : 0;
  const char* save_path = argc > 4 ? argv[4] : nullptr;
  try {
    if (multigrid) return run(MultiCartesianGrid<2>(0.3, n), n, steps, save_path);
    return run(CartesianGrid<2>(0.3, n), n, steps, save_path);
  } catch (const qsf_error& e) {
    std::fprintf(stderr, "qsf error %d: %s\n", e.code, e.what());
    return 1;
  }
}
