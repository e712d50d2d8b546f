// physics.hpp — host-side FP64 physics of the step: fields and dipole couplings, potential
// closed forms, absorber node profiles and flux-border vectors. Everything here is scalar
// host code; its outputs are the per-substep coupling values and the tables the kernels read.
#pragma once
#include <cmath>
#include <vector>

#include "../../include/qsf_types.h"

namespace qsf {
namespace phys {

constexpr double kPi = M_PI;  // basics/constants.h:38

// Power as the reference's recursive product (math_basic.h:20-29): b*(b*(b*b)) for 4.
template <typename T> inline T ipow(T b, int e) {
  if (e == 0) return T(1);
  if (e == 1) return b;
  if (e == 2) return b * b;
  return b * ipow(b, e - 1);
}

// One Field<Axis, Envelope<Pulse>> (field.h:9-53, field_p.h:15-194).
class Field {
 public:
  explicit Field(const qsf_field_desc& d) : d_(d) {
    phi_ = d.phase_in_pi * kPi;
    pulse_time_ = 2.0 * kPi * d.ncycles / d.omega;
    const double cycle = 2.0 * kPi / d.omega;
    (void)cycle;
    delay_ = d.delay_in_cycles * 2.0 * kPi / d.omega;
    end_ = d.delay_in_cycles * 2.0 * kPi / d.omega + pulse_time_;
    env_ = d.omega / (2.0 * d.ncycles);
    if (d.envelope == QSF_ENV_SIN) sin_pow_ = std::log(2) / std::log(1 / std::sin(0.5 * kPi * (1 - d.fwhm_percent)));
    if (d.envelope == QSF_ENV_GAUSS) gauss_pow_ = -4.0 * std::log(2) / ipow(d.fwhm_percent * pulse_time_, 2);
    if (d.envelope == QSF_ENV_RAMP) {
      left_ = (-2.0 * (-1.0 + d.fwhm_percent)) * pulse_time_;
      right_ = (-1.0 + 2.0 * d.fwhm_percent) * pulse_time_;
      ramp_ = 0.5 * (1.0 / (pulse_time_ * (1.0 - d.fwhm_percent)));
    }
  }
  int axis() const { return d_.axis; }
  double end_time() const { return end_; }
  // Field value at time t (field.h:23-35): zero outside [delay, end].
  double at(double t) const { return !(t > end_ || t < delay_) ? shape(t - delay_) : 0.0; }

 private:
  double carrier(double t) const {
    if (d_.pulse == QSF_PULSE_SIN) return d_.field * std::sin(d_.omega * t + phi_);
    if (d_.pulse == QSF_PULSE_CONST) return d_.field;
    if (d_.pulse == QSF_PULSE_CHEMPHYS) {
      const double em = d_.omega / (2.0 * d_.ncycles);
      const double tmp = d_.omega * t + phi_ - kPi * d_.ncycles;
      return d_.field * (std::sin(em * t) * std::cos(tmp) + std::cos(em * t) * std::sin(tmp) / d_.ncycles);
    }
    return 0.0;
  }
  double shape(double t) const {
    const bool bare = d_.pulse == QSF_PULSE_EMPTY;
    switch (d_.envelope) {
      case QSF_ENV_SIN2: {
        const double s2 = ipow(std::sin(env_ * t), 2);
        return bare ? s2 : carrier(t) * s2;
      }
      case QSF_ENV_CHEMPHYS: return bare ? std::sin(env_ * t) : carrier(t) * std::sin(env_ * t);
      case QSF_ENV_SIN: {
        const double e = std::pow(std::sin(env_ * t), sin_pow_);
        return bare ? e : carrier(t) * e;
      }
      case QSF_ENV_GAUSS: {
        const double e = std::exp(gauss_pow_ * ipow(t - 0.5 * pulse_time_, 2));
        return bare ? e : carrier(t) * e;
      }
      case QSF_ENV_RAMP: {
        double e = 1.0;
        if (t < left_) e = ramp_ * t;
        else if (t > right_) e = ramp_ * (pulse_time_ - t);
        if (bare) return e;
        if (t < left_) return carrier(t) * ramp_ * t;
        if (t > right_) return carrier(t) * ramp_ * (pulse_time_ - t);
        return carrier(t) * 1.0;
      }
      default: return carrier(t);
    }
  }
  qsf_field_desc d_;
  double phi_, pulse_time_, delay_, end_, env_;
  double sin_pow_ = 0, gauss_pow_ = 0, left_ = 0, right_ = 0, ramp_ = 0;
};

// DipoleCoupling over a set of fields. Velocity gauge promotes each field to A(t) with the
// reference's running right-Riemann sum (PromoteFieldToA field.h:103-123); the length gauge
// couples F(t) in X (new here; the reference only declares the tag, coupling.h:9-11).
class Coupling {
 public:
  Coupling() = default;
  Coupling(int gauge, const qsf_field_desc* f, int nf) : gauge_(gauge) {
    for (int i = 0; i < nf; ++i) {
      fields_.emplace_back(f[i]);
      last_.push_back(0.0);
      prev_time_.push_back(0.0);
    }
  }
  // representation the coupling acts in (no gauge = the reference's default velocity gauge)
  int rep() const { return gauge_ == QSF_GAUGE_LENGTH ? QSF_REP_X : QSF_REP_P; }
  double max_pulse_duration() const {
    double mx = 0.0;
    for (auto& f : fields_) mx = std::max(mx, f.end_time());
    return mx;
  }
  void precalc(double t) {
    for (size_t i = 0; i < fields_.size(); ++i) {
      const double F = fields_[i].at(t);
      if (gauge_ == QSF_GAUGE_LENGTH) {
        last_[i] = F;
      } else {
        last_[i] = last_[i] + (prev_time_[i] - t) * F;
        prev_time_[i] = t;
      }
    }
  }
  // sum of the fields whose axis mask meets `ax` (coupling.h:63-67), folded from the right
  double on(int ax) const {
    double acc = 0.0;
    for (size_t k = fields_.size(); k-- > 0;) {
      const double v = (fields_[k].axis() & ax) ? last_[k] : 0;
      acc = (k + 1 == fields_.size()) ? v : v + acc;
    }
    return acc;
  }
  void values(double* c3) const {
    for (int d = 0; d < 3; ++d) c3[d] = on(1 << d);
  }

 private:
  int gauge_ = QSF_GAUGE_NONE;
  std::vector<Field> fields_;
  std::vector<double> last_, prev_time_;
};

// ---- flux borders (fluxes/borders_g.h, borders.h:30-257, borders_t.h:1-50) ----
namespace bord {
constexpr long kNS = 25, kSD = 14, kDT = 10, kCAP = 40;  // borders.h:30-36 (nodes)
inline double sgn(long x) { return (double)((x > 0) - (x < 0)); }
inline bool in(long i, long lo, long hi) { return i >= lo && i < hi; }
inline bool in2(long i, long lo, long hi) { return in(i, lo, hi) || in(i, -hi, -lo); }
inline double pt(long i, long at) { return i == at - 1 ? sgn(i) : 0; }
inline double pt2(long i, long at) { return (i == at - 1 || i == -at) ? sgn(i) : 0; }
inline double seg(long i, long j, long at, long lo, long hi) { return in(i, lo, hi) ? pt(j, at) : 0; }
inline double seg22(long i, long j, long at, long lo, long hi) { return in2(i, lo, hi) ? pt2(j, at) : 0; }
inline double pat(long i, long j, long k, long at, long il, long ir, long jl, long jr) {
  return in2(i, il, ir) ? seg22(j, k, at, jl, jr) : 0;
}

struct Vec {
  double c[3] = {0, 0, 0};
};

// borVec<D>(a...) normalises when |a| > 0.001 (borders_t.h:7-13)
inline Vec make(int D, double a, double b = 0, double c = 0) {
  Vec v;
  v.c[0] = a;
  v.c[1] = b;
  v.c[2] = c;
  double nn = D == 1 ? std::sqrt((a * a) + 0.0)
            : D == 2 ? std::sqrt((a * a) + ((b * b) + 0.0))
                     : std::sqrt((a * a) + ((b * b) + ((c * c) + 0.0)));
  if (nn > 0.001)
    for (int i = 0; i < D; ++i) v.c[i] /= nn;
  for (int i = D; i < 3; ++i) v.c[i] = 0;
  return v;
}

// Border vector of a family at centred indices (i, j, k).
inline Vec border(int kind, long s, int D, long i, long j, long k) {
  // 3D families are the symmetrised patch sums of the reference, one per component
  auto tri = [&](long at, long il, long ir, long jl, long jr, bool sym) {
    if (!sym)
      return make(3, pat(j, k, i, at, il, ir, jl, jr), pat(k, i, j, at, il, ir, jl, jr), pat(i, j, k, at, il, ir, jl, jr));
    return make(3, pat(j, k, i, at, il, ir, jl, jr) + pat(k, j, i, at, il, ir, jl, jr),
                pat(k, i, j, at, il, ir, jl, jr) + pat(i, k, j, at, il, ir, jl, jr),
                pat(i, j, k, at, il, ir, jl, jr) + pat(j, i, k, at, il, ir, jl, jr));
  };
  auto duo = [&](long at, long lo, long hi) { return make(2, seg22(j, i, at, lo, hi), seg22(i, j, at, lo, hi)); };
  switch (kind) {
    case QSF_BORDER_BOX:
      if (D == 1) return make(1, pt2(i, s));
      if (D == 2) return duo(s, 0, s);
      return tri(s, 0, s, 0, s, false);
    case QSF_BORDER_S2CAP: return D == 2 ? duo(kCAP, 0, kSD) : tri(kCAP, 0, kDT, 0, kSD, true);
    case QSF_BORDER_D2CAP: return D == 2 ? duo(kCAP, kSD, kCAP) : tri(kCAP, kSD, kCAP, 0, kDT, true);
    case QSF_BORDER_N2S: return D == 2 ? duo(kNS, 0, kSD) : tri(kNS, 0, kSD, 0, kDT - 1, true);
    case QSF_BORDER_N2D: return D == 2 ? duo(kSD, kSD, kNS) : tri(kSD, kSD, kNS, 0, kDT - 1, true);
    case QSF_BORDER_N2D_SYM:
      return make(2, seg(j + 1, i, kSD, kSD, kNS + 1) - seg(-j, -i - 1, kSD, kSD, kNS + 1),
                  seg(i + 1, j, kSD, kSD, kNS + 1) - seg(-i, -j - 1, kSD, kSD, kNS + 1));
    case QSF_BORDER_N2D_ASYM:
      return make(2, -seg(j + 1, -i - 1, kSD, kSD, kNS + 1) + seg(-j, i, kSD, kSD, kNS + 1),
                  -seg(i + 1, -j - 1, kSD, kSD, kNS + 1) + seg(-i, j, kSD, kSD, kNS + 1));
    case QSF_BORDER_S2D: return D == 2 ? duo(kSD, kNS, kCAP) : tri(kSD, kNS, kCAP, 0, kDT, true);
    case QSF_BORDER_S2D_SYM:
      return make(2, seg(j, i, kSD, kNS, kCAP) - seg(-j - 1, -i - 1, kSD, kNS, kCAP),
                  seg(i, j, kSD, kNS, kCAP) - seg(-i - 1, -j - 1, kSD, kNS, kCAP));
    case QSF_BORDER_S2D_ASYM:
      return make(2, -seg(j, -i - 1, kSD, kNS, kCAP) + seg(-j - 1, i, kSD, kNS, kCAP),
                  -seg(i, -j - 1, kSD, kNS, kCAP) + seg(-i - 1, j, kSD, kNS, kCAP));
    case QSF_BORDER_N2T: return tri(kDT, kDT - 1, kNS, kDT - 1, kSD, true);
    case QSF_BORDER_S2T: return tri(kDT, kNS, kCAP, kDT, kSD, true);
    case QSF_BORDER_D2T: return tri(kDT, kSD, kCAP, kSD, kCAP, false);
    case QSF_BORDER_T2CAP: return tri(kCAP, kDT, kCAP, kDT, kCAP, false);
    default: return Vec{};
  }
}
}  // namespace bord

}  // namespace phys
}  // namespace qsf
