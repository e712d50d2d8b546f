// qsf_oracle.cpp — CPU FP64 restatement of QSF's split-operator step.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU baseline. The
// product under qsf_amd/ never links, loads or calls this file.
//
// It restates the reference (a header-only C++17/MPI/FFTW library that cannot be built in
// this image: fftw3.h is absent) on one process. Every block cites the reference file:line it
// follows. The FFT is this file's own mixed-radix transform (the reference calls FFTW): it is
// pinned against a long-double brute-force DFT.
//
// Parity pins for every function live in tests/test_oracle_*.py.

#include "qsf_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <vector>

#include <omp.h>

namespace orc {

using cxd = std::complex<double>;
using ind = ptrdiff_t;
static constexpr double pi = M_PI;   // basics/constants.h:38

// basics/math_basic.h:14
inline double Sign(int x) { return (double)((x > 0) - (x < 0)); }
// basics/math_basic.h:20-29 (recursive product: b*(b*(b*b)) for pow=4)
template <typename T> T Power(T base, ind pow) {
  if (pow == 0) return 1;
  if (pow == 1) return base;
  if (pow == 2) return base * base;
  return base * Power(base, pow - 1);
}

// ----------------------------------------------------------------------------------------
// FFT. Stands in for FFTW's c2c transforms (wf/grid.h:251-274, 369-378): forward = sign -1,
// e^{-2πi jk/n}, unnormalised. Recursive decimation in time over radices 4,2,3,5,7 and a
// direct DFT for any other prime factor.
// ----------------------------------------------------------------------------------------
struct FFT1D {
  ind n;
  std::vector<cxd> root;  // root[j] = exp(-2πi j / n)
  explicit FFT1D(ind n_) : n(n_), root(n_) {
    for (ind j = 0; j < n; ++j) {
      double a = -2.0 * pi * double(j) / double(n);
      root[j] = cxd(std::cos(a), std::sin(a));
    }
  }
  cxd w(ind j, int sign) const {  // exp(sign 2πi j/n)
    cxd r = root[((j % n) + n) % n];
    return sign < 0 ? r : std::conj(r);
  }
  static ind pick_radix(ind m) {
    if (m % 4 == 0 && m != 8) return 4;
    if (m % 2 == 0) return 2;
    for (ind r : {3, 5, 7})
      if (m % r == 0) return r;
    for (ind r = 11; r <= 13; r += 2)
      if (m % r == 0) return r;
    return m;  // prime length beyond 13: direct DFT (m == 1 branch)
  }
  // out[0..len) = DFT of in[0], in[s], ..., in[(len-1)s]; scale = n/len for the root table.
  void rec(const cxd* in, cxd* out, ind len, ind s, int sign) const {
    if (len == 1) { out[0] = in[0]; return; }
    ind r = pick_radix(len);
    ind m = len / r;
    ind scale = n / len;
    if (m == 1) {  // direct DFT of length r
      for (ind p = 0; p < r; ++p) {
        cxd acc = 0;
        for (ind q = 0; q < r; ++q) acc += in[q * s] * w(scale * ((p * q) % r), sign);
        out[p] = acc;
      }
      return;
    }
    for (ind q = 0; q < r; ++q) rec(in + q * s, out + q * m, m, s * r, sign);
    if (r == 2) {  // X[k] = A[k] + w^k B[k], X[k+m] = A[k] - w^k B[k]
      for (ind k = 0; k < m; ++k) {
        cxd a = out[k], b = out[m + k] * w(scale * k, sign);
        out[k] = a + b;
        out[m + k] = a - b;
      }
      return;
    }
    if (r == 4) {
      const cxd j = sign < 0 ? cxd(0, -1) : cxd(0, 1);  // w_4
      for (ind k = 0; k < m; ++k) {
        cxd a = out[k], b = out[m + k] * w(scale * k, sign), c = out[2 * m + k] * w(scale * 2 * k, sign),
            d = out[3 * m + k] * w(scale * 3 * k, sign);
        cxd s0 = a + c, s1 = a - c, s2 = b + d, s3 = j * (b - d);
        out[k] = s0 + s2;
        out[m + k] = s1 + s3;
        out[2 * m + k] = s0 - s2;
        out[3 * m + k] = s1 - s3;
      }
      return;
    }
    cxd t[16], y[16];
    for (ind k = 0; k < m; ++k) {
      for (ind q = 0; q < r; ++q) t[q] = out[q * m + k] * w(scale * q * k, sign);
      for (ind p = 0; p < r; ++p) {
        cxd acc = 0;
        for (ind q = 0; q < r; ++q) acc += t[q] * w(scale * m * ((q * p) % r), sign);
        y[p] = acc;
      }
      for (ind p = 0; p < r; ++p) out[k + m * p] = y[p];
    }
  }
  void exec(cxd* data, int sign) const {
    std::vector<cxd> in(data, data + n);
    rec(in.data(), data, n, 1, sign);
  }
};

// FFT along `axis` of a row-major array with shape sh[0..2] (unused dims = 1).
static void fft_axis(cxd* a, const ind* sh, int axis, int sign) {
  ind n = sh[axis];
  if (n == 1) return;
  FFT1D f(n);
  ind inner = 1;
  for (int d = axis + 1; d < 3; ++d) inner *= sh[d];
  ind outer = 1;
  for (int d = 0; d < axis; ++d) outer *= sh[d];
  // lines are independent: threads split them, each line's arithmetic is unchanged
#pragma omp parallel
  {
    std::vector<cxd> line(n);
#pragma omp for schedule(static)
    for (ind q = 0; q < outer * inner; ++q) {
      const ind o = q / inner, i = q % inner;
      cxd* base = a + o * n * inner + i;
      for (ind k = 0; k < n; ++k) line[k] = base[k * inner];
      f.exec(line.data(), sign);
      for (ind k = 0; k < n; ++k) base[k * inner] = line[k];
    }
  }
}

// Swap axes 0 and 1 of a row-major (n0, n1, n2) array -> (n1, n0, n2).
// This is FFTW_MPI_TRANSPOSED_OUT / _IN's P-space layout (wf/grid.h:106-112, 238-248).
static void swap01(std::vector<cxd>& a, ind n0, ind n1, ind n2) {
  std::vector<cxd> b(a.size());
  for (ind i = 0; i < n0; ++i)
    for (ind j = 0; j < n1; ++j)
      for (ind k = 0; k < n2; ++k) b[(j * n0 + i) * n2 + k] = a[(i * n1 + j) * n2 + k];
  a.swap(b);
}

// ----------------------------------------------------------------------------------------
// Fields (field_p.h, field.h) and the dipole couplings (coupling.h).
// ----------------------------------------------------------------------------------------
struct Pulse {  // PulsePrototype field_p.h:15-34 + one shape/envelope
  qsf_field_desc c;
  double phi, pulse_time, onecycle_time;
  double env_mult = 0, sin_pow = 0, gauss_pow = 0, leftEdge = 0, rightEdge = 0, ramp = 0;
  explicit Pulse(const qsf_field_desc& d) : c(d) {
    phi = c.phase_in_pi * pi;                       // field_p.h:31
    pulse_time = 2.0 * pi * c.ncycles / c.omega;    // field_p.h:32
    onecycle_time = 2.0 * pi / c.omega;             // field_p.h:33
    switch (c.envelope) {
      case QSF_ENV_SIN2: case QSF_ENV_CHEMPHYS:
        env_mult = c.omega / (2.0 * c.ncycles);     // field_p.h:108, 91
        break;
      case QSF_ENV_SIN:                             // field_p.h:161-163
        env_mult = c.omega / (2.0 * c.ncycles);
        sin_pow = std::log(2) / std::log(1 / std::sin(0.5 * pi * (1 - c.fwhm_percent)));
        break;
      case QSF_ENV_GAUSS:                           // field_p.h:183
        gauss_pow = -4.0 * std::log(2) / Power(c.fwhm_percent * pulse_time, 2);
        break;
      case QSF_ENV_RAMP:                            // field_p.h:127-130
        leftEdge = (-2.0 * (-1.0 + c.fwhm_percent)) * pulse_time;
        rightEdge = (-1.0 + 2.0 * c.fwhm_percent) * pulse_time;
        ramp = 0.5 * (1.0 / (pulse_time * (1.0 - c.fwhm_percent)));
        break;
      default: break;
    }
  }
  double carrier(double t) const {
    switch (c.pulse) {
      case QSF_PULSE_SIN: return c.field * std::sin(c.omega * t + phi);   // field_p.h:54
      case QSF_PULSE_CONST: return c.field;                               // field_p.h:64
      case QSF_PULSE_CHEMPHYS: {                                          // field_p.h:74-78
        double em = c.omega / (2.0 * c.ncycles);
        double tmp = c.omega * t + phi - pi * c.ncycles;
        return c.field * (std::sin(em * t) * std::cos(tmp) + std::cos(em * t) * std::sin(tmp) / c.ncycles);
      }
      default: return 0.0;
    }
  }
  double operator()(double t) const {
    bool empty = c.pulse == QSF_PULSE_EMPTY;
    switch (c.envelope) {
      case QSF_ENV_SIN2: {                                                // field_p.h:111-115
        double s = Power(std::sin(env_mult * t), 2);
        return empty ? s : carrier(t) * s;
      }
      case QSF_ENV_CHEMPHYS:                                              // field_p.h:92-96
        return empty ? std::sin(env_mult * t) : carrier(t) * std::sin(env_mult * t);
      case QSF_ENV_SIN:                                                   // field_p.h:166-172
        return empty ? std::pow(std::sin(env_mult * t), sin_pow)
                     : carrier(t) * std::pow(std::sin(env_mult * t), sin_pow);
      case QSF_ENV_GAUSS:                                                 // field_p.h:187-193
        return empty ? std::exp(gauss_pow * Power(t - 0.5 * pulse_time, 2))
                     : carrier(t) * std::exp(gauss_pow * Power(t - 0.5 * pulse_time, 2));
      case QSF_ENV_RAMP:                                                  // field_p.h:134-148
        if (empty) {
          if (t < leftEdge) return ramp * t;
          if (t > rightEdge) return ramp * (pulse_time - t);
          return 1.0;
        }
        if (t < leftEdge) return carrier(t) * ramp * t;
        if (t > rightEdge) return carrier(t) * ramp * (pulse_time - t);
        return carrier(t) * 1.0;
      default: return carrier(t);
    }
  }
};

struct FieldObj {  // Field<Ax, Pulse> (field.h:9-53), one pulse per field
  Pulse p;
  int axis;
  double delay, endtime, lastVal = 0;
  double prevVal = 0, prevTime = 0;  // PromoteFieldToA state (field.h:111-112)
  explicit FieldObj(const qsf_field_desc& d) : p(d), axis(d.axis) {
    delay = p.c.delay_in_cycles * 2.0 * pi / p.c.omega;                  // field.h:48
    endtime = p.c.delay_in_cycles * 2.0 * pi / p.c.omega + p.pulse_time; // field.h:49
  }
  double field(double t) {  // field.h:23-35
    lastVal = 0;
    lastVal += (!(t > endtime || t < delay) ? p(t - delay) : 0.0);
    return lastVal;
  }
  double promoted(double t) {  // PromoteFieldToA::operator() field.h:114-123
    prevVal = lastVal;
    field(t);
    lastVal = prevVal + (prevTime - t) * lastVal;
    prevTime = t;
    return lastVal;
  }
};

struct Coupling {  // CouplingBase<DipoleApprox,...> coupling.h:43-78 (+ length gauge, new)
  int gauge = QSF_GAUGE_NONE;
  std::vector<FieldObj> fields;
  // No gauge = the reference's default DipoleCoupling<VelocityGauge> with no fields
  // (hamiltonian.h:98): couples in P, so the timer still advances on P substeps.
  int couplesInRep() const { return gauge == QSF_GAUGE_LENGTH ? QSF_REP_X : QSF_REP_P; }
  double maxPulseDuration() const {  // coupling.h:54 with Field::maxPulseDuration field.h:37-44
    double mx = 0.0;
    for (auto& f : fields) mx = std::max(mx, f.endtime);
    return mx;
  }
  void precalc(double t) {  // coupling.h:56-59
    for (auto& f : fields) {
      if (gauge == QSF_GAUGE_VELOCITY) f.promoted(t);
      else f.field(t);
    }
  }
  double operator[](int ax) const {  // coupling.h:63-67 (unary right fold f0 + (f1 + f2))
    double acc = 0.0;
    bool first = true;
    for (ind i = (ind)fields.size() - 1; i >= 0; --i) {
      double v = (fields[i].axis & ax) ? fields[i].lastVal : 0;
      acc = first ? v : v + acc;
      first = false;
    }
    return acc;
  }
};

// ----------------------------------------------------------------------------------------
// Flux borders (fluxes/borders_g.h, borders.h, borders_t.h).
// ----------------------------------------------------------------------------------------
static const ind p_NS = (ind)(12.5 / 0.5), p_SD = (ind)(7. / 0.5), p_DT = (ind)(5. / 0.5),
                 p_CAP = (ind)(20.0 / 0.5);  // borders.h:30-36
inline bool Rn(ind i, ind il, ind ir) { return (i >= il && i < ir); }
inline bool Rn2(ind i, ind il, ind ir) { return Rn(i, il, ir) || Rn(i, -ir, -il); }
inline double PointB(ind i, ind iat) { return i == iat - 1 ? Sign(i) : 0; }
inline double PointB2(ind i, ind iat) { return i == iat - 1 || i == -iat ? Sign(i) : 0; }
inline double SegmentB(ind i, ind j, ind jat, ind il, ind ir) { return Rn(i, il, ir) ? PointB(j, jat) : 0; }
inline double SegmentBi2j2(ind i, ind j, ind jat, ind il, ind ir) { return Rn2(i, il, ir) ? PointB2(j, jat) : 0; }
inline double PatchBi2j2k2(ind i, ind j, ind k, ind kat, ind il, ind ir, ind jl, ind jr) {
  return Rn2(i, il, ir) ? SegmentBi2j2(j, k, kat, jl, jr) : 0;
}

// borVec constructor (borders_t.h:7-13): normalise when the norm exceeds 0.001.
static std::array<double, 3> borvec(int D, double a0, double a1 = 0, double a2 = 0) {
  std::array<double, 3> v{a0, a1, a2};
  double norm;
  if (D == 1) norm = std::sqrt((a0 * a0) + 0.0);
  else if (D == 2) norm = std::sqrt((a0 * a0) + ((a1 * a1) + 0.0));
  else norm = std::sqrt((a0 * a0) + ((a1 * a1) + ((a2 * a2) + 0.0)));
  if (norm > 0.001)
    for (int i = 0; i < D; ++i) v[i] /= norm;
  for (int i = D; i < 3; ++i) v[i] = 0;
  return v;
}

static std::array<double, 3> border(int kind, int size, int D, ind i, ind j = 0, ind k = 0) {
  const ind s = size;
  auto P3 = [](ind a, ind b, ind c, ind kat, ind il, ind ir, ind jl, ind jr) {
    return PatchBi2j2k2(a, b, c, kat, il, ir, jl, jr);
  };
  switch (kind) {
    case QSF_BORDER_BOX:  // borders.h:38-71
      if (D == 1) return borvec(1, PointB2(i, s));
      if (D == 2) return borvec(2, SegmentBi2j2(j, i, s, 0, s), SegmentBi2j2(i, j, s, 0, s));
      return borvec(3, P3(j, k, i, s, 0, s, 0, s), P3(k, i, j, s, 0, s, 0, s), P3(i, j, k, s, 0, s, 0, s));
    case QSF_BORDER_S2CAP:  // borders.h:74-91
      if (D == 2) return borvec(2, SegmentBi2j2(j, i, p_CAP, 0, p_SD), SegmentBi2j2(i, j, p_CAP, 0, p_SD));
      return borvec(3, P3(j, k, i, p_CAP, 0, p_DT, 0, p_SD) + P3(k, j, i, p_CAP, 0, p_DT, 0, p_SD),
                    P3(k, i, j, p_CAP, 0, p_DT, 0, p_SD) + P3(i, k, j, p_CAP, 0, p_DT, 0, p_SD),
                    P3(i, j, k, p_CAP, 0, p_DT, 0, p_SD) + P3(j, i, k, p_CAP, 0, p_DT, 0, p_SD));
    case QSF_BORDER_D2CAP:  // borders.h:93-110
      if (D == 2) return borvec(2, SegmentBi2j2(j, i, p_CAP, p_SD, p_CAP), SegmentBi2j2(i, j, p_CAP, p_SD, p_CAP));
      return borvec(3, P3(j, k, i, p_CAP, p_SD, p_CAP, 0, p_DT) + P3(k, j, i, p_CAP, p_SD, p_CAP, 0, p_DT),
                    P3(k, i, j, p_CAP, p_SD, p_CAP, 0, p_DT) + P3(i, k, j, p_CAP, p_SD, p_CAP, 0, p_DT),
                    P3(i, j, k, p_CAP, p_SD, p_CAP, 0, p_DT) + P3(j, i, k, p_CAP, p_SD, p_CAP, 0, p_DT));
    case QSF_BORDER_N2S:  // borders.h:111-127
      if (D == 2) return borvec(2, SegmentBi2j2(j, i, p_NS, 0, p_SD), SegmentBi2j2(i, j, p_NS, 0, p_SD));
      return borvec(3, P3(j, k, i, p_NS, 0, p_SD, 0, p_DT - 1) + P3(k, j, i, p_NS, 0, p_SD, 0, p_DT - 1),
                    P3(k, i, j, p_NS, 0, p_SD, 0, p_DT - 1) + P3(i, k, j, p_NS, 0, p_SD, 0, p_DT - 1),
                    P3(i, j, k, p_NS, 0, p_SD, 0, p_DT - 1) + P3(j, i, k, p_NS, 0, p_SD, 0, p_DT - 1));
    case QSF_BORDER_N2D:  // borders.h:128-144
      if (D == 2) return borvec(2, SegmentBi2j2(j, i, p_SD, p_SD, p_NS), SegmentBi2j2(i, j, p_SD, p_SD, p_NS));
      return borvec(3, P3(j, k, i, p_SD, p_SD, p_NS, 0, p_DT - 1) + P3(k, j, i, p_SD, p_SD, p_NS, 0, p_DT - 1),
                    P3(k, i, j, p_SD, p_SD, p_NS, 0, p_DT - 1) + P3(i, k, j, p_SD, p_SD, p_NS, 0, p_DT - 1),
                    P3(i, j, k, p_SD, p_SD, p_NS, 0, p_DT - 1) + P3(j, i, k, p_SD, p_SD, p_NS, 0, p_DT - 1));
    case QSF_BORDER_N2D_SYM:  // borders.h:145-155 (2D only)
      return borvec(2, SegmentB(j + 1, i, p_SD, p_SD, p_NS + 1) - SegmentB(-j, -i - 1, p_SD, p_SD, p_NS + 1),
                    SegmentB(i + 1, j, p_SD, p_SD, p_NS + 1) - SegmentB(-i, -j - 1, p_SD, p_SD, p_NS + 1));
    case QSF_BORDER_N2D_ASYM:  // borders.h:156-166
      return borvec(2, -SegmentB(j + 1, -i - 1, p_SD, p_SD, p_NS + 1) + SegmentB(-j, i, p_SD, p_SD, p_NS + 1),
                    -SegmentB(i + 1, -j - 1, p_SD, p_SD, p_NS + 1) + SegmentB(-i, j, p_SD, p_SD, p_NS + 1));
    case QSF_BORDER_S2D:  // borders.h:168-183
      if (D == 2) return borvec(2, SegmentBi2j2(j, i, p_SD, p_NS, p_CAP), SegmentBi2j2(i, j, p_SD, p_NS, p_CAP));
      return borvec(3, P3(j, k, i, p_SD, p_NS, p_CAP, 0, p_DT) + P3(k, j, i, p_SD, p_NS, p_CAP, 0, p_DT),
                    P3(k, i, j, p_SD, p_NS, p_CAP, 0, p_DT) + P3(i, k, j, p_SD, p_NS, p_CAP, 0, p_DT),
                    P3(i, j, k, p_SD, p_NS, p_CAP, 0, p_DT) + P3(j, i, k, p_SD, p_NS, p_CAP, 0, p_DT));
    case QSF_BORDER_S2D_SYM:  // borders.h:184-195
      return borvec(2, SegmentB(j, i, p_SD, p_NS, p_CAP) - SegmentB(-j - 1, -i - 1, p_SD, p_NS, p_CAP),
                    SegmentB(i, j, p_SD, p_NS, p_CAP) - SegmentB(-i - 1, -j - 1, p_SD, p_NS, p_CAP));
    case QSF_BORDER_S2D_ASYM:  // borders.h:196-207
      return borvec(2, -SegmentB(j, -i - 1, p_SD, p_NS, p_CAP) + SegmentB(-j - 1, i, p_SD, p_NS, p_CAP),
                    -SegmentB(i, -j - 1, p_SD, p_NS, p_CAP) + SegmentB(-i - 1, j, p_SD, p_NS, p_CAP));
    case QSF_BORDER_N2T:  // borders.h:209-220 (3D only)
      return borvec(3, P3(j, k, i, p_DT, p_DT - 1, p_NS, p_DT - 1, p_SD) + P3(k, j, i, p_DT, p_DT - 1, p_NS, p_DT - 1, p_SD),
                    P3(k, i, j, p_DT, p_DT - 1, p_NS, p_DT - 1, p_SD) + P3(i, k, j, p_DT, p_DT - 1, p_NS, p_DT - 1, p_SD),
                    P3(i, j, k, p_DT, p_DT - 1, p_NS, p_DT - 1, p_SD) + P3(j, i, k, p_DT, p_DT - 1, p_NS, p_DT - 1, p_SD));
    case QSF_BORDER_S2T:  // borders.h:221-232
      return borvec(3, P3(j, k, i, p_DT, p_NS, p_CAP, p_DT, p_SD) + P3(k, j, i, p_DT, p_NS, p_CAP, p_DT, p_SD),
                    P3(k, i, j, p_DT, p_NS, p_CAP, p_DT, p_SD) + P3(i, k, j, p_DT, p_NS, p_CAP, p_DT, p_SD),
                    P3(i, j, k, p_DT, p_NS, p_CAP, p_DT, p_SD) + P3(j, i, k, p_DT, p_NS, p_CAP, p_DT, p_SD));
    case QSF_BORDER_D2T:  // borders.h:233-244
      return borvec(3, P3(j, k, i, p_DT, p_SD, p_CAP, p_SD, p_CAP), P3(k, i, j, p_DT, p_SD, p_CAP, p_SD, p_CAP),
                    P3(i, j, k, p_DT, p_SD, p_CAP, p_SD, p_CAP));
    case QSF_BORDER_T2CAP:  // borders.h:245-257
      return borvec(3, P3(j, k, i, p_CAP, p_DT, p_CAP, p_DT, p_CAP), P3(k, i, j, p_CAP, p_DT, p_CAP, p_DT, p_CAP),
                    P3(i, j, k, p_CAP, p_DT, p_CAP, p_DT, p_CAP));
    default: return {0, 0, 0};
  }
}

// ----------------------------------------------------------------------------------------
// Grid: CartesianGrid_ (coords.h) + LocalGrid (grid.h) + absorber + Spin0 (hamiltonian.h)
// + potential (potential.h), on one rank.
// ----------------------------------------------------------------------------------------
struct Grid {
  qsf_grid_desc d;
  int D;
  ind n[3], n2[3], m;
  double dx[3], inv_m, dV, dVm, L[3], xmin[3], inv_dx[3], inv_2dx[3], dp[3];
  double eta = 0;            // CAP η (absorbers.h:30,36)
  double NEcharge, EEcharge;  // potential.h:43,123-124
  int freeMask;
  bool bounded[3];
  bool isMain, many;
  int group;
  std::vector<cxd> psi;
  int rep = QSF_REP_X;
  std::vector<std::array<double, 3>> curr;

  explicit Grid(const qsf_grid_desc& desc) : d(desc) {
    D = d.dim;
    for (int i = 0; i < 3; ++i) { n[i] = i < D ? d.n[i] : 1; dx[i] = i < D ? d.dx[i] : 1.0; }
    // CartesianGrid_::init coords.h:55-93
    m = 1;
    for (int i = 0; i < D; ++i) { n2[i] = n[i] / 2; m *= n[i]; }
    for (int i = D; i < 3; ++i) n2[i] = 0;
    inv_m = 1.0 / m;
    dV = 1.0;
    for (int i = 0; i < D; ++i) {
      dV *= dx[i];
      L[i] = (dx[i] * (n[i] - 1));
      xmin[i] = -0.5 * L[i];
      inv_dx[i] = 1.0 / dx[i];
      inv_2dx[i] = inv_dx[i] / 2.0;
      dp[i] = 2.0 * pi / double(n[i]) / dx[i];
    }
    dVm = dV * inv_m;
    if (d.absorber == QSF_ABS_CAP) eta = Power(3.0 / double(d.nCAP), 4);  // absorbers.h:36
    NEcharge = d.Ncharge * d.Echarge;
    EEcharge = d.Echarge * d.Echarge;
    // Regions (mpi_logic.h:142-150)
    freeMask = d.region_free;
    int freeCount = 0;
    for (int i = 0; i < D; ++i) {
      bounded[i] = !(freeMask & (1 << i));
      freeCount += !bounded[i];
    }
    group = freeCount;
    isMain = (freeCount == 0);
    many = d.multigrid != 0;
    psi.assign(m, cxd(0, 0));
  }

  // Layout: X rep is row-major (n0,n1,n2); P rep of the main region of a D>1 grid is
  // (n1,n0,n2) (grid.h:238-248, canLeaveTransposed grid.h:118).
  bool transposedP() const { return D > 1 && isMain; }
  void shape(int R, ind* s) const {
    for (int i = 0; i < 3; ++i) s[i] = n[i];
    if (R == QSF_REP_P && transposedP()) std::swap(s[0], s[1]);
  }
  ind strideX(int dir) const {
    ind s = 1;
    for (int i = dir + 1; i < D; ++i) s *= n[i];
    return s;
  }

  // coords.h:132-138 — uses the storage axis' own n/dx/dp (the reference's P loop does too).
  double pos(int R, int dir, ind i) const {
    if (R == QSF_REP_X) return xmin[dir] + i * dx[dir];
    return dp[dir] * (i < n2[dir] ? i : i - n[dir]);
  }
  // grid.h:181-185
  ind neg_dist_from_edge(int dir, ind i) const { return i < n2[dir] ? -i : i - n[dir] + 1; }
  // grid.h:171-177 (X rep)
  ind abs_centered_index_X(int dir, ind i) const { return i - n2[dir]; }

  // ---- potential.h ----
  double Ne_i(double x) const { return NEcharge / std::sqrt(x * x + d.Nsoft); }  // :137
  double ee_i(double x, double y) const {                                        // :130-136
    if (d.potential == QSF_POT_ECKHARDT_SACHA) return EEcharge / std::sqrt((x - y) * (x - y) + x * y + d.Esoft);
    return EEcharge / std::sqrt((x - y) * (x - y) + d.Esoft);
  }
  double coulomb(int k, const double* c) const {  // potential.h:43 (right fold of squares)
    double s;
    if (k == 1) s = c[0] * c[0];
    else if (k == 2) s = (c[0] * c[0]) + (c[1] * c[1]);
    else s = (c[0] * c[0]) + ((c[1] * c[1]) + (c[2] * c[2]));
    return NEcharge / std::sqrt(s + d.Nsoft);
  }
  double full(const double* c) const {
    switch (d.potential) {
      case QSF_POT_COULOMB: return coulomb(D, c);
      case QSF_POT_EBERLY: case QSF_POT_ECKHARDT_SACHA:
        if (D == 1) return Ne_i(c[0]);                                               // :140
        if (D == 2) return ee_i(c[0], c[1]) + Ne_i(c[0]) + Ne_i(c[1]);               // :139
        return ee_i(c[0], c[1]) + ee_i(c[1], c[2]) + ee_i(c[2], c[0]) + Ne_i(c[0]) + Ne_i(c[1]) + Ne_i(c[2]);  // :138
      default: return 0.0;
    }
  }
  double partial(int fc, const double* c) const {  // potential.h:62-98 and :141-178
    if (d.potential == QSF_POT_NONE) return 0.0;
    if (fc == QSF_AXIS_NO) return full(c);
    if (d.potential == QSF_POT_COULOMB) {
      double a[3];
      if (D == 2) {
        if (fc == QSF_AXIS_Y) { a[0] = c[0]; return coulomb(1, a); }
        if (fc == QSF_AXIS_X) { a[0] = c[1]; return coulomb(1, a); }
        return 0.0;
      }
      if (D == 3) {
        if (fc == QSF_AXIS_Z) { a[0] = c[0]; a[1] = c[1]; return coulomb(2, a); }
        if (fc == QSF_AXIS_Y) { a[0] = c[0]; a[1] = c[2]; return coulomb(2, a); }
        if (fc == QSF_AXIS_X) { a[0] = c[1]; a[1] = c[2]; return coulomb(2, a); }
        // potential.h:72-77 as written: YZ and XY keep z, XZ keeps y.
        if (fc == QSF_AXIS_YZ) { a[0] = c[2]; return coulomb(1, a); }
        if (fc == QSF_AXIS_XZ) { a[0] = c[1]; return coulomb(1, a); }
        if (fc == QSF_AXIS_XY) { a[0] = c[2]; return coulomb(1, a); }
        return 0.0;
      }
      return 0.0;
    }
    if (D == 1) return 0.0;
    if (D == 2) {
      if (fc == QSF_AXIS_Y) return Ne_i(c[0]);
      if (fc == QSF_AXIS_X) return Ne_i(c[1]);
      return 0.0;
    }
    if (fc == QSF_AXIS_Z) return ee_i(c[0], c[1]) + Ne_i(c[0]) + Ne_i(c[1]);
    if (fc == QSF_AXIS_Y) return ee_i(c[0], c[2]) + Ne_i(c[0]) + Ne_i(c[2]);
    if (fc == QSF_AXIS_X) return ee_i(c[1], c[2]) + Ne_i(c[1]) + Ne_i(c[2]);
    if (fc == QSF_AXIS_YZ) return Ne_i(c[0]);
    if (fc == QSF_AXIS_XZ) return Ne_i(c[1]);
    if (fc == QSF_AXIS_XY) return Ne_i(c[2]);
    return 0.0;
  }
  // Spin0::potential (hamiltonian.h:60-71)
  double potentialX(const double* c) const { return many ? partial(freeMask, c) : full(c); }

  // ---- absorbers ----
  // cos^(1/8) edge profile over nCAP nodes (new): cos(π/2 · node/nCAP) written as the sine of
  // the distance from the edge, sin(π/2 · (nCAP − node)/nCAP), which is exactly 0 at the edge
  // node; the cosine form can round to a tiny negative value there and pow(·, 1/8) is then NaN.
  double cos18_node(ind node) const {
    if (node <= 0) return 1.0;
    return std::pow(std::sin(0.5 * pi * double(d.nCAP - node) / double(d.nCAP)), 0.125);
  }
  // CAP::mask (absorbers.h:48-54) / cos^(1/8) product over axes
  double mask(int k, const ind* nd) const {
    ind nodes[3];
    bool allneg = true;
    for (int i = 0; i < k; ++i) { nodes[i] = d.nCAP + nd[i]; allneg = allneg && nodes[i] < 0; }
    if (d.absorber == QSF_ABS_COS18) {
      double r = 1.0;
      for (int i = 0; i < k; ++i) r = r * cos18_node(nodes[i]);
      return r;
    }
    if (allneg) return 1.0;
    double S = 0.0;  // binary right fold p0 + (p1 + (p2 + 0)); terms are exact integers
    for (int i = k - 1; i >= 0; --i) S = (nodes[i] > 0 ? double(Power(nodes[i], 4)) : 0.0) + S;
    return std::exp(-double(S) * eta);
  }
  double inv_mask(int k, const ind* nd) const { return 1.0 - mask(k, nd); }  // absorbers.h:56-59
  // CAP::absorb (absorbers.h:40-46)
  double absorb(double delta, int k, const ind* nd) const {
    ind nodes[3];
    bool allneg = true;
    for (int i = 0; i < k; ++i) { nodes[i] = d.nCAP + nd[i]; allneg = allneg && nodes[i] < 0; }
    if (d.absorber == QSF_ABS_COS18) {
      // One full mask per absorber_tau of propagated time: mask^(δ/τ).
      double r = 1.0;
      for (int i = 0; i < k; ++i) r = r * cos18_node(nodes[i]);
      return std::pow(r, delta / d.absorber_tau);
    }
    if (allneg) return 1.0;
    double S = 0.0;
    for (int i = k - 1; i >= 0; --i) S = (nodes[i] > 0 ? double(Power(nodes[i], 4)) : 0.0) + S;
    return std::exp(-delta * double(S) * eta);
  }

  // ---- Spin0::operator() (hamiltonian.h:44-79) ----
  double H(int R, const double* c, const double* C, int couplesIn) const {
    double coup = 0.0;
    if (couplesIn == R) {  // ((coords * C) + ...) unary right fold
      if (D == 1) coup = c[0] * C[0];
      else if (D == 2) coup = (c[0] * C[0]) + (c[1] * C[1]);
      else coup = (c[0] * C[0]) + ((c[1] * C[1]) + (c[2] * C[2]));
    }
    if (R == QSF_REP_P) {
      double kin;
      if (D == 1) kin = 0.5 * c[0] * c[0];
      else if (D == 2) kin = (0.5 * c[0] * c[0]) + (0.5 * c[1] * c[1]);
      else kin = (0.5 * c[0] * c[0]) + ((0.5 * c[1] * c[1]) + (0.5 * c[2] * c[2]));
      return kin + coup;
    }
    return potentialX(c) + coup;
  }

  // grid.h:369-378 (main) and multigrid.h:660-680 (partial)
  void fourier(int to) {
    if (to == rep) return;
    ind sh[3] = {n[0], n[1], n[2]};
    if (isMain) {
      if (to == QSF_REP_P) {
        for (int a = 0; a < D; ++a) fft_axis(psi.data(), sh, a, -1);
        if (transposedP()) swap01(psi, n[0], n[1], n[2]);
      } else {
        if (transposedP()) swap01(psi, n[1], n[0], n[2]);
        for (int a = 0; a < D; ++a) fft_axis(psi.data(), sh, a, +1);
#pragma omp parallel for schedule(static)
        for (ind c = 0; c < m; ++c) psi[c] *= inv_m;  // normalizeAfterTwoFFT grid.h:363-367
      }
    } else {
      int boundedCount = 0;
      double partial_inv_m = 1.0;  // multigrid.h:148-150
      for (int a = 0; a < D; ++a)
        if (bounded[a]) { partial_inv_m *= 1.0 / n[a]; ++boundedCount; }
      for (int a = 0; a < D; ++a)
        if (bounded[a]) fft_axis(psi.data(), sh, a, to == QSF_REP_P ? -1 : +1);
      if (to == QSF_REP_X && boundedCount > 0)
        for (auto& v : psi) v *= partial_inv_m;
    }
    rep = to;
  }

  // grid.h:394-418
  void evolve(int M, int R, double delta, const double* C, int couplesIn) {
    ind s[3];
    shape(R, s);
    // row-major odometer of grid.h:414-417, one independent factor per point
#pragma omp parallel for schedule(static)
    for (ind c = 0; c < s[0] * s[1] * s[2]; ++c) {
      const ind i0 = c / (s[1] * s[2]), i1 = (c / s[2]) % s[1], i2 = c % s[2];
      double co[3] = {0, 0, 0};
      ind idx[3] = {i0, i1, i2};
      for (int a = 0; a < D; ++a) co[a] = pos(R, a, idx[a]);
      double v = delta * H(R, co, C, couplesIn);
      if (M == QSF_MODE_IM) psi[c] *= std::exp(-v);
      else psi[c] *= cxd{std::cos(-v), std::sin(-v)};
      if (d.absorber != QSF_ABS_NONE && R == QSF_REP_X && !many) {
        ind nd[3];
        for (int a = 0; a < D; ++a) nd[a] = neg_dist_from_edge(a, idx[a]);
        psi[c] *= absorb(delta, D, nd);
      }
    }
  }

  // grid.h:425-441 with Spin0::call (hamiltonian.h:81-90)
  double average(int R, int op) const {
    ind s[3];
    shape(R, s);
    double res = 0.0;
    ind c = 0;
    double co[3] = {0, 0, 0};
    for (ind i0 = 0; i0 < s[0]; ++i0)
      for (ind i1 = 0; i1 < s[1]; ++i1)
        for (ind i2 = 0; i2 < s[2]; ++i2, ++c) {
          ind idx[3] = {i0, i1, i2};
          double val = 1.0;
          if (op == 1) {  // PotentialEnergy -> potential<X>
            for (int a = 0; a < D; ++a) co[a] = pos(R, a, idx[a]);
            val = potentialX(co);
          } else if (op == 2) {  // KineticEnergy -> kinetic<P>
            for (int a = 0; a < D; ++a) co[a] = pos(R, a, idx[a]);
            if (D == 1) val = 0.5 * co[0] * co[0];
            else if (D == 2) val = (0.5 * co[0] * co[0]) + (0.5 * co[1] * co[1]);
            else val = (0.5 * co[0] * co[0]) + ((0.5 * co[1] * co[1]) + (0.5 * co[2] * co[2]));
          }
          res += std::norm(psi[c]) * val;
        }
    return res * (R == QSF_REP_X ? dV : dVm);  // vol<R> coords.h:48-53
  }

  void normalize() {  // grid.h:601-606
    double res = average(QSF_REP_X, 0);
    double f = std::sqrt(1.0 / res);
    for (auto& v : psi) v *= f;
  }

  // grid.h:467-528 on one rank, then velocityGaugeCorrection (propagator.h:249-257)
  void current_map(const double* vgC) {
    curr.assign(m, {0, 0, 0});
#pragma omp parallel for schedule(static)
    for (ind c = 0; c < m; ++c)
      for (int dir = 0; dir < D; ++dir) {
        ind s = strideX(dir);
        if ((c - s > 0) && (c + s < m)) {
          cxd tmp = psi[c + s] - psi[c - s];
          curr[c][dir] = inv_2dx[dir] * std::imag(std::conj(psi[c]) * tmp);
        } else curr[c][dir] = 0.0;
      }
    if (vgC)
      for (ind c = 0; c < m; ++c)
        for (int dir = 0; dir < D; ++dir) curr[c][dir] += vgC[dir] * std::norm(psi[c]);
  }
  // grid.h:559-573
  double flux(int kind, int size) const {
    double res = 0.0;
    ind c = 0;
    for (ind i0 = 0; i0 < n[0]; ++i0)
      for (ind i1 = 0; i1 < n[1]; ++i1)
        for (ind i2 = 0; i2 < n[2]; ++i2, ++c) {
          auto b = border(kind, size, D, abs_centered_index_X(0, i0), D > 1 ? abs_centered_index_X(1, i1) : 0,
                          D > 2 ? abs_centered_index_X(2, i2) : 0);
          double dot = 0.0;  // borVec::operator* borders_t.h:29-35
          for (int a = 0; a < D; ++a) dot += curr[c][a] * b[a];
          res += dot;
        }
    return res * dV / dx[0];
  }

  // grid.h:582-597
  void maskRegion() {
    ind c = 0;
    for (ind i0 = 0; i0 < n[0]; ++i0)
      for (ind i1 = 0; i1 < n[1]; ++i1)
        for (ind i2 = 0; i2 < n[2]; ++i2, ++c) {
          ind idx[3] = {i0, i1, i2}, nd[3];
          for (int a = 0; a < D; ++a) nd[a] = bounded[a] ? neg_dist_from_edge(a, idx[a]) : -n2[a];
          psi[c] *= mask(D, nd);
        }
  }
};

// ----------------------------------------------------------------------------------------
// Region sets and the multigrid transfer (wf/multigrid.h, basics/mpi_logic.h:103-263).
// ----------------------------------------------------------------------------------------
struct Source { int region; int fc; };

struct GridSet {
  std::vector<std::unique_ptr<Grid>> g;
  std::vector<std::vector<Source>> sources;  // per region (multigrid.h:128-146)
  int D = 1, groupCount = 1;
  bool multi = false;

  explicit GridSet(const qsf_grid_desc& desc) {
    D = desc.dim;
    multi = desc.multigrid != 0;
    if (!multi) {
      g.emplace_back(new Grid(desc));
      return;
    }
    // MultiRegionsReduced region order (mpi_logic.h:249-263)
    std::vector<int> freeAxes;
    if (D == 1) freeAxes = {QSF_AXIS_NO, QSF_AXIS_X};
    else if (D == 2) freeAxes = {QSF_AXIS_NO, QSF_AXIS_X, QSF_AXIS_Y, QSF_AXIS_XY};
    else freeAxes = {QSF_AXIS_NO, QSF_AXIS_X, QSF_AXIS_Y, QSF_AXIS_Z, QSF_AXIS_XY, QSF_AXIS_XZ, QSF_AXIS_YZ, QSF_AXIS_XYZ};
    auto cnt = [](int a) { return __builtin_popcount(a); };
    int maxc = 0;
    for (int fa : freeAxes) {
      qsf_grid_desc dd = desc;
      dd.region_free = fa;
      g.emplace_back(new Grid(dd));
      maxc = std::max(maxc, cnt(fa));
    }
    groupCount = maxc + 1;
    sources.resize(freeAxes.size());
    for (size_t r = 0; r < freeAxes.size(); ++r)
      for (size_t i = 0; i < freeAxes.size(); ++i)
        if (cnt(freeAxes[r]) - cnt(freeAxes[i]) == 1) {
          int x = freeAxes[r] ^ freeAxes[i];
          if (x == QSF_AXIS_X || x == QSF_AXIS_Y || x == QSF_AXIS_Z) sources[r].push_back({(int)i, x});
        }
  }

  // --- slice geometry, single rank (multigrid.h:214-349) ---
  struct SliceGeo {
    int dirFree;
    ind shape[3], stride[3];
    ind nCAP;
    int D;
  };
  bool slidesIn(int dirFree, int dir) const {  // multigrid.h:297-309
    if (D == 3) {
      if (dir == 1 && (dirFree == 0 || dirFree == 2)) return true;
      if (dir == 2 && dirFree == 1) return true;
    }
    return false;
  }
  SliceGeo geo(const Grid& rg, int dirFree) const {
    SliceGeo s;
    s.dirFree = dirFree;
    s.D = D;
    s.nCAP = rg.d.nCAP;
    for (int d = 0; d < 3; ++d) s.shape[d] = d < D ? (slidesIn(dirFree, d) ? s.nCAP : rg.n[d]) : 1;
    s.stride[D - 1] = 1;  // initSliceStrides multigrid.h:214-220
    for (int d = D - 2; d >= 0; --d) s.stride[d] = s.shape[d + 1] * s.stride[d + 1];
    for (int d = D; d < 3; ++d) s.stride[d] = 0;
    return s;
  }
  ind slider(const SliceGeo& s, int dir, ind box) const { return slidesIn(s.dirFree, dir) ? box * s.nCAP : 0; }
  ind sliceLimited(const SliceGeo& s, int dir) const {  // multigrid.h:332-338
    if (s.dirFree == dir && s.dirFree) return s.nCAP;
    return s.shape[dir];
  }
  ind top(const Grid& rg, const SliceGeo& s, int dir) const {  // multigrid.h:339-343
    return (s.dirFree == dir && s.dirFree) ? rg.n[dir] - s.nCAP : 0;
  }
  ind tslider(const Grid& rg, const SliceGeo& s, int dir, ind box) const { return slider(s, dir, box) + top(rg, s, dir); }
  ind slice_offset(const SliceGeo& s, const ind* c) const {
    ind o = 0;
    for (int d = 0; d < D; ++d) o += c[d] * s.stride[d];
    return o;
  }
  static ind data_offset(const Grid& g, const ind* c) {
    ind o = 0;
    for (int d = 0; d < g.D; ++d) o += c[d] * g.strideX(d);
    return o;
  }

  // getBox (multigrid.h:454-482): copy the source's edge rows into the slice.
  void getBox(const Grid& recv, const Grid& src, const SliceGeo& s, ind box, std::vector<cxd>& slice) const {
    const ind nCAP = s.nCAP;
    const bool isBottomX = 0 < nCAP;
    const ind sizeX = std::min<ind>(recv.n[0], nCAP);
    const bool isTopX = recv.n[0] - 1 > recv.n[0] - 1 - nCAP;
    const ind startX = std::max<ind>(0, recv.n[0] - nCAP);
    const int Dm1 = D - 1;
    ind cnt[3] = {0, 0, 0};
    ind lim[3] = {1, 1, 1};
    for (int d = 0; d < Dm1; ++d) lim[d] = sliceLimited(s, d);
    const ind count = sliceLimited(s, Dm1);
    for (;;) {
      if (s.dirFree != 0 || (isBottomX && cnt[0] < sizeX)) {
        ind sc[3] = {0, 0, 0}, wc[3] = {0, 0, 0};
        for (int d = 0; d < Dm1; ++d) { sc[d] = cnt[d]; wc[d] = cnt[d] + slider(s, d, box); }
        sc[Dm1] = 0;
        wc[Dm1] = slider(s, Dm1, box);
        ind so = slice_offset(s, sc), wo = data_offset(src, wc);
        for (ind q = 0; q < count; ++q) slice[so + q] = src.psi[wo + q];
      }
      if (s.dirFree != 0 || (isTopX && cnt[0] >= startX)) {
        ind sc[3] = {0, 0, 0}, wc[3] = {0, 0, 0};
        for (int d = 0; d < Dm1; ++d) { sc[d] = cnt[d] + top(recv, s, d); wc[d] = cnt[d] + tslider(recv, s, d, box); }
        sc[Dm1] = top(recv, s, Dm1);
        wc[Dm1] = tslider(recv, s, Dm1, box);
        ind so = slice_offset(s, sc), wo = data_offset(src, wc);
        for (ind q = 0; q < count; ++q) slice[so + q] = src.psi[wo + q];
      }
      // odometer over dirs 0..D-2, last fastest (rev<Is> = DIM-2-Is)
      int d = Dm1 - 1;
      for (; d >= 0; --d) {
        if (++cnt[d] < lim[d]) break;
        cnt[d] = 0;
      }
      if (d < 0) break;
    }
  }

  // mask_correction (multigrid.h:519-530)
  double mask_correction(const Grid& rg, int dirFree, const ind* nd) const {
    const double onesixth = 1.0 / 6.0, onehalf = 0.5;
    if (D == 3) {
      double sum = 0;  // binary right fold t0 + (t1 + (t2 + 0))
      for (int d = 2; d >= 0; --d) sum = (d == dirFree ? 0 : rg.mask(1, &nd[d])) + sum;
      double prod = 1.0;  // unary right fold u0 * (u1 * u2)
      for (int d = 2; d >= 0; --d) {
        double u = d == dirFree ? 1.0 : rg.mask(1, &nd[d]);
        prod = d == 2 ? u : u * prod;
      }
      return onehalf * sum + onesixth * (1 - prod);
    }
    if (D == 2) {
      double prod = 1.0;
      for (int d = 1; d >= 0; --d) {
        double u = d == dirFree ? 1.0 : rg.inv_mask(1, &nd[d]);
        prod = d == 1 ? u : u * prod;
      }
      return 1 - onehalf * prod;
    }
    return 1.0;
  }

  // maskSlice (multigrid.h:532-569)
  void maskSlice(const Grid& rg, const SliceGeo& s, ind box, std::vector<cxd>& slice) const {
    const ind nCAP = s.nCAP;
    const bool isBottomX = 0 < nCAP;
    const ind sizeX = std::min<ind>(rg.n[0], nCAP);
    const bool isTopX = rg.n[0] - 1 > rg.n[0] - 1 - nCAP;
    const ind startX = std::max<ind>(0, rg.n[0] - nCAP);
    const int f = s.dirFree;
    ind cnt[3] = {0, 0, 0}, lim[3] = {1, 1, 1};
    for (int d = 0; d < D; ++d) lim[d] = sliceLimited(s, d);
    double corr = 1.0;
    for (;;) {
      if (f != 0 || (isBottomX && cnt[0] < sizeX)) {
        ind ndf = rg.neg_dist_from_edge(f, cnt[f]);
        double mv = rg.inv_mask(1, &ndf);
        if (rg.group == 1) {
          ind nd[3];
          for (int d = 0; d < D; ++d) nd[d] = rg.neg_dist_from_edge(d, cnt[d] + slider(s, d, box));
          corr = mask_correction(rg, f, nd);
        }
        slice[slice_offset(s, cnt)] *= ((rg.group == 1) ? (mv * corr) : mv);
      }
      if (f != 0 || (isTopX && cnt[0] >= startX)) {
        ind ndf = rg.neg_dist_from_edge(f, cnt[f] + top(rg, s, f));
        double mv = rg.inv_mask(1, &ndf);
        if (rg.group == 1) {
          ind nd[3];
          for (int d = 0; d < D; ++d) nd[d] = rg.neg_dist_from_edge(d, cnt[d] + tslider(rg, s, d, box));
          corr = mask_correction(rg, f, nd);
        }
        ind tc[3];
        for (int d = 0; d < D; ++d) tc[d] = cnt[d] + top(rg, s, d);
        slice[slice_offset(s, tc)] *= ((rg.group == 1) ? (mv * corr) : mv);
      }
      int d = D - 1;
      for (; d >= 0; --d) {
        if (++cnt[d] < lim[d]) break;
        cnt[d] = 0;
      }
      if (d < 0) break;
    }
  }

  // add (multigrid.h:571-585)
  void add(Grid& rg, const SliceGeo& s, ind box, const std::vector<cxd>& slice) const {
    ind cnt[3] = {0, 0, 0};
    for (;;) {
      ind wc[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) wc[d] = cnt[d] + slider(s, d, box);
      rg.psi[data_offset(rg, wc)] += slice[slice_offset(s, cnt)];
      int d = D - 1;
      for (; d >= 0; --d) {
        if (++cnt[d] < s.shape[d]) break;
        cnt[d] = 0;
      }
      if (d < 0) break;
    }
  }

  // Slice of one (receiver, source, box): gathered, masked, not yet transformed.
  std::vector<cxd> build_slice(int r, int sI, ind box, SliceGeo& s) const {
    Grid& rg = *g[r];
    const Source& src = sources[r][sI];
    int dirFree = src.fc == QSF_AXIS_X ? 0 : src.fc == QSF_AXIS_Y ? 1 : 2;
    s = geo(rg, dirFree);
    ind m_slice = D == 3 ? rg.m / rg.n[2] * s.nCAP : rg.m;  // multigrid.h:277
    std::vector<cxd> slice(m_slice, cxd(0, 0));            // resetSlice :484-488
    getBox(rg, *g[src.region], s, box, slice);
    maskSlice(rg, s, box, slice);
    return slice;
  }

  // transfer (multigrid.h:627-657), one process per region emulated in group order.
  void transfer() {
    int boxesCount = D < 3 ? 1 : (int)(g[0]->n[0] / g[0]->d.nCAP);  // multigrid.h:188
    for (int gI = 1; gI < groupCount; ++gI)
      for (int bI = 0; bI < boxesCount; ++bI)
        for (int sI = 0; sI < gI; ++sI)
          for (size_t r = 0; r < g.size(); ++r) {
            if (g[r]->group != gI) continue;
            SliceGeo s;
            std::vector<cxd> slice = build_slice((int)r, sI, bI, s);
            fft_axis(slice.data(), s.shape, s.dirFree, -1);  // transf_slice: FFTW_FORWARD
            add(*g[r], s, bI, slice);
          }
  }
  void postCompute() {  // multigrid.h:617-625
    if (!multi) return;
    transfer();
    for (auto& gr : g)
      if (gr->group != groupCount - 1) gr->maskRegion();
  }
};

// ----------------------------------------------------------------------------------------
// Split tables (splitting.h) and the propagator (propagator.h).
// ----------------------------------------------------------------------------------------
struct Chain {
  std::vector<int> reps;
  std::vector<double> mults;
  double value;
};
// MultiProductSplit<Split3Base<X,P,X>, Order>::Chain<c> (splitting.h:18-27, 43-58, 63-87)
static Chain make_chain(int c, int order) {
  Chain ch;
  const double base[3] = {0.5, 1.0, 0.5};
  ch.reps.push_back(QSF_REP_X);  // canJoin: X, (P, X) repeated c times
  for (int k = 0; k < c; ++k) { ch.reps.push_back(QSF_REP_P); ch.reps.push_back(QSF_REP_X); }
  ch.mults.assign(ch.reps.size(), 0.0);
  int i = 0;
  for (int k = 0; k < c; ++k) {
    double v = 1.0 / c;  // SplitCoeff::value
    for (int j = 0; j < 3; ++j) ch.mults[i + j] += base[j] * v;
    i += 3 - 1;
  }
  double val = 1.0;  // unary left fold over Is = 1..Order
  bool first = true;
  for (int Is = 1; Is <= order; ++Is) {
    double t = (Is != c ? double(c * c) / (double(c * c) - double(Is * Is)) : 1.0);
    val = first ? t : val * t;
    first = false;
  }
  ch.value = val;
  return ch;
}

struct Prop {
  GridSet* set;
  qsf_prop_desc pd;
  Coupling coupling;
  std::vector<Chain> chains;
  int64_t max_steps = 0, step = 0;
  double dt, timer = 0, timer_copy = 0, state_accuracy = 0;
  double dif_energy = 1e300, last_change = 0;
  std::vector<std::vector<cxd>> psi_copy, psi_acc;
  std::vector<double> records;
  int width;

  Prop(GridSet* s, const qsf_prop_desc& p, const qsf_field_desc* f, int nf) : set(s), pd(p) {
    coupling.gauge = p.gauge;
    for (int i = 0; i < nf; ++i) coupling.fields.emplace_back(f[i]);
    dt = p.dt;
    for (int c = 1; c <= p.order; ++c) chains.push_back(make_chain(c, p.order));
    if (p.mode == QSF_MODE_RE) {
      max_steps = (int64_t)(coupling.maxPulseDuration() / dt + 1);  // propagator.h:153
      state_accuracy = 0;
    } else {
      max_steps = p.max_steps;
      state_accuracy = p.state_accuracy;
    }
    if (p.order > 1) {
      psi_copy.resize(s->g.size());
      psi_acc.resize(s->g.size());
    }
    width = 7 + p.nborders;
  }
  Grid& main() { return *set->g[0]; }
  int couplesIn() const { return coupling.couplesInRep(); }
  void C3(double* C) const {
    for (int d = 0; d < 3; ++d) C[d] = coupling[1 << d];
  }

  void fourier(int rep) { for (auto& gr : set->g) gr->fourier(rep); }
  void evolve(int rep, double delta) {
    double C[3];
    C3(C);
    for (auto& gr : set->g) gr->evolve(pd.mode, rep, delta, C, couplesIn());
  }
  // propagator.h:311-332
  void chainEvolve(const Chain& ch) {
    for (size_t SI = 0; SI < ch.reps.size(); ++SI) {
      int rep = ch.reps[SI];
      if (SI > 0) fourier(rep);
      if (couplesIn() == rep) coupling.precalc(timer);
      evolve(rep, dt * ch.mults[SI]);
      if (couplesIn() == rep) timer += dt * (ch.mults[SI] * 1.0);
    }
  }
  // propagator.h:334-346, 371-409
  void makeStep() {
    const size_t CC = chains.size();
    if (CC > 1) {
      for (size_t r = 0; r < set->g.size(); ++r) {
        psi_copy[r] = set->g[r]->psi;
        psi_acc[r].assign(set->g[r]->psi.size(), cxd(0, 0));
      }
      timer_copy = timer;
    }
    for (size_t c = 1; c <= CC; ++c) {
      if (c > 1) {
        timer = timer_copy;
        for (size_t r = 0; r < set->g.size(); ++r) set->g[r]->psi = psi_copy[r];
      }
      chainEvolve(chains[c - 1]);
      if (CC > 1) {
        double v = chains[c - 1].value;
        for (size_t r = 0; r < set->g.size(); ++r) {
          auto& w = set->g[r]->psi;
          if (c == CC - 1)
            for (size_t i = 0; i < w.size(); ++i) w[i] = v * w[i] + psi_acc[r][i];
          else
            for (size_t i = 0; i < w.size(); ++i) psi_acc[r][i] += v * w[i];
        }
      }
    }
    ++step;
  }

  // computeEach (propagator.h:275-300) with the canonical output order
  // OPERATION<Normalize>, AVG<Identity>, AVG<PotentialEnergy>, AVG<KineticEnergy>,
  // ENERGY_TOTAL, ENERGY_DIFFERENCE, FLUX<borders...>.
  void computeEach() {
    Grid& g = main();
    const double nan = std::nan("");
    std::vector<double> row(width, nan);
    row[0] = (double)step;
    row[1] = timer;
    int o = pd.outputs;
    double epot = nan, ekin = nan;
    if (o & QSF_OUT_NORMALIZE) g.normalize();
    if (o & QSF_OUT_NORM) row[2] = g.average(QSF_REP_X, 0);
    if (o & QSF_OUT_EPOT) row[3] = epot = g.average(QSF_REP_X, 1);
    if (o & QSF_OUT_FLUX) {
      double C[3];
      C3(C);
      g.current_map(couplesIn() == QSF_REP_P ? C : nullptr);
      for (int b = 0; b < pd.nborders; ++b) row[7 + b] = g.flux(pd.borders[b].kind, pd.borders[b].size);
    }
    if (o & QSF_OUT_EKIN) {
      g.fourier(QSF_REP_P);
      row[4] = ekin = g.average(QSF_REP_P, 2);
      g.fourier(QSF_REP_X);
    }
    if (o & QSF_OUT_ETOT) row[5] = epot + ekin;  // SUM: getLastValue of both
    if (o & QSF_OUT_EDIFF) {                      // CHANGE: current minus previous call
      double curr = epot + ekin;
      row[6] = curr - last_change;
      dif_energy = curr - last_change;
      last_change = curr;
    }
    records.insert(records.end(), row.begin(), row.end());
  }
  bool shouldContinue() const {  // propagator.h:304-309
    if (pd.mode == QSF_MODE_RE) return step <= max_steps;
    return std::fabs(dif_energy) > state_accuracy && step != max_steps;
  }
  void start() { computeEach(); }
  void loopBody() { makeStep(); computeEach(); set->postCompute(); }
};

}  // namespace orc

using namespace orc;

struct orc_grid { Grid g; };
struct orc_set { GridSet s; explicit orc_set(const qsf_grid_desc& d) : s(d) {} };
struct orc_prop { Prop p; };

static Prop& P(orc_prop* p) { return *reinterpret_cast<Prop*>(p); }
static const Prop& P(const orc_prop* p) { return *reinterpret_cast<const Prop*>(p); }
static Grid& G(orc_grid* g) { return *reinterpret_cast<Grid*>(g); }
static const Grid& G(const orc_grid* g) { return *reinterpret_cast<const Grid*>(g); }

extern "C" {

void orc_fft1d(int64_t n, int sign, const double* in, double* out) {
  std::vector<cxd> a(n);
  for (int64_t i = 0; i < n; ++i) a[i] = cxd(in[2 * i], in[2 * i + 1]);
  FFT1D f(n);
  f.exec(a.data(), sign);
  for (int64_t i = 0; i < n; ++i) { out[2 * i] = a[i].real(); out[2 * i + 1] = a[i].imag(); }
}

// OpenMP threads of the parallel loops (<= 0: the runtime's default); returns the count in use.
int orc_set_threads(int k) {
  if (k > 0) omp_set_num_threads(k);
  return omp_get_max_threads();
}

void orc_dft_bruteforce(int64_t n, int sign, const double* in, double* out) {
  const long double tp = 2.0L * 3.14159265358979323846264338327950288L;
  for (int64_t k = 0; k < n; ++k) {
    long double re = 0, im = 0;
    for (int64_t j = 0; j < n; ++j) {
      long double a = sign * tp * (long double)((j * k) % n) / (long double)n;
      long double c = cosl(a), s = sinl(a);
      re += in[2 * j] * c - in[2 * j + 1] * s;
      im += in[2 * j] * s + in[2 * j + 1] * c;
    }
    out[2 * k] = (double)re;
    out[2 * k + 1] = (double)im;
  }
}

orc_set* orc_set_create(const qsf_grid_desc* d) { return new orc_set(*d); }
void orc_set_destroy(orc_set* s) { delete s; }
int orc_set_count(const orc_set* s) { return (int)s->s.g.size(); }
orc_grid* orc_set_grid(orc_set* s, int r) { return reinterpret_cast<orc_grid*>(s->s.g[r].get()); }

int64_t orc_grid_size(const orc_grid* g) { return G(g).m; }
int orc_grid_rep(const orc_grid* g) { return G(g).rep; }
void orc_grid_set_psi(orc_grid* g, const double* a) {
  auto& p = G(g).psi;
  for (size_t i = 0; i < p.size(); ++i) p[i] = cxd(a[2 * i], a[2 * i + 1]);
}
void orc_grid_get_psi(const orc_grid* g, double* a) {
  auto& p = G(g).psi;
  for (size_t i = 0; i < p.size(); ++i) { a[2 * i] = p[i].real(); a[2 * i + 1] = p[i].imag(); }
}
void orc_fourier(orc_grid* g, int to) { G(g).fourier(to); }
void orc_evolve(orc_grid* g, int mode, int rep, double delta, const double* C) {
  // Single-grid evolve with an explicit coupling; couplesIn derived from the nonzero entries
  // is ambiguous, so the caller passes C only for the representation it couples in.
  double z[3] = {0, 0, 0};
  G(g).evolve(mode, rep, delta, C ? C : z, C ? rep : QSF_REP_NONE);
}
double orc_average(orc_grid* g, int rep, int op) { return G(g).average(rep, op); }
void orc_normalize(orc_grid* g) { G(g).normalize(); }
void orc_current_map(orc_grid* g, const double* vgC, double* out) {
  G(g).current_map(vgC);
  auto& c = G(g).curr;
  int D = G(g).D;
  for (size_t i = 0; i < c.size(); ++i)
    for (int d = 0; d < D; ++d) out[i * D + d] = c[i][d];
}
void orc_flux(orc_grid* g, const double* vgC, const qsf_border_desc* b, int nb, double* out) {
  G(g).current_map(vgC);
  for (int i = 0; i < nb; ++i) out[i] = G(g).flux(b[i].kind, b[i].size);
}
void orc_mask_region(orc_grid* g) { G(g).maskRegion(); }
void orc_set_transfer(orc_set* s) { s->s.transfer(); }
void orc_set_post_compute(orc_set* s) { s->s.postCompute(); }
int orc_set_transfer_slice(orc_set* s, int region, int sI, double* out) {
  if (!s->s.multi || region >= (int)s->s.g.size() || sI >= (int)s->s.sources[region].size()) return -1;
  GridSet::SliceGeo geo;
  auto slice = s->s.build_slice(region, sI, 0, geo);
  for (size_t i = 0; i < slice.size(); ++i) { out[2 * i] = slice[i].real(); out[2 * i + 1] = slice[i].imag(); }
  return (int)slice.size();
}

double orc_absorber_mask(const orc_grid* g, int nd, const int64_t* nodes) {
  ind a[3] = {0, 0, 0};
  for (int i = 0; i < nd; ++i) a[i] = nodes[i];
  return G(g).mask(nd, a);
}
double orc_absorber_absorb(const orc_grid* g, double delta, int nd, const int64_t* nodes) {
  ind a[3] = {0, 0, 0};
  for (int i = 0; i < nd; ++i) a[i] = nodes[i];
  return G(g).absorb(delta, nd, a);
}
void orc_border_vec(int kind, int size, int dim, const int64_t* idx, double* out3) {
  auto v = border(kind, size, dim, idx[0], dim > 1 ? idx[1] : 0, dim > 2 ? idx[2] : 0);
  for (int i = 0; i < 3; ++i) out3[i] = v[i];
}
double orc_potential(const orc_grid* g, const double* c) { return G(g).full(c); }
double orc_potential_partial(const orc_grid* g, int fc, const double* c) { return G(g).partial(fc, c); }

void orc_coupling_sequence(const qsf_field_desc* f, int nf, int gauge, const double* t, int nt, double* out) {
  Coupling c;
  c.gauge = gauge;
  for (int i = 0; i < nf; ++i) c.fields.emplace_back(f[i]);
  for (int k = 0; k < nt; ++k) {
    c.precalc(t[k]);
    for (int d = 0; d < 3; ++d) out[3 * k + d] = c[1 << d];
  }
}
double orc_max_pulse_duration(const qsf_field_desc* f, int nf) {
  Coupling c;
  for (int i = 0; i < nf; ++i) c.fields.emplace_back(f[i]);
  return c.maxPulseDuration();
}

orc_prop* orc_prop_create(orc_set* s, const qsf_prop_desc* p, const qsf_field_desc* f, int nf) {
  return reinterpret_cast<orc_prop*>(new Prop(&s->s, *p, f, nf));
}
void orc_prop_destroy(orc_prop* p) { delete reinterpret_cast<Prop*>(p); }
int64_t orc_prop_max_steps(const orc_prop* p) { return P(p).max_steps; }
void orc_prop_start(orc_prop* p) { P(p).start(); }
int orc_prop_should_continue(orc_prop* p) { return P(p).shouldContinue(); }
void orc_prop_step(orc_prop* p) { P(p).loopBody(); }
void orc_prop_finish(orc_prop* p) { P(p).loopBody(); }
void orc_prop_make_step(orc_prop* p) { P(p).makeStep(); }
int64_t orc_prop_step_count(const orc_prop* p) { return P(p).step; }
double orc_prop_time(const orc_prop* p) { return P(p).timer; }
int orc_prop_nrecords(const orc_prop* p) { return (int)(P(p).records.size() / P(p).width); }
int orc_prop_record_width(const orc_prop* p) { return P(p).width; }
void orc_prop_records(const orc_prop* p, double* out) {
  std::memcpy(out, P(p).records.data(), P(p).records.size() * sizeof(double));
}
void orc_prop_coupling(const orc_prop* p, double* out3) { P(p).C3(out3); }

}  // extern "C"
