// axis_pass_inst.hip — explicit instantiations of the fused axis pass for one precision, one
// mode and one group of axis lengths. The Makefile compiles this file once per
// (QSF_PREC: 0 = double, 1 = float, 2 = double arithmetic over a float state; QSF_ROWMODE: 1 =
// row, 0 = column; QSF_GROUP: 0..11), so the
// kernels of the long transforms build in parallel translation units.
#include "axis_pass_impl.hpp"

#if !defined(QSF_PREC) || !defined(QSF_ROWMODE) || !defined(QSF_GROUP)
#error "QSF_PREC, QSF_ROWMODE and QSF_GROUP must be defined"
#endif

#if QSF_PREC == 1
using QsfReal = float;
#else
using QsfReal = double;
#endif

#if QSF_GROUP == 0
#define QSF_GROUP_SIZES(X) X(8) X(16) X(32) X(64) X(128) X(12) X(24) X(48) X(96)
#elif QSF_GROUP == 1
#define QSF_GROUP_SIZES(X) X(192) X(384)
#elif QSF_GROUP == 2
#define QSF_GROUP_SIZES(X) X(3072)
#elif QSF_GROUP == 3
#define QSF_GROUP_SIZES(X) X(256)
#elif QSF_GROUP == 4
#define QSF_GROUP_SIZES(X) X(512)
#elif QSF_GROUP == 5
#define QSF_GROUP_SIZES(X) X(768)
#elif QSF_GROUP == 6
#define QSF_GROUP_SIZES(X) X(1024)
#elif QSF_GROUP == 7
#define QSF_GROUP_SIZES(X) X(1536)
#elif QSF_GROUP == 8
#define QSF_GROUP_SIZES(X) X(2048)
#elif QSF_GROUP == 9
#define QSF_GROUP_SIZES(X) X(4096)
#elif QSF_GROUP == 10
#define QSF_GROUP_SIZES(X) X(10) X(20) X(40) X(80) X(160) X(320)
#elif QSF_GROUP == 11
#define QSF_GROUP_SIZES(X) X(640) X(1280) X(2560)
#else
#error "QSF_GROUP out of range"
#endif

namespace qsf {
#if QSF_PREC == 2
// the tile counts are the FP64 kernels' (instantiated with QSF_PREC = 0)
#define QSF_INST(NN) \
  template int launch_size<double, (QSF_ROWMODE != 0), NN, true>(int, const PassParams&, hipStream_t);
#else
#define QSF_INST(NN)                                                                            \
  template int launch_size<QsfReal, (QSF_ROWMODE != 0), NN>(int, const PassParams&, hipStream_t); \
  template int64_t tiles_size<QsfReal, (QSF_ROWMODE != 0), NN>(int, const PassParams&);
#endif
QSF_GROUP_SIZES(QSF_INST)
#undef QSF_INST
}  // namespace qsf
