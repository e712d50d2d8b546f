/* qsf_types.h — plain-C descriptors shared by the HIP library's C ABI (qsf_hip.h)
 * and by anything that drives it (the C++ host mirror, ctypes bindings, the CPU oracle).
 *
 * The enumerations mirror the reference's compile-time vocabulary so a maintainer can map
 * one template instantiation onto one descriptor:
 *   REP   include/basics/enums.h:36-42     MODE  include/basics/enums.h:15-21
 *   AXIS  include/basics/enums.h:46-62     potentials include/potential.h:21-179
 *   gauges include/coupling.h:6-17, 101-118   absorber include/wf/absorbers.h:16-60
 *   pulses/envelopes include/field_p.h:47-194, include/field.h:9-53
 * No torch or HIP types appear here: every field is a fixed-width integer or a double.
 */
#ifndef QSF_TYPES_H
#define QSF_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* REP (enums.h:36-42) */
enum { QSF_REP_NONE = 0, QSF_REP_X = 1, QSF_REP_P = 2, QSF_REP_BOTH = 3 };
/* MODE (enums.h:15-21) */
enum { QSF_MODE_IM = 1, QSF_MODE_RE = 2 };
/* AXIS bit masks (enums.h:46-62) */
enum {
  QSF_AXIS_NO = 0, QSF_AXIS_X = 1, QSF_AXIS_Y = 2, QSF_AXIS_Z = 4,
  QSF_AXIS_XY = 3, QSF_AXIS_XZ = 5, QSF_AXIS_YZ = 6, QSF_AXIS_XYZ = 7
};

/* Interaction potentials (potential.h). */
enum {
  QSF_POT_NONE = 0,
  QSF_POT_COULOMB = 1,          /* CoulombInteraction            potential.h:21-82   */
  QSF_POT_EBERLY = 2,           /* ReducedDimInteraction<Eberly> potential.h:97-179  */
  QSF_POT_ECKHARDT_SACHA = 3    /* ReducedDimInteraction<EckhardSacha>               */
};

/* Gauge of the dipole coupling (coupling.h). LENGTH has no reference specialisation
 * (coupling.h:9-11 is a tag only); it couples x·F(t) in the X representation. */
enum { QSF_GAUGE_NONE = 0, QSF_GAUGE_VELOCITY = 1, QSF_GAUGE_LENGTH = 2 };

/* Absorbers. CAP follows absorbers.h:16-60; COS18 is the cos^(1/8) edge mask, new here. */
enum { QSF_ABS_NONE = 0, QSF_ABS_CAP = 1, QSF_ABS_COS18 = 2 };

/* Arithmetic of the device step. FP32_MIXED: complex FP32 state and FP32 butterflies with
 * compensated twiddles and constants, FP64 phase tables and reductions. FP32_STATE: the same
 * FP32 state in HBM (half the bytes of FP64), every transform and phase product in FP64, one
 * rounding per element at each store; single-region grids and batches (multigrid sets refuse it). */
enum { QSF_FP64 = 0, QSF_FP32_MIXED = 1, QSF_FP32_STATE = 2 };

/* Pulse shapes (field_p.h:47-84) and envelopes (field_p.h:88-194). */
enum {
  QSF_PULSE_EMPTY = 0, QSF_PULSE_SIN = 1, QSF_PULSE_CONST = 2, QSF_PULSE_CHEMPHYS = 3
};
enum {
  QSF_ENV_NONE = 0, QSF_ENV_SIN2 = 1, QSF_ENV_CHEMPHYS = 2, QSF_ENV_RAMP = 3,
  QSF_ENV_SIN = 4, QSF_ENV_GAUSS = 5
};

/* Flux border families (fluxes/borders.h:38-257). BOX takes a size in nodes. */
enum {
  QSF_BORDER_BOX = 0,
  QSF_BORDER_S2CAP = 1, QSF_BORDER_D2CAP = 2, QSF_BORDER_N2S = 3, QSF_BORDER_N2D = 4,
  QSF_BORDER_N2D_SYM = 5, QSF_BORDER_N2D_ASYM = 6, QSF_BORDER_S2D = 7,
  QSF_BORDER_S2D_SYM = 8, QSF_BORDER_S2D_ASYM = 9, QSF_BORDER_N2T = 10,
  QSF_BORDER_S2T = 11, QSF_BORDER_D2T = 12, QSF_BORDER_T2CAP = 13,
  QSF_BORDER_COUNT = 14
};

/* Observables of one propagation (wf/computations.h, propagator.h:26-27). They are
 * evaluated as if listed to BufferedOutputs in this order: OPERATION<Normalize>,
 * AVG<Identity>, AVG<PotentialEnergy>, AVG<KineticEnergy>, ENERGY_TOTAL, ENERGY_DIFFERENCE,
 * FLUX<borders...> (computeEach propagator.h:275-300 evaluates in list order). */
enum {
  QSF_OUT_NORM = 1 << 0,        /* AVG<Identity>                      */
  QSF_OUT_EPOT = 1 << 1,        /* AVG<PotentialEnergy>               */
  QSF_OUT_EKIN = 1 << 2,        /* AVG<KineticEnergy> (P rep)         */
  QSF_OUT_ETOT = 1 << 3,        /* ENERGY_TOTAL  = SUM<EPOT, EKIN>    */
  QSF_OUT_EDIFF = 1 << 4,       /* ENERGY_DIFFERENCE = CHANGE<ETOT>   */
  QSF_OUT_FLUX = 1 << 5,        /* FLUX<borders...>                   */
  QSF_OUT_NORMALIZE = 1 << 6    /* OPERATION<Normalize>               */
};

/* Reference flux-border node constants (fluxes/borders.h:30-36, tuned for dx=0.5). */
#define QSF_P_NS 25
#define QSF_P_SD 14
#define QSF_P_DT 10
#define QSF_P_CAP 40

/* One grid (CartesianGrid_ coords.h:14-139 + LocalGrid grid.h:63-123 + absorber + potential). */
typedef struct qsf_grid_desc {
  int32_t dim;            /* 1..3                                                   */
  int32_t precision;      /* QSF_FP64 | QSF_FP32_MIXED | QSF_FP32_STATE             */
  int64_t n[3];           /* nodes per axis (unused axes: 1)                        */
  double dx[3];           /* spacing per axis                                       */
  int32_t potential;      /* QSF_POT_*                                              */
  int32_t absorber;       /* QSF_ABS_*                                              */
  double Ncharge, Echarge, Nsoft, Esoft;   /* InteractionBase potential.h:7-12      */
  int64_t nCAP;           /* absorber width in nodes                                */
  double absorber_tau;    /* COS18 only: time over which one full mask is applied   */
  int32_t region_free;    /* AXIS mask of free (momentum-resident) axes; 0 = main   */
  int32_t multigrid;      /* 1 when the grid is one region of a multi-region set    */
} qsf_grid_desc;

/* One field acting on an axis mask (Field<AXIS, Envelope<Pulse>> field.h:9-53). */
typedef struct qsf_field_desc {
  int32_t axis;           /* AXIS mask                                              */
  int32_t pulse;          /* QSF_PULSE_*                                            */
  int32_t envelope;       /* QSF_ENV_*                                              */
  int32_t pad_;
  double field, omega, ncycles, fwhm_percent, phase_in_pi, delay_in_cycles;
} qsf_field_desc;

/* One flux border request (FLUX<BOX<size>> or a region border). */
typedef struct qsf_border_desc {
  int32_t kind;           /* QSF_BORDER_*                                           */
  int32_t size;           /* BOX size in nodes (ignored otherwise)                  */
} qsf_border_desc;

/* Propagation settings (SplitPropagator propagator.h:51-161). */
typedef struct qsf_prop_desc {
  int32_t mode;           /* QSF_MODE_IM | QSF_MODE_RE                              */
  int32_t order;          /* MultiProductSplit<Split3Base<X,P,X>, order>            */
  double dt;
  int64_t max_steps;      /* IM only; RE derives it from the fields (propagator.h:153) */
  double state_accuracy;  /* IM stop criterion                                      */
  int32_t gauge;          /* QSF_GAUGE_*                                            */
  int32_t outputs;        /* QSF_OUT_* mask                                         */
  int32_t comp_interval;  /* observables are recorded every comp_interval steps     */
  int32_t nborders;       /* number of flux borders in borders[]                    */
  qsf_border_desc borders[16];
} qsf_prop_desc;

#ifdef __cplusplus
}
#endif
#endif
