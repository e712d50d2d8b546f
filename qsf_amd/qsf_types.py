"""ctypes mirror of include/qsf_types.h (descriptors of the C ABI).

The enumerations follow the reference's compile-time vocabulary: REP/MODE/AXIS
(include/basics/enums.h:15-62), potentials (include/potential.h:21-179), gauges
(include/coupling.h:6-17), absorber (include/wf/absorbers.h:16-60), pulses
(include/field_p.h:47-194), flux borders (include/fluxes/borders.h:38-257).
"""
import ctypes as C

REP_NONE, REP_X, REP_P, REP_BOTH = 0, 1, 2, 3
MODE_IM, MODE_RE = 1, 2
AXIS_NO, AXIS_X, AXIS_Y, AXIS_Z, AXIS_XY, AXIS_XZ, AXIS_YZ, AXIS_XYZ = 0, 1, 2, 4, 3, 5, 6, 7
POT_NONE, POT_COULOMB, POT_EBERLY, POT_ECKHARDT_SACHA = 0, 1, 2, 3
GAUGE_NONE, GAUGE_VELOCITY, GAUGE_LENGTH = 0, 1, 2
ABS_NONE, ABS_CAP, ABS_COS18 = 0, 1, 2
FP64, FP32_MIXED, FP32_STATE = 0, 1, 2
PULSE_EMPTY, PULSE_SIN, PULSE_CONST, PULSE_CHEMPHYS = 0, 1, 2, 3
ENV_NONE, ENV_SIN2, ENV_CHEMPHYS, ENV_RAMP, ENV_SIN, ENV_GAUSS = 0, 1, 2, 3, 4, 5
(BORDER_BOX, BORDER_S2CAP, BORDER_D2CAP, BORDER_N2S, BORDER_N2D, BORDER_N2D_SYM,
 BORDER_N2D_ASYM, BORDER_S2D, BORDER_S2D_SYM, BORDER_S2D_ASYM, BORDER_N2T, BORDER_S2T,
 BORDER_D2T, BORDER_T2CAP) = range(14)
OUT_NORM, OUT_EPOT, OUT_EKIN, OUT_ETOT, OUT_EDIFF, OUT_FLUX, OUT_NORMALIZE = (1 << i for i in range(7))

# ZOA_FLUX_2D = FLUX<N2S, N2D_SYM, N2D_ASYM, S2D_SYM, S2D_ASYM, S2CAP, D2CAP> (computations.h:70)
ZOA_FLUX_2D = [BORDER_N2S, BORDER_N2D_SYM, BORDER_N2D_ASYM, BORDER_S2D_SYM, BORDER_S2D_ASYM,
               BORDER_S2CAP, BORDER_D2CAP]
# ZOA_FLUX_3D = FLUX<N2S, N2D, N2T, S2D, S2T, S2CAP, D2CAP, T2CAP> (computations.h:68)
ZOA_FLUX_3D = [BORDER_N2S, BORDER_N2D, BORDER_N2T, BORDER_S2D, BORDER_S2T, BORDER_S2CAP,
               BORDER_D2CAP, BORDER_T2CAP]

# Record columns of a propagation (one row per computeEach): step, time, then observables.
REC_STEP, REC_TIME, REC_NORM, REC_EPOT, REC_EKIN, REC_ETOT, REC_EDIFF, REC_FLUX0 = range(8)


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("precision", C.c_int32),
                ("n", C.c_int64 * 3), ("dx", C.c_double * 3),
                ("potential", C.c_int32), ("absorber", C.c_int32),
                ("Ncharge", C.c_double), ("Echarge", C.c_double),
                ("Nsoft", C.c_double), ("Esoft", C.c_double),
                ("nCAP", C.c_int64), ("absorber_tau", C.c_double),
                ("region_free", C.c_int32), ("multigrid", C.c_int32)]


class FieldDesc(C.Structure):
    _fields_ = [("axis", C.c_int32), ("pulse", C.c_int32), ("envelope", C.c_int32),
                ("pad_", C.c_int32),
                ("field", C.c_double), ("omega", C.c_double), ("ncycles", C.c_double),
                ("fwhm_percent", C.c_double), ("phase_in_pi", C.c_double),
                ("delay_in_cycles", C.c_double)]


class BorderDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("size", C.c_int32)]


class PropDesc(C.Structure):
    _fields_ = [("mode", C.c_int32), ("order", C.c_int32), ("dt", C.c_double),
                ("max_steps", C.c_int64), ("state_accuracy", C.c_double),
                ("gauge", C.c_int32), ("outputs", C.c_int32),
                ("comp_interval", C.c_int32), ("nborders", C.c_int32),
                ("borders", BorderDesc * 16)]


def grid_desc(dim, n, dx, *, precision=FP64, potential=POT_NONE, absorber=ABS_NONE,
              Ncharge=1.0, Echarge=-1.0, Nsoft=1.0, Esoft=1.0, nCAP=0, absorber_tau=1.0,
              region_free=0, multigrid=0):
    d = GridDesc()
    d.dim = dim
    d.precision = precision
    ns = list(n) if hasattr(n, "__len__") else [n] * dim
    dxs = list(dx) if hasattr(dx, "__len__") else [dx] * dim
    for i in range(3):
        d.n[i] = ns[i] if i < dim else 1
        d.dx[i] = dxs[i] if i < dim else 1.0
    d.potential, d.absorber = potential, absorber
    d.Ncharge, d.Echarge, d.Nsoft, d.Esoft = Ncharge, Echarge, Nsoft, Esoft
    d.nCAP, d.absorber_tau = nCAP, absorber_tau
    d.region_free, d.multigrid = region_free, multigrid
    return d


def field_desc(axis, field, omega, ncycles, *, pulse=PULSE_SIN, envelope=ENV_SIN2,
               fwhm_percent=0.0, phase_in_pi=0.0, delay_in_cycles=0.0):
    f = FieldDesc()
    f.axis, f.pulse, f.envelope = axis, pulse, envelope
    f.field, f.omega, f.ncycles = field, omega, ncycles
    f.fwhm_percent, f.phase_in_pi, f.delay_in_cycles = fwhm_percent, phase_in_pi, delay_in_cycles
    return f


def prop_desc(mode, order, dt, *, max_steps=0, state_accuracy=0.0, gauge=GAUGE_VELOCITY,
              outputs=0, comp_interval=1, borders=()):
    p = PropDesc()
    p.mode, p.order, p.dt = mode, order, dt
    p.max_steps, p.state_accuracy = max_steps, state_accuracy
    p.gauge, p.outputs, p.comp_interval = gauge, outputs, comp_interval
    p.nborders = len(borders)
    for i, b in enumerate(borders):
        kind, size = (b if isinstance(b, tuple) else (b, 0))
        p.borders[i].kind, p.borders[i].size = kind, size
    return p


def field_array(fields):
    arr = (FieldDesc * max(1, len(fields)))()
    for i, f in enumerate(fields):
        arr[i] = f
    return arr
