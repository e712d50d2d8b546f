"""Every compiled kernel variant against the FP64 oracle.

The autotuner picks, per (precision, pass class, axis length, grid size), one of the six
compiled variants of an axis pass (two Stockham factorisations x three tile widths), and the
timed workloads run whichever variant won. Each variant is forced here through qsf_tune_set on
the key its pass uses, and then

* the forward transform of a random state (fourier<P>, wf/grid.h:369-378), and
* a fused propagation of a few steps (inverse transform, X halves, observation with a flux
  border gather, forward transform, P phase: propagator.h:311-346, wf/grid.h:394-423)

are compared with the oracle on the same input. Rectangular 2D grids put the tuned length on
the row pass (16 x N) or the column pass (N x 64); 3D grids put it on the row pass with the
patched (i0, i1) line order (32 x 64 x N), the first column pass (N x 32 x 64) or the middle
pass (32 x N x 64); 2D multigrid grids (N x N, regions NO/X/Y/XY) add the region masks and
multigrid edge writes of the row pass and the fused transfers.

Bars: FP64 L2(psi) relative error <= 1e-10; FP32-mixed <= 2e-5 for one transform and <= 1e-4
after the propagation (single-precision state, FP64 tables); records to 1e-10 (FP64) / 1e-4.
FP32 state with FP64 arithmetic (the FP64 kernels with FP32 loads and stores; single-region
grids): 3e-7 for one transform, 2e-6 after the propagation, records to 1e-5.
"""
import numpy as np
import pytest

from oracle import oracle as O
from qsf_amd import qsf_types as T

pytestmark = pytest.mark.gpu

TUNED = [256, 512, 768, 1024, 1536, 2048, 4096]
VARIANTS = range(6)
FP64_TOL = 1e-10
FP32_FFT_TOL = 2e-5
FP32_PROP_TOL = 1e-4
S32_FFT_TOL = 3e-7
S32_PROP_TOL = 2e-6
STEPS = 3


def fft_tol(prec):
    return {T.FP64: FP64_TOL, T.FP32_MIXED: FP32_FFT_TOL, T.FP32_STATE: S32_FFT_TOL}[prec]


def prop_tol(prec):
    return {T.FP64: FP64_TOL, T.FP32_MIXED: FP32_PROP_TOL, T.FP32_STATE: S32_PROP_TOL}[prec]


def rec_tol(prec):
    return {T.FP64: 1e-9, T.FP32_MIXED: 1e-4, T.FP32_STATE: 1e-5}[prec]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    from qsf_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def state(m, seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=m) + 1j * rng.normal(size=m)) * 0.01


_ORACLE = {}


def oracle_fft(shape, dx, seed):
    key = ("fft", shape, seed)
    if key not in _ORACLE:
        d = T.grid_desc(len(shape), list(shape), dx)
        g = O.GridSet(d)
        g[0].set_psi(state(int(np.prod(shape)), seed))
        g[0].fourier(T.REP_P)
        _ORACLE[key] = g[0].psi()
    return _ORACLE[key]


def prop_setup(shape, dx, multigrid):
    dim = len(shape)
    if multigrid:
        # the two-electron model on a square grid: pair table, region masks and transfers
        desc = dict(potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_CAP, nCAP=16, multigrid=1)
        f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
        pd = T.prop_desc(T.MODE_RE, 2, 0.02, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_FLUX,
                         borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    else:
        desc = dict(absorber=T.ABS_CAP, nCAP=4)
        axis = T.AXIS_XY if dim == 2 else T.AXIS_XYZ
        f = [T.field_desc(axis, 0.1, 0.057, 2)]
        pd = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH,
                         outputs=T.OUT_NORM | T.OUT_EKIN | T.OUT_FLUX, borders=[(T.BORDER_BOX, 6)])
    return desc, f, pd


def oracle_prop(shape, dx, multigrid, seed):
    key = ("prop", shape, multigrid, seed)
    if key not in _ORACLE:
        desc, f, pd = prop_setup(shape, dx, multigrid)
        g = O.GridSet(T.grid_desc(len(shape), list(shape), dx, **desc))
        psi0 = state(int(np.prod(shape)), seed)
        for r in range(len(g)):
            g[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
        rec = O.Propagator(g, pd, f).run(STEPS)
        _ORACLE[key] = ([g[r].psi() for r in range(len(g))], rec)
    return _ORACLE[key]


def force(ctx, prec, row, n, m, v):
    ctx.tune_set(prec, row, n, m, v)


def device_fft(ctx, shape, dx, prec, seed):
    from qsf_amd import hip
    d = T.grid_desc(len(shape), list(shape), dx, precision=prec)
    g = hip.GridSet(ctx, d)
    g[0].set_psi(state(int(np.prod(shape)), seed))
    g[0].fourier(T.REP_P)
    out = g[0].psi()
    g.close()
    return out


def device_prop(ctx, shape, dx, prec, multigrid, seed):
    from qsf_amd import hip
    desc, f, pd = prop_setup(shape, dx, multigrid)
    g = hip.GridSet(ctx, T.grid_desc(len(shape), list(shape), dx, precision=prec, **desc))
    psi0 = state(int(np.prod(shape)), seed)
    for r in range(len(g)):
        g[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
    p = hip.Propagator(g, pd, f)
    rec = p.run(STEPS)
    out = [g[r].psi() for r in range(len(g))]
    p.close()
    g.close()
    return out, rec


def check_records(rd, ro, tol):
    assert rd.shape == ro.shape
    assert np.allclose(rd[:, :2], ro[:, :2], rtol=0, atol=1e-12)
    for c in range(2, ro.shape[1]):
        col = ro[:, c]
        if np.all(np.isnan(col)):
            continue
        scale = max(np.abs(col).max(), 1e-300)
        assert np.abs(rd[:, c] - col).max() <= tol * scale, f"record column {c}"


# (shape, tuning class of the pass that runs the tuned length) per dimensionality: 1 row pass,
# 0 column pass of an inner axis, 2 column pass of the outermost axis
def shapes_2d(n):
    return [((16, n), 1), ((n, 64), 2)]


def shapes_3d(n):
    return [((32, 64, n), 1), ((n, 32, 64), 2), ((32, n, 64), 0)]


def dxs(shape):
    return [0.3, 0.2, 0.25][:len(shape)]


@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED, T.FP32_STATE])
@pytest.mark.parametrize("n", TUNED)
def test_variants_fft_2d(ctx, n, prec):
    for shape, row in shapes_2d(n):
        m = int(np.prod(shape))
        ref = oracle_fft(shape, dxs(shape), 7)
        tol = fft_tol(prec)
        for v in VARIANTS:
            print(f"[case] fft shape={shape} prec={prec} row={row} variant={v}", flush=True)
            force(ctx, prec, row, n, m, v)
            err = rel(device_fft(ctx, shape, dxs(shape), prec, 7), ref)
            assert err < tol, f"shape {shape} variant {v}: {err:.3e}"


@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED, T.FP32_STATE])
@pytest.mark.parametrize("n", TUNED)
def test_variants_propagation_2d(ctx, n, prec):
    for shape, row in shapes_2d(n):
        m = int(np.prod(shape))
        ref_psi, ref_rec = oracle_prop(shape, dxs(shape), 0, 11)
        tol = prop_tol(prec)
        for v in VARIANTS:
            print(f"[case] propagation shape={shape} prec={prec} row={row} variant={v}", flush=True)
            force(ctx, prec, row, n, m, v)
            psi, rec = device_prop(ctx, shape, dxs(shape), prec, 0, 11)
            err = rel(psi[0], ref_psi[0])
            assert err < tol, f"shape {shape} variant {v}: {err:.3e}"
            check_records(rec, ref_rec, rec_tol(prec))


@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED, T.FP32_STATE])
@pytest.mark.parametrize("n", [256, 512])
def test_variants_3d(ctx, n, prec):
    for shape, row in shapes_3d(n):
        m = int(np.prod(shape))
        ref_fft = oracle_fft(shape, dxs(shape), 13)
        ref_psi, ref_rec = oracle_prop(shape, dxs(shape), 0, 17)
        for v in VARIANTS:
            print(f"[case] 3d shape={shape} prec={prec} row={row} variant={v}", flush=True)
            force(ctx, prec, row, n, m, v)
            err = rel(device_fft(ctx, shape, dxs(shape), prec, 13), ref_fft)
            assert err < fft_tol(prec), f"fft {shape} v{v}: {err:.3e}"
            psi, rec = device_prop(ctx, shape, dxs(shape), prec, 0, 17)
            err = rel(psi[0], ref_psi[0])
            assert err < prop_tol(prec), f"prop {shape} v{v}: {err:.3e}"
            check_records(rec, ref_rec, rec_tol(prec))


@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED])
@pytest.mark.parametrize("n", [256, 512])
def test_variants_multigrid_2d(ctx, n, prec):
    shape = (n, n)
    m = n * n
    ref_psi, ref_rec = oracle_prop(shape, [0.2, 0.2], 1, 19)
    tol = FP64_TOL if prec == T.FP64 else FP32_PROP_TOL
    for v in VARIANTS:
        print(f"[case] multigrid n={n} prec={prec} variant={v} (rows and columns)", flush=True)
        for cls in (0, 1, 2):
            force(ctx, prec, cls, n, m, v)
        psi, rec = device_prop(ctx, shape, [0.2, 0.2], prec, 1, 19)
        for r in range(4):
            err = rel(psi[r], ref_psi[r])
            assert err < tol, f"variant {v} region {r}: {err:.3e}"
        check_records(rec, ref_rec, 1e-9 if prec == T.FP64 else 1e-4)
