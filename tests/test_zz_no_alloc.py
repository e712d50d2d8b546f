"""qsf_average_noalloc (the entry point that allocates no device memory) against qsf_average.

Both run the same observation pass (same kernel variant, hence the same blocks) and the same
fixed-order reduction, so the results are bit-identical in FP64 and FP32-mixed. Two grids in a
row, the second smaller, so the second call finds the first one's partial sums in the shared
scratch and must not depend on them; the pending-step refusal is checked from the host.
"""
import numpy as np
import pytest

from qsf_amd import qsf_types as T

pytestmark = pytest.mark.gpu


def crand(m, seed):
    r = np.random.default_rng(seed)
    return (r.normal(size=m) + 1j * r.normal(size=m)) * 0.01


CASES = [
    # (description, grid) — the first large (many blocks, two reduction levels), then smaller
    ("3D 128^3 Eckhardt-Sacha FP64", T.grid_desc(3, 128, 0.3, potential=T.POT_ECKHARDT_SACHA, Ncharge=3,
                                                 Echarge=-1, Nsoft=1, Esoft=1, absorber=T.ABS_CAP, nCAP=8)),
    ("2D 96^2 Eberly FP64", T.grid_desc(2, 96, 0.2, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5,
                                         Esoft=0.33)),
    ("2D 64^2 Eberly FP32-mixed", T.grid_desc(2, 64, 0.2, precision=T.FP32_MIXED, potential=T.POT_EBERLY,
                                              Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33)),
    ("1D 256 Coulomb FP64", T.grid_desc(1, 256, 0.1, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1)),
]


def test_noalloc_average_equals_average():
    from qsf_amd import hip
    ctx = hip.Context(0)
    for k, (name, d) in enumerate(CASES):
        print(f"[case] noalloc {name}", flush=True)
        gs = hip.GridSet(ctx, d)
        gs[0].set_psi(crand(gs[0].m, k))
        for op in (0, 1):
            a = gs[0].average(T.REP_X, op)
            b = gs[0].average_noalloc(T.REP_X, op)
            assert a == b, (name, op, a, b)
            # twice in a row: the scratch keeps the last call's contents, the result does not
            assert gs[0].average_noalloc(T.REP_X, op) == a
        gs[0].fourier(T.REP_P)
        for op in (0, 2):
            a = gs[0].average(T.REP_P, op)
            assert gs[0].average_noalloc(T.REP_P, op) == a, (name, op)
        gs.close()
    ctx.close()


def test_noalloc_refuses_pending_step():
    from qsf_amd import hip
    ctx = hip.Context(0)
    d = T.grid_desc(2, 64, 0.2, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_CAP, nCAP=8)
    gs = hip.GridSet(ctx, d)
    gs[0].set_psi(crand(gs[0].m, 9))
    p = hip.Propagator(gs, T.prop_desc(T.MODE_RE, 1, 0.02, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM),
                       [T.field_desc(T.AXIS_XY, 0.1, 0.057, 2)])
    p.start()
    p.run_steps(3)
    with pytest.raises(hip.QSFError, match="pending"):
        gs[0].average_noalloc(T.REP_X, 0)
    a = gs[0].average(T.REP_X, 0)          # settles the grid
    assert gs[0].average_noalloc(T.REP_X, 0) == a
    p.close()
    gs.close()
    ctx.close()
