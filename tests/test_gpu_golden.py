"""The HIP path against the committed fixtures under tests/golden/ (no oracle run needed).

Bars: FFT and trajectories within the FP64 bar (relative L2 <= 1e-10); region masks bit-exact.
"""
import os
import sys

import numpy as np
import pytest

from qsf_amd import qsf_types as T

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import make_golden as MG  # noqa: E402

FP64_TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    from qsf_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("n", MG.DFT_SIZES)
def test_fft_matches_dft_fixture(ctx, n):
    from qsf_amd import hip
    g = np.load(os.path.join(GOLD, "dft.npz"))
    print(f"[case] dft n={n}", flush=True)
    gs = hip.GridSet(ctx, T.grid_desc(1, n, 0.1))
    gs[0].set_psi(g[f"x{n}"])
    gs[0].fourier(T.REP_P)
    assert rel(gs[0].psi(), g[f"X{n}"]) < 1e-13
    gs.close()


@pytest.mark.parametrize("case", MG.MASK_CASES)
def test_region_masks_match_fixture_bitwise(ctx, case):
    from qsf_amd import hip
    n, c, a = case
    tab = np.load(os.path.join(GOLD, "masks.npz"))[f"n{n}_c{c}_a{a}"]
    print(f"[case] masks n={n} nCAP={c} absorber={a}", flush=True)
    gs = hip.GridSet(ctx, MG.mask_desc(n, c, a))
    for r in range(4):
        gs[r].set_psi(np.ones(n * n, np.complex128))
        gs[r].mask_region()
        out = gs[r].psi()
        assert np.array_equal(out.real.reshape(n, n), tab[r]), f"region {r}"
        assert np.all(out.imag == 0.0)
    gs.close()


@pytest.mark.parametrize("case", ["1d", "2d"])
@pytest.mark.parametrize("exec_mode", [0, 1])
def test_trajectory_matches_fixture(ctx, case, exec_mode):
    from qsf_amd import hip
    g = np.load(os.path.join(GOLD, "traj.npz"))
    d, p, f, psi0 = MG.traj1d_case() if case == "1d" else MG.traj2d_case()
    print(f"[case] trajectory {case} exec={exec_mode}", flush=True)
    gs = hip.GridSet(ctx, d)
    for r in range(len(gs)):
        gs[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
    prop = hip.Propagator(gs, p, f, exec_mode=exec_mode)
    rec = prop.run(MG.TRAJ_STEPS)
    ref = g[f"rec{case}"]
    assert rec.shape == ref.shape
    np.testing.assert_allclose(rec[:, :2], ref[:, :2], rtol=0, atol=1e-12)
    np.testing.assert_allclose(rec[:, 2:7], ref[:, 2:7], rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(rec[:, 7:], ref[:, 7:], rtol=1e-8, atol=1e-12 * np.abs(ref[:, 7:]).max())
    if case == "1d":
        assert rel(gs[0].psi(), g["psi1d"]) < FP64_TOL
    else:
        for r in range(4):
            assert rel(gs[r].psi(), g[f"psi2d_r{r}"]) < FP64_TOL, f"region {r}"
    prop.close()
    gs.close()
