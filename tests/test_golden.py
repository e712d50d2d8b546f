"""The committed fixtures under tests/golden/ against the oracle and against what fixes them
independently of it (numpy's FFT, symmetries and ranges of the mask and border tables).

The fixtures are written by tests/golden/make_golden.py from the oracle; the reference itself
ships no golden vectors. GPU checks of the same fixtures live in tests/test_gpu_golden.py.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from qsf_amd import qsf_types as T

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_dft_fixture_is_the_definition():
    g = load("dft.npz")
    for n in (8, 12, 16, 24, 64):
        x, X = g[f"x{n}"], g[f"X{n}"]
        # numpy's FFT uses the same sign and scale as FFTW_FORWARD
        assert rel(np.fft.fft(x), X) < 1e-14, n
        assert rel(O.fft1d(x, -1), X) < 1e-14, n
        # backward unnormalised transform returns n*x
        assert rel(O.fft1d(X, +1) / n, x) < 1e-14, n


def test_mask_fixture_reproduced_bitwise_and_its_invariants():
    import sys
    sys.path.insert(0, GOLD)
    import make_golden as MG
    g = load("masks.npz")
    for n, c, a in MG.MASK_CASES:
        tab = g[f"n{n}_c{c}_a{a}"]
        assert np.array_equal(MG.region_masks(n, c, a), tab)
        for r in range(4):
            m = tab[r]
            assert np.array_equal(m, m[::-1, :]) and np.array_equal(m, m[:, ::-1]), (n, c, a, r)
            assert m.min() >= 0.0 and m.max() == 1.0
            # interior nodes (further than nCAP from every edge) are untouched
            assert np.all(m[c:n - c, c:n - c] == 1.0)
        # the last region (both axes free) is not masked (multigrid.h:617-625)
        assert np.all(tab[3] == 1.0)
        # the main region's corner node, both axes at the edge: CAP exp(-81 * 2); cos^(1/8) 0
        corner = tab[0][0, 0]
        if a == T.ABS_CAP:
            assert np.isclose(corner, np.exp(-162.0), rtol=1e-12)
        else:
            assert corner == 0.0


def test_border_fixture_reproduced_and_normalised():
    import sys
    sys.path.insert(0, GOLD)
    import make_golden as MG
    g = load("borders.npz")
    n = MG.BORDER_N
    for k in MG.BORDER_KINDS:
        size = 20 if k == T.BORDER_BOX else 0
        tab = g[f"k{k}_s{size}"]
        np.testing.assert_array_equal(MG.border_table(k, size, n), tab)
        norms = np.linalg.norm(tab, axis=-1)
        nz = norms > 0
        assert nz.any(), k
        assert np.allclose(norms[nz], 1.0, rtol=0, atol=1e-15), k
    box = g[f"k{T.BORDER_BOX}_s20"]
    # BOX<20>: faces at centred index 19 and -20 (borders.h:48-70); corners point diagonally
    i19, im20 = n // 2 + 19, n // 2 - 20
    assert np.allclose(box[i19, n // 2], [1, 0]) and np.allclose(box[im20, n // 2], [-1, 0])
    assert np.allclose(box[i19, i19], [np.sqrt(0.5), np.sqrt(0.5)])


@pytest.mark.parametrize("case", ["1d", "2d"])
def test_trajectory_fixture_reproduced(case):
    import sys
    sys.path.insert(0, GOLD)
    import make_golden as MG
    g = load("traj.npz")
    if case == "1d":
        rec, psis = MG.run_oracle(MG.traj1d_case())
        assert rel(psis[0], g["psi1d"]) < 1e-13
        np.testing.assert_allclose(rec, g["rec1d"], rtol=1e-12, atol=1e-15)
    else:
        rec, psis = MG.run_oracle(MG.traj2d_case())
        for r in range(4):
            assert rel(psis[r], g[f"psi2d_r{r}"]) < 1e-13, r
        np.testing.assert_allclose(rec, g["rec2d"], rtol=1e-12, atol=1e-15)
    rec = g[f"rec{case}"]
    # steps are consecutive and time advances by dt per step
    assert np.array_equal(rec[:, T.REC_STEP], np.arange(len(rec)))
    # the absorber only removes norm
    assert np.all(np.diff(rec[:, T.REC_NORM]) <= 1e-14)
