"""Writes the committed fixtures under tests/golden/ from the FP64 oracle (CPU only).

The reference holds no golden vectors (its one test hook writes mask CSVs at run time under
MPI+FFTW, wf/tests/multigrid.h), so these fixtures are regression vectors of the oracle, each
anchored by an independent check in tests/test_golden.py:
  dft.npz      seeded inputs, n in {8, 12, 16, 24, 64}; expected = long-double brute-force DFT
               (sign -1, unnormalised: FFTW_FORWARD as grid.h:257-267 plans it)
  masks.npz    maskRegion of every region of a 2D multigrid (grid.h:582-597), n in {64, 128},
               nCAP in {8, 16}, CAP (absorbers.h:16-60) and cos^(1/8)
  borders.npz  borVec of the ZOA_FLUX_2D region borders and BOX<20> on a 128x128 grid
               (fluxes/borders.h:38-257, borders_t.h:1-50), indexed by abs_centered_index
  traj.npz     20 loop bodies of SplitPropagator::run: 1D n=256 length gauge Strang with CAP and
               BOX flux; 2D 64x64 multigrid, velocity gauge, MultiProductSplit order 2, ZOA fluxes

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from qsf_amd import qsf_types as T  # noqa: E402

DFT_SIZES = (8, 12, 16, 24, 64)
MASK_CASES = [(64, 8, T.ABS_CAP), (64, 16, T.ABS_CAP), (128, 8, T.ABS_CAP), (128, 16, T.ABS_CAP),
              (64, 8, T.ABS_COS18), (128, 16, T.ABS_COS18)]
BORDER_N = 128
BORDER_KINDS = list(T.ZOA_FLUX_2D) + [T.BORDER_BOX]
TRAJ_STEPS = 20


def mask_desc(n, nCAP, absorber):
    return T.grid_desc(2, n, 0.3, absorber=absorber, nCAP=nCAP, absorber_tau=0.05, multigrid=1)


def traj1d_case():
    n, dx = 256, 0.2
    d = T.grid_desc(1, n, dx, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1,
                    absorber=T.ABS_CAP, nCAP=24)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    psi0 = (np.exp(-0.5 * (x - 2.0) ** 2) * np.exp(0.7j * x)).astype(np.complex128)
    f = [T.field_desc(T.AXIS_X, 0.0534, 0.057, 4)]
    p = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH,
                    outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_FLUX,
                    borders=[(T.BORDER_BOX, 12)])
    return d, p, f, psi0


def traj2d_case():
    n, dx = 64, 0.3
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_CAP, nCAP=8, multigrid=1)
    rng = np.random.default_rng(64)
    psi0 = (rng.normal(size=n * n) + 1j * rng.normal(size=n * n)) * 0.01
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    p = T.prop_desc(T.MODE_RE, 2, 0.02, gauge=T.GAUGE_VELOCITY,
                    outputs=T.OUT_NORM | T.OUT_FLUX, borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    return d, p, f, psi0


def run_oracle(case, steps=TRAJ_STEPS):
    d, p, f, psi0 = case
    gs = O.GridSet(d)
    for r in range(len(gs)):
        gs[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
    prop = O.Propagator(gs, p, f)
    rec = prop.run(steps)
    return rec, [gs[r].psi() for r in range(len(gs))]


def region_masks(n, nCAP, absorber):
    gs = O.GridSet(mask_desc(n, nCAP, absorber))
    out = []
    for r in range(len(gs)):
        gs[r].set_psi(np.ones(n * n, np.complex128))
        gs[r].mask_region()
        out.append(gs[r].psi().real.reshape(n, n))
    return np.stack(out)


def border_table(kind, size, n):
    tab = np.zeros((n, n, 2))
    for i in range(n):
        for j in range(n):
            tab[i, j] = O.border_vec_d(kind, size, 2, (i - n // 2, j - n // 2))
    return tab


def main():
    rng = np.random.default_rng(8)
    dft = {}
    for n in DFT_SIZES:
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        dft[f"x{n}"] = x
        dft[f"X{n}"] = O.dft_bruteforce(x, -1)
    np.savez_compressed(os.path.join(HERE, "dft.npz"), **dft)

    masks = {}
    for n, c, a in MASK_CASES:
        masks[f"n{n}_c{c}_a{a}"] = region_masks(n, c, a)
    np.savez_compressed(os.path.join(HERE, "masks.npz"), **masks)

    borders = {}
    for k in BORDER_KINDS:
        size = 20 if k == T.BORDER_BOX else 0
        borders[f"k{k}_s{size}"] = border_table(k, size, BORDER_N)
    np.savez_compressed(os.path.join(HERE, "borders.npz"), **borders)

    traj = {}
    rec, psis = run_oracle(traj1d_case())
    traj["rec1d"], traj["psi1d"] = rec, psis[0]
    rec, psis = run_oracle(traj2d_case())
    traj["rec2d"] = rec
    for r, p in enumerate(psis):
        traj[f"psi2d_r{r}"] = p
    np.savez_compressed(os.path.join(HERE, "traj.npz"), **traj)
    for f in ("dft.npz", "masks.npz", "borders.npz", "traj.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
