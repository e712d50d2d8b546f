"""3D multigrid regions on the device against the oracle (wf/multigrid.h, basics/mpi_logic.h).

The eight regions of MultiRegionsReduced in 3D (NO, X, Y, Z, XY, XZ, YZ, XYZ; mpi_logic.h:249-263)
each keep their bounded axes in X and their free axes in P. postCompute (multigrid.h:617-625) is
transfer() in group order, then maskRegion() of every region but the last group: the slice of a
receiver gathers the freed axis' edge nodes of its source (getBox :454-482), masks them with
inv_mask and, for group-1 receivers, the 3D mask_correction (:519-530), transforms them along the
freed axis (FFTW forward, unnormalised) and adds them (:571-585). The device builds all boxes of
a slice at once (they tile the sliding axis; the transform never crosses them).

Bars: masks bit-exact (array_equal); transfers and propagation FP64 L2(psi) <= 1e-10 per region.
"""
import numpy as np
import pytest

from oracle import oracle as O
from qsf_amd import qsf_types as T

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(303)
FP64_TOL = 1e-10


def crand(n):
    return RNG.normal(size=n) + 1j * RNG.normal(size=n)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    from qsf_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def pair(ctx, d):
    from qsf_amd import hip
    print(f"[case] 3d multigrid n={list(d.n)} nCAP={d.nCAP} abs={d.absorber} prec={d.precision}", flush=True)
    return hip.GridSet(ctx, d), O.GridSet(d)


@pytest.mark.parametrize("absorber", [T.ABS_CAP, T.ABS_COS18])
def test_mask_region_3d_bit_exact(ctx, absorber):
    d = T.grid_desc(3, 32, 0.4, absorber=absorber, nCAP=8, absorber_tau=0.05, multigrid=1)
    dev, ref = pair(ctx, d)
    assert len(dev) == 8 and len(ref) == 8
    for r in range(8):
        x = crand(dev[r].m)
        dev[r].set_psi(x)
        ref[r].set_psi(x)
        dev[r].mask_region()
        ref[r].mask_region()
        assert np.array_equal(dev[r].psi(), ref[r].psi()), f"region {r}"


@pytest.mark.parametrize("absorber,n,nCAP", [(T.ABS_CAP, 32, 8), (T.ABS_COS18, 32, 16), (T.ABS_CAP, 48, 12)])
def test_post_compute_3d(ctx, absorber, n, nCAP):
    d = T.grid_desc(3, n, 0.4, absorber=absorber, nCAP=nCAP, absorber_tau=0.05, multigrid=1)
    dev, ref = pair(ctx, d)
    for r in range(8):
        x = crand(dev[r].m)
        dev[r].set_psi(x)
        ref[r].set_psi(x)
    dev.post_compute()
    ref.post_compute()
    for r in range(8):
        assert rel(dev[r].psi(), ref[r].psi()) < FP64_TOL, f"region {r}"


def test_post_compute_3d_fp32(ctx):
    d = T.grid_desc(3, 32, 0.4, precision=T.FP32_MIXED, absorber=T.ABS_CAP, nCAP=8, multigrid=1)
    dev = pair(ctx, d)[0]
    ref = O.GridSet(T.grid_desc(3, 32, 0.4, absorber=T.ABS_CAP, nCAP=8, multigrid=1))
    for r in range(8):
        x = crand(dev[r].m).astype(np.complex64).astype(np.complex128)
        dev[r].set_psi(x)
        ref[r].set_psi(x)
    dev.post_compute()
    ref.post_compute()
    for r in range(8):
        assert rel(dev[r].psi(), ref[r].psi()) < 2e-6, f"region {r}"


def test_propagation_multigrid_3d(ctx):
    from qsf_amd import hip
    n, dx = 32, 0.4
    d = T.grid_desc(3, n, dx, potential=T.POT_ECKHARDT_SACHA, Ncharge=3, Echarge=-1, Nsoft=1, Esoft=1,
                    absorber=T.ABS_CAP, nCAP=8, multigrid=1)
    dev, ref = pair(ctx, d)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    g = np.exp(-np.sqrt(x * x + 1.0))
    psi0 = (g[:, None, None] * g[None, :, None] * g[None, None, :]).ravel().astype(np.complex128)
    psi0 *= (1 + 0.1j * np.sin(np.arange(psi0.size)))
    psi0 /= np.sqrt(np.sum(np.abs(psi0) ** 2) * dx ** 3)
    for r in range(8):
        z = psi0 if r == 0 else np.zeros_like(psi0)
        dev[r].set_psi(z)
        ref[r].set_psi(z)
    f = [T.field_desc(T.AXIS_XYZ, 0.12, 0.057, 2)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM)
    pd = hip.Propagator(dev, pdesc, f)
    po = O.Propagator(ref, pdesc, f)
    rd = pd.run(6)
    ro = po.run(6)
    for r in range(8):
        assert rel(dev[r].psi(), ref[r].psi()) < FP64_TOL, f"region {r}"
    assert np.allclose(rd[:, 2], ro[:, 2], rtol=1e-11)
    pd.close()


def test_multigrid_3d_refusals(ctx):
    from qsf_amd import hip
    for d in [T.grid_desc(3, [32, 32, 16], 0.4, absorber=T.ABS_CAP, nCAP=4, multigrid=1),   # not cubic
              T.grid_desc(3, 32, 0.4, absorber=T.ABS_CAP, nCAP=12, multigrid=1),          # 12 does not divide 32
              T.grid_desc(3, 32, 0.4, potential=T.POT_COULOMB, absorber=T.ABS_CAP, nCAP=8, multigrid=1)]:
        with pytest.raises(Exception):
            hip.GridSet(ctx, d)
