"""CoulombInteraction over two or more bounded coordinates on the device, against the oracle.

V = NEcharge / sqrt(sum_d x_d^2 + Nsoft) over the bounded coordinates (potential.h:40-44, the
parameter-pack fold of squares) does not separate into per-axis factors. The device applies it
as a pass of its own after the per-axis tables (coupling and absorber), so propagations of such
grids run unfused (one standalone call per reference member); multigrid regions with one bounded
coordinate keep the per-axis form of partial() (potential.h:62-98).

Bars: FP64 L2(psi) <= 1e-10 and records to 1e-11 (norm, energies) / 1e-8 (fluxes);
FP32-mixed L2(psi) <= 1e-4 and records to 1e-5.
"""
import numpy as np
import pytest

from oracle import oracle as O
from qsf_amd import qsf_types as T

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(77)
FP64_TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    from qsf_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def packet(shape, dx):
    axes = [-0.5 * dx * (n - 1) + dx * np.arange(n) for n in shape]
    grids = np.meshgrid(*axes, indexing="ij")
    r2 = sum(g * g for g in grids)
    psi = np.exp(-np.sqrt(r2 + 0.5)) * (1 + 0.2j * grids[-1])
    psi = psi.ravel().astype(np.complex128)
    return psi / np.sqrt(np.sum(np.abs(psi) ** 2) * dx ** len(shape))


def run_both(ctx, desc, pdesc, fields, psi0, steps):
    from qsf_amd import hip
    print(f"[case] coulomb dim={desc.dim} n={list(desc.n)[:desc.dim]} prec={desc.precision} "
          f"mg={desc.multigrid} gauge={pdesc.gauge}", flush=True)
    dev = hip.GridSet(ctx, desc)
    ref = O.GridSet(T.grid_desc(desc.dim, list(desc.n)[:desc.dim], list(desc.dx)[:desc.dim],
                                potential=desc.potential, Ncharge=desc.Ncharge, Echarge=desc.Echarge,
                                Nsoft=desc.Nsoft, Esoft=desc.Esoft, absorber=desc.absorber,
                                nCAP=desc.nCAP, absorber_tau=desc.absorber_tau, multigrid=desc.multigrid))
    for r in range(len(ref)):
        z = psi0 if r == 0 else np.zeros_like(psi0)
        dev[r].set_psi(z)
        ref[r].set_psi(z)
    pd = hip.Propagator(dev, pdesc, fields)
    po = O.Propagator(ref, pdesc, fields)
    rd = pd.run(steps)
    ro = po.run(steps)
    return dev, ref, rd, ro


@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED])
@pytest.mark.parametrize("gauge", [T.GAUGE_LENGTH, T.GAUGE_VELOCITY])
def test_coulomb_2d_propagation(ctx, prec, gauge):
    n, dx = 64, 0.25
    d = T.grid_desc(2, n, dx, precision=prec, potential=T.POT_COULOMB, Ncharge=2, Echarge=-1,
                    Nsoft=0.5, absorber=T.ABS_CAP, nCAP=8)
    f = [T.field_desc(T.AXIS_XY, 0.1, 0.057, 2)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=gauge,
                        outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_FLUX,
                        borders=[(T.BORDER_BOX, 20)])
    dev, ref, rd, ro = run_both(ctx, d, pdesc, f, packet((n, n), dx), 12)
    err = rel(dev[0].psi(), ref[0].psi())
    assert rd.shape == ro.shape
    if prec == T.FP64:
        assert err < FP64_TOL, err
        assert np.allclose(rd[:, 2:5], ro[:, 2:5], rtol=1e-11)
        assert np.allclose(rd[:, 7:], ro[:, 7:], rtol=1e-8, atol=1e-12 * np.abs(ro[:, 7:]).max())
    else:
        assert err < 1e-4, err
        assert np.allclose(rd[:, 2:5], ro[:, 2:5], rtol=1e-5)


def test_coulomb_3d_propagation(ctx):
    n, dx = 32, 0.3
    d = T.grid_desc(3, n, dx, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1,
                    absorber=T.ABS_COS18, nCAP=6, absorber_tau=0.05)
    f = [T.field_desc(T.AXIS_XYZ, 0.1, 0.057, 2)]
    pdesc = T.prop_desc(T.MODE_RE, 2, 0.05, gauge=T.GAUGE_LENGTH,
                        outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_FLUX, borders=[(T.BORDER_BOX, 10)])
    dev, ref, rd, ro = run_both(ctx, d, pdesc, f, packet((n, n, n), dx), 6)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    assert np.allclose(rd[:, 2:4], ro[:, 2:4], rtol=1e-11)
    assert np.allclose(rd[:, 7:], ro[:, 7:], rtol=1e-8, atol=1e-12 * np.abs(ro[:, 7:]).max())


def test_coulomb_2d_multigrid(ctx):
    # main region: both coordinates bounded (the non-separable factor); regions X, Y: one bounded
    # coordinate (partial() as per-axis tables); XY: none
    n, dx = 64, 0.3
    d = T.grid_desc(2, n, dx, potential=T.POT_COULOMB, Ncharge=2, Echarge=-1, Nsoft=0.5,
                    absorber=T.ABS_CAP, nCAP=8, multigrid=1)
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.02, gauge=T.GAUGE_VELOCITY,
                        outputs=T.OUT_NORM | T.OUT_FLUX, borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    dev, ref, rd, ro = run_both(ctx, d, pdesc, f, packet((n, n), dx), 8)
    for r in range(4):
        assert rel(dev[r].psi(), ref[r].psi()) < FP64_TOL, f"region {r}"
    assert np.allclose(rd[:, 2], ro[:, 2], rtol=1e-11)


def test_coulomb_ground_state_imaginary_time(ctx):
    # IM steps with renormalisation: the radial factor's real exp(-delta V) and E0 of the 2D model
    n, dx = 48, 0.3
    d = T.grid_desc(2, n, dx, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1)
    pdesc = T.prop_desc(T.MODE_IM, 1, 0.05, max_steps=40,
                        outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_NORMALIZE)
    f = [T.field_desc(T.AXIS_XY, 0.0, 0.057, 1)]
    dev, ref, rd, ro = run_both(ctx, d, pdesc, f, packet((n, n), dx), 40)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    assert np.allclose(rd[:, 2:6], ro[:, 2:6], rtol=1e-11)


def test_coulomb_fused_exec_refused(ctx):
    from qsf_amd import hip
    d = T.grid_desc(2, 32, 0.3, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1)
    g = hip.GridSet(ctx, d)
    g[0].set_psi(packet((32, 32), 0.3))
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, outputs=T.OUT_NORM)
    p = hip.Propagator(g, pdesc, [T.field_desc(T.AXIS_XY, 0.1, 0.057, 1)])
    with pytest.raises(Exception):
        from qsf_amd.hip import load
        p.ctx.check(load().qsf_prop_set_exec(p.h, 0), "prop_set_exec")
    p.close()
    g.close()
