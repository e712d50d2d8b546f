"""Parity of the HIP path (libqsf_hip.so through its C ABI) with the CPU FP64 oracle.

Tolerances: FP64 kernels L2(ψ) relative error ≤ 1e-10 (the required bar for the FP64 path;
actual errors are ~1e-14); FP32-mixed kernels ≤ 2e-5 relative per operation. Bit-exact
operations (region masks, border tables and their indexing) are compared with array_equal.
"""
import numpy as np
import pytest

from oracle import oracle as O
from qsf_amd import qsf_types as T

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2024)
FP64_TOL = 1e-10
FP32_TOL = 2e-5


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


def pair(ctx, desc):
    from qsf_amd import hip
    print(f"[case] dim={desc.dim} n={list(desc.n)[:desc.dim]} prec={desc.precision} "
          f"pot={desc.potential} abs={desc.absorber} mg={desc.multigrid}", flush=True)
    return hip.GridSet(ctx, desc), O.GridSet(desc)


# ------------------------------------------------------------------------------------ FFT
SIZES_1D = [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 12, 24, 48, 96, 192, 384, 768, 1536, 3072,
            10, 20, 40, 80, 160, 320, 640, 1280, 2560]


@pytest.mark.parametrize("n", SIZES_1D)
@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED])
def test_fft_1d(ctx, n, prec):
    d = T.grid_desc(1, n, 0.1, precision=prec)
    dev, ref = pair(ctx, d)
    x = crand(n)
    dev[0].set_psi(x)
    ref[0].set_psi(x)
    dev[0].fourier(T.REP_P)
    ref[0].fourier(T.REP_P)
    tol = FP64_TOL if prec == T.FP64 else FP32_TOL
    assert rel(dev[0].psi(), ref[0].psi()) < tol
    dev[0].fourier(T.REP_X)
    assert rel(dev[0].psi(), x) < tol


@pytest.mark.parametrize("shape", [(64, 48), (256, 256), (2048, 16), (12, 1024), (80, 320), (640, 40)])
@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED])
def test_fft_2d_transposed_layout(ctx, shape, prec):
    d = T.grid_desc(2, list(shape), [0.2, 0.3], precision=prec)
    dev, ref = pair(ctx, d)
    x = crand(shape[0] * shape[1])
    for g in (dev[0], ref[0]):
        g.set_psi(x)
        g.fourier(T.REP_P)
    tol = FP64_TOL if prec == T.FP64 else FP32_TOL
    assert rel(dev[0].psi(), ref[0].psi()) < tol   # both in FFTW's (n1, n0) P layout


@pytest.mark.parametrize("shape", [(16, 32, 8), (32, 32, 32), (64, 16, 128), (40, 20, 80)])
def test_fft_3d(ctx, shape):
    d = T.grid_desc(3, list(shape), [0.3, 0.3, 0.3])
    dev, ref = pair(ctx, d)
    x = crand(int(np.prod(shape)))
    for g in (dev[0], ref[0]):
        g.set_psi(x)
        g.fourier(T.REP_P)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    for g in (dev[0], ref[0]):
        g.fourier(T.REP_X)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL


# ------------------------------------------------------------------------------------ evolve
CASES = [
    dict(dim=1, n=256, dx=0.1, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1,
         absorber=T.ABS_CAP, nCAP=26),
    dict(dim=1, n=512, dx=0.1, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1,
         absorber=T.ABS_COS18, nCAP=51, absorber_tau=0.05),
    dict(dim=2, n=64, dx=0.2, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5,
         Esoft=0.33, absorber=T.ABS_CAP, nCAP=8),
    dict(dim=3, n=32, dx=0.3, potential=T.POT_ECKHARDT_SACHA, Ncharge=3, Echarge=-1, Nsoft=1,
         Esoft=1, absorber=T.ABS_CAP, nCAP=6),
    # CoulombInteraction over every coordinate (no per-axis separation: potential.h:40-44)
    dict(dim=2, n=64, dx=0.2, potential=T.POT_COULOMB, Ncharge=2, Echarge=-1, Nsoft=0.5,
         absorber=T.ABS_CAP, nCAP=8),
    dict(dim=3, n=32, dx=0.3, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1,
         absorber=T.ABS_COS18, nCAP=6, absorber_tau=0.05),
]


def desc_of(c, prec=T.FP64, **kw):
    c = dict(c)
    c.update(kw)
    dim, n, dx = c.pop("dim"), c.pop("n"), c.pop("dx")
    return T.grid_desc(dim, n, dx, precision=prec, **c)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("mode", [T.MODE_RE, T.MODE_IM])
def test_evolve_x_and_p(ctx, case, mode):
    d = desc_of(CASES[case])
    dev, ref = pair(ctx, d)
    x = crand(dev[0].m) * 0.1
    for g in (dev[0], ref[0]):
        g.set_psi(x)
    C = np.array([0.031, -0.017, 0.009])
    for g in (dev[0], ref[0]):
        g.evolve(mode, T.REP_X, 0.025, C, T.REP_X)   # length gauge coupling in X
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    for g in (dev[0], ref[0]):
        g.fourier(T.REP_P)
        g.evolve(mode, T.REP_P, 0.05, C, T.REP_P)    # velocity gauge coupling in P
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    for g in (dev[0], ref[0]):
        g.fourier(T.REP_X)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL


@pytest.mark.parametrize("case", range(len(CASES)))
def test_averages(ctx, case):
    d = desc_of(CASES[case])
    dev, ref = pair(ctx, d)
    x = crand(dev[0].m) * 0.1
    for g in (dev[0], ref[0]):
        g.set_psi(x)
    for op in (0, 1):
        assert np.isclose(dev[0].average(T.REP_X, op), ref[0].average(T.REP_X, op), rtol=1e-12)
    for g in (dev[0], ref[0]):
        g.fourier(T.REP_P)
    assert np.isclose(dev[0].average(T.REP_P, 2), ref[0].average(T.REP_P, 2), rtol=1e-12)
    assert np.isclose(dev[0].average(T.REP_P, 0), ref[0].average(T.REP_P, 0), rtol=1e-12)
    for g in (dev[0], ref[0]):
        g.fourier(T.REP_X)
        g.normalize()
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL


# ------------------------------------------------------------------------------------ flux
@pytest.mark.parametrize("dim,n,borders", [
    (1, 256, [(T.BORDER_BOX, 100)]),
    (2, 128, [(T.BORDER_BOX, 40)] + [(b, 0) for b in T.ZOA_FLUX_2D]),
    (3, 32, [(T.BORDER_BOX, 10), (T.BORDER_N2S, 0), (T.BORDER_D2T, 0)]),
])
def test_flux(ctx, dim, n, borders):
    d = T.grid_desc(dim, n, 0.5)
    dev, ref = pair(ctx, d)
    x = crand(dev[0].m)
    for g in (dev[0], ref[0]):
        g.set_psi(x)
    A = [0.2, -0.1, 0.05]
    for vg in (None, A):
        a = dev[0].flux(borders, vg)
        b = ref[0].flux(borders, vg)
        assert np.allclose(a, b, rtol=1e-11, atol=1e-12 * np.abs(b).max())


# ------------------------------------------------------------------------------------ masks
@pytest.mark.parametrize("absorber", [T.ABS_CAP, T.ABS_COS18])
def test_mask_region_bit_exact(ctx, absorber):
    d = T.grid_desc(2, 64, 0.5, absorber=absorber, nCAP=12, multigrid=1)
    dev, ref = pair(ctx, d)
    for r in range(4):
        x = crand(dev[r].m)
        dev[r].set_psi(x)
        ref[r].set_psi(x)
        dev[r].mask_region()
        ref[r].mask_region()
        assert np.array_equal(dev[r].psi(), ref[r].psi()), f"region {r}"


def test_multigrid_post_compute(ctx):
    d = T.grid_desc(2, 64, 0.5, absorber=T.ABS_CAP, nCAP=8, multigrid=1)
    dev, ref = pair(ctx, d)
    for r in range(4):
        x = crand(dev[r].m)
        dev[r].set_psi(x)
        ref[r].set_psi(x)
    dev.post_compute()
    ref.post_compute()
    for r in range(4):
        assert rel(dev[r].psi(), ref[r].psi()) < FP64_TOL, f"region {r}"


# ------------------------------------------------------------------------------------ propagation
def run_both(ctx, desc, pdesc, fields, psi0, limit, exec_mode=0):
    from qsf_amd import hip
    dev = hip.GridSet(ctx, desc)
    ref = O.GridSet(desc)
    for r in range(len(ref)):
        dev[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
        ref[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
    pd = hip.Propagator(dev, pdesc, fields, exec_mode=exec_mode)
    po = O.Propagator(ref, pdesc, fields)
    rd = pd.run(limit)
    ro = po.run(limit)
    return dev, ref, pd, po, rd, ro


def check_flux_integral(pd, rd, ro, dt, rtol, atol):
    # the device's FP64 flux time integral: every record row (the start row included) adds
    # flux * dt, on the fused and the member-by-member path alike. Against the device's own rows
    # up to summation order, and against the oracle's rows within the row tolerance (rtol, atol)
    got = pd.flux_integral()
    fd, fo = rd[:, T.REC_FLUX0:], ro[:, T.REC_FLUX0:]
    assert got.shape == (fd.shape[1],)
    assert np.allclose(got, fd.sum(axis=0) * dt, rtol=1e-12, atol=1e-15 * np.abs(fd).sum() * dt)
    assert np.allclose(got, fo.sum(axis=0) * dt, rtol=rtol, atol=atol * dt * fo.shape[0])


@pytest.mark.parametrize("exec_mode", [0, 1])
def test_propagation_1d_length_gauge(ctx, exec_mode):
    n, dx = 512, 0.1
    d = T.grid_desc(1, n, dx, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1,
                    absorber=T.ABS_COS18, nCAP=51, absorber_tau=0.05)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    psi0 = (np.exp(-x * x) + 0j) / np.sqrt(np.sqrt(np.pi / 2))
    f = [T.field_desc(T.AXIS_X, 0.0534, 0.057, 4)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH,
                        outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_FLUX,
                        borders=[(T.BORDER_BOX, 150)])
    dev, ref, pd, po, rd, ro = run_both(ctx, d, pdesc, f, psi0, 20, exec_mode)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    assert rd.shape == ro.shape
    assert np.allclose(rd[:, :2], ro[:, :2], rtol=0, atol=1e-12)   # step, time
    assert np.allclose(rd[:, 2:6], ro[:, 2:6], rtol=1e-11, atol=1e-13)
    assert np.allclose(rd[:, 7:], ro[:, 7:], rtol=1e-9, atol=1e-13)
    check_flux_integral(pd, rd, ro, 0.05, 1e-9, 1e-13)


@pytest.mark.parametrize("exec_mode", [0, 1])
def test_propagation_2d_velocity_gauge_order2(ctx, exec_mode):
    n, dx = 128, 0.2
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_CAP, nCAP=16)
    psi0 = crand(n * n) * 0.01
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    pdesc = T.prop_desc(T.MODE_RE, 2, 0.02, gauge=T.GAUGE_VELOCITY,
                        outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_FLUX,
                        borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    dev, ref, pd, po, rd, ro = run_both(ctx, d, pdesc, f, psi0, 15, exec_mode)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    assert np.allclose(rd[:, 2:4], ro[:, 2:4], rtol=1e-11)
    assert np.allclose(rd[:, 7:], ro[:, 7:], rtol=1e-8, atol=1e-12 * np.abs(ro[:, 7:]).max())
    check_flux_integral(pd, rd, ro, 0.02, 1e-8, 1e-12 * np.abs(ro[:, 7:]).max())


@pytest.mark.parametrize("exec_mode,n", [(0, 32), (1, 32), (0, 64)])
def test_propagation_3d_length_gauge(ctx, exec_mode, n):
    # n = 64 takes the patched (i0, i1) line order of the 3D row pass; 32 the linear one
    dx = 0.3
    d = T.grid_desc(3, n, dx, potential=T.POT_ECKHARDT_SACHA, Ncharge=3, Echarge=-1, Nsoft=1, Esoft=1,
                    absorber=T.ABS_CAP, nCAP=6)
    psi0 = crand(n ** 3) * 0.01
    f = [T.field_desc(T.AXIS_XYZ, 0.1, 0.057, 2)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH,
                        outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_FLUX,
                        borders=[(T.BORDER_BOX, 10)])
    dev, ref, pd, po, rd, ro = run_both(ctx, d, pdesc, f, psi0, 10, exec_mode)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    assert np.allclose(rd[:, 2:5], ro[:, 2:5], rtol=1e-11)
    assert np.allclose(rd[:, 7:], ro[:, 7:], rtol=1e-8, atol=1e-12 * np.abs(ro[:, 7:]).max())


def test_propagation_multigrid_2d(ctx):
    n, dx = 64, 0.3
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_CAP, nCAP=8, multigrid=1)
    psi0 = crand(n * n) * 0.01
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    pdesc = T.prop_desc(T.MODE_RE, 2, 0.02, gauge=T.GAUGE_VELOCITY,
                        outputs=T.OUT_NORM | T.OUT_FLUX, borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    for exec_mode in (0, 1):
        dev, ref, pd, po, rd, ro = run_both(ctx, d, pdesc, f, psi0, 8, exec_mode)
        for r in range(4):
            assert rel(dev[r].psi(), ref[r].psi()) < FP64_TOL, f"exec {exec_mode} region {r}"
        assert np.allclose(rd[:, 2], ro[:, 2], rtol=1e-11)


@pytest.mark.parametrize("multigrid", [0, 1])
def test_state_between_steps(ctx, multigrid):
    # A worker at WHEN::DURING (propagator.h:446-452) reads and may change the state between
    # steps. The grid entry points first apply what the fused pipeline left pending (inverse
    # transform, trailing X half, multigrid transfer), so the state read is the oracle's after the
    # same steps, a state written is the one the next step starts from, and the records of the
    # interrupted run equal those of an uninterrupted one.
    from qsf_amd import hip
    n, dx, steps = 64, 0.3, 10
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_CAP, nCAP=8, multigrid=multigrid)
    psi0 = crand(n * n) * 0.01
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    pdesc = T.prop_desc(T.MODE_RE, 2, 0.02, gauge=T.GAUGE_VELOCITY,
                        outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_FLUX, borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    print(f"[case] state between steps multigrid={multigrid}", flush=True)

    def loaded(gs):
        for r in range(len(gs)):
            gs[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
        return gs

    dev, ref = loaded(hip.GridSet(ctx, d)), loaded(O.GridSet(d))
    pd, po = hip.Propagator(dev, pdesc, f), O.Propagator(ref, pdesc, f)
    pd.start()
    po.start()
    for k in range(steps):
        pd.step()
        po.step()
        if k in (3, 6):
            for r in range(len(ref)):
                assert rel(dev[r].psi(), ref[r].psi()) < FP64_TOL, f"step {k} region {r}"
    rd, ro = pd.records(), po.records()
    np.testing.assert_allclose(rd[:, 2:4], ro[:, 2:4], rtol=1e-11)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL

    plain = loaded(hip.GridSet(ctx, d))
    rp = hip.Propagator(plain, pdesc, f).run(steps)
    assert rd.shape == rp.shape
    np.testing.assert_allclose(rd, rp, rtol=1e-12, atol=1e-14 * np.nanmax(np.abs(rp)))
    for r in range(len(ref)):
        assert rel(dev[r].psi(), plain[r].psi()) < 1e-12

    # a state written between steps is where the next step starts
    dev, ref = loaded(hip.GridSet(ctx, d)), loaded(O.GridSet(d))
    pd, po = hip.Propagator(dev, pdesc, f), O.Propagator(ref, pdesc, f)
    pd.start()
    po.start()
    for k in range(6):
        pd.step()
        po.step()
        if k == 2:
            half = 0.5 * ref[0].psi()
            dev[0].set_psi(half)
            ref[0].set_psi(half)
    rd, ro = pd.records(), po.records()
    for r in range(len(ref)):
        assert rel(dev[r].psi(), ref[r].psi()) < FP64_TOL, f"region {r}"
    np.testing.assert_allclose(rd[:, 2:4], ro[:, 2:4], rtol=1e-11)


def test_imaginary_time_ground_state(ctx):
    n, dx = 256, 0.2
    d = T.grid_desc(1, n, dx, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    psi0 = np.exp(-x * x / 2) + 0j
    pdesc = T.prop_desc(T.MODE_IM, 1, 0.05, max_steps=200,
                        outputs=T.OUT_NORMALIZE | T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT
                        | T.OUT_EDIFF)
    dev, ref, pd, po, rd, ro = run_both(ctx, d, pdesc, [], psi0, None)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    assert np.allclose(rd[:, 2:7], ro[:, 2:7], rtol=1e-10, atol=1e-13)
    assert abs(rd[-1, T.REC_ETOT] - (-0.6698)) < 2e-3


def test_fp32_mixed_step_close_to_fp64(ctx):
    n, dx = 128, 0.2
    base = dict(potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                absorber=T.ABS_CAP, nCAP=16)
    psi0 = crand(n * n) * 0.01
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.02, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM)
    from qsf_amd import hip
    out = []
    for prec in (T.FP64, T.FP32_MIXED):
        g = hip.GridSet(ctx, T.grid_desc(2, n, dx, precision=prec, **base))
        g[0].set_psi(psi0)
        p = hip.Propagator(g, pdesc, f)
        rec = p.run(50)
        out.append((g[0].psi(), rec))
    assert rel(out[1][0], out[0][0]) < 1e-4
    assert np.allclose(out[1][1][:, 2], out[0][1][:, 2], rtol=1e-5)


def test_grid_teardown_with_unbuilt_tables(ctx):
    # a propagator stopped mid-run leaves phase tables queued for the next pass; they belong to
    # their grid and must be dropped with it, not built into freed memory by another grid
    from qsf_amd import hip
    base = dict(potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                absorber=T.ABS_CAP, nCAP=16)
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    for prec in (T.FP64, T.FP32_MIXED, T.FP64):
        g = hip.GridSet(ctx, T.grid_desc(2, 128, 0.2, precision=prec, **base))
        g[0].set_psi(crand(128 * 128) * 0.01)
        p = hip.Propagator(g, T.prop_desc(T.MODE_RE, 1, 0.02, outputs=0), f)
        p.start()
        p.run_steps(3)
        p.close()
        g.close()
    ctx.synchronize()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_single_solves(ctx, use_graph):
    # uneven sizes and fields in one batch; every solve equals its own single-solve propagation
    # (same kernels: up to summation order of the observables) and the FP64 oracle
    from qsf_amd import hip
    base = dict(potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                absorber=T.ABS_CAP, nCAP=8)
    sizes = [64, 96, 64, 128, 48]
    descs = [T.grid_desc(2, n, 0.3, **base) for n in sizes]
    fields = [T.field_desc(T.AXIS_XY, 0.05 + 0.02 * i, 2.0, 1 + (i % 2), phase_in_pi=0.3 * i)
              for i in range(len(sizes))]
    borders = [(b, 0) for b in T.ZOA_FLUX_2D] + [(T.BORDER_BOX, 18)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=borders)
    psis = [crand(n * n) * 0.01 for n in sizes]
    b = hip.Batch(ctx, descs, fields, pdesc)
    for s, p in enumerate(psis):
        b.set_psi(s, p)
    b.run_all(use_graph=use_graph)
    for s, n in enumerate(sizes):
        print(f"[case] batch solve {s} n={n}", flush=True)
        ref = O.GridSet(descs[s])
        ref[0].set_psi(psis[s])
        ro = O.Propagator(ref, pdesc, [fields[s]]).run()
        rb = b.records(s)
        assert rb.shape[0] == ro.shape[0]
        assert rel(b.psi(s), ref[0].psi()) < FP64_TOL
        assert np.allclose(rb[:, 2], ro[:, T.REC_NORM], rtol=1e-11)
        assert np.allclose(rb[:, 3:], ro[:, T.REC_FLUX0:], rtol=1e-8, atol=1e-12 * np.abs(ro[:, T.REC_FLUX0:]).max())
    b.close()


@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED, T.FP32_STATE])
def test_batch_mixed_dimensions(ctx, prec):
    # 1D and 2D solves in one batch (one pass per step for the 1D groups, two for the 2D groups),
    # uneven sizes including 3*2^k and 5*2^k, replayed as a hipGraph; each solve against its own
    # FP64 oracle propagation
    from qsf_amd import hip
    specs = [(1, 256, dict(potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1.0,
                           absorber=T.ABS_COS18, nCAP=26, absorber_tau=0.05), 0.1),
             (1, 384, dict(potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1.0,
                           absorber=T.ABS_CAP, nCAP=38), 0.1),
             (2, 64, dict(potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                          absorber=T.ABS_CAP, nCAP=8), 0.3),
             (1, 320, dict(potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1.0,
                           absorber=T.ABS_CAP, nCAP=32), 0.1),
             (2, 96, dict(potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                          absorber=T.ABS_COS18, nCAP=10, absorber_tau=0.02), 0.3),
             (1, 256, dict(potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1.0,
                           absorber=T.ABS_COS18, nCAP=26, absorber_tau=0.05), 0.1)]
    descs = [T.grid_desc(dim, n, dx, precision=prec, **kw) for dim, n, kw, dx in specs]
    descs64 = [T.grid_desc(dim, n, dx, **kw) for dim, n, kw, dx in specs]
    fields = [T.field_desc(T.AXIS_X if dim == 1 else T.AXIS_XY, 0.05 + 0.02 * i, 2.0, 1 + (i % 2),
                           phase_in_pi=0.3 * i) for i, (dim, n, kw, dx) in enumerate(specs)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=[(T.BORDER_BOX, 12)])
    psis = []
    for dim, n, kw, dx in specs:
        x = crand(n ** dim) * 0.01
        psis.append(x.astype(np.complex64).astype(np.complex128) if prec != T.FP64 else x)
    b = hip.Batch(ctx, descs, fields, pdesc)
    for s, p in enumerate(psis):
        b.set_psi(s, p)
    b.run_all(use_graph=True)
    # FP32 state with FP64 arithmetic: one rounding per element per store
    tol = {T.FP64: FP64_TOL, T.FP32_MIXED: 1e-4, T.FP32_STATE: 2e-6}[prec]
    norm_rtol = {T.FP64: 1e-11, T.FP32_MIXED: 1e-5, T.FP32_STATE: 1e-7}[prec]
    for s, (dim, n, kw, dx) in enumerate(specs):
        print(f"[case] mixed batch solve {s} dim={dim} n={n} prec={prec}", flush=True)
        ref = O.GridSet(descs64[s])
        ref[0].set_psi(psis[s])
        ro = O.Propagator(ref, pdesc, [fields[s]]).run()
        rb = b.records(s)
        assert rb.shape[0] == ro.shape[0]
        assert rel(b.psi(s), ref[0].psi()) < tol
        assert np.allclose(rb[:, 2], ro[:, T.REC_NORM], rtol=norm_rtol)
        if prec == T.FP64:
            assert np.allclose(rb[:, 3:], ro[:, T.REC_FLUX0:], rtol=1e-8,
                               atol=1e-12 * max(np.abs(ro[:, T.REC_FLUX0:]).max(), 1e-300))
        if prec == T.FP32_STATE:
            fl = ro[:, T.REC_FLUX0:]
            assert np.abs(rb[:, 3:] - fl).max() <= 1e-5 * max(np.abs(fl).max(), 1e-300)
    b.close()


@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED])
def test_batch_stream_count_bit_identical(ctx, prec, monkeypatch):
    # size groups share nothing within a step but the P tables and the step counter, so running
    # them on one, three or four streams (QSF_BATCH_STREAMS) gives the same bits: states and
    # record rows of every solve, under hipGraph replay
    from qsf_amd import hip
    kw2 = dict(potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
               absorber=T.ABS_CAP, nCAP=8)
    kw1 = dict(potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1.0, absorber=T.ABS_CAP, nCAP=30)
    specs = [(2, 128, kw2, 0.3), (2, 64, kw2, 0.3), (2, 96, kw2, 0.3), (1, 256, kw1, 0.1),
             (2, 80, kw2, 0.3), (2, 128, kw2, 0.3), (2, 48, kw2, 0.3), (1, 384, kw1, 0.1)]
    descs = [T.grid_desc(dim, n, dx, precision=prec, **kw) for dim, n, kw, dx in specs]
    fields = [T.field_desc(T.AXIS_X if dim == 1 else T.AXIS_XY, 0.05 + 0.01 * i, 2.0, 1 + (i % 2),
                           phase_in_pi=0.25 * i) for i, (dim, n, kw, dx) in enumerate(specs)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=[(T.BORDER_BOX, 12)])
    psis = [crand(n ** dim) * 0.01 for dim, n, kw, dx in specs]
    results = {}
    for streams in (1, 3, 4):
        print(f"[case] batch on {streams} stream(s), prec={prec}", flush=True)
        monkeypatch.setenv("QSF_BATCH_STREAMS", str(streams))
        b = hip.Batch(ctx, descs, fields, pdesc)
        for s, p in enumerate(psis):
            b.set_psi(s, p)
        b.run_all(use_graph=True)
        results[streams] = [(b.psi(s), b.records(s)) for s in range(len(specs))]
        b.close()
    for streams in (3, 4):
        for s in range(len(specs)):
            assert np.array_equal(results[streams][s][0], results[1][s][0])
            assert np.array_equal(results[streams][s][1], results[1][s][1], equal_nan=True)


def test_cos18_edge_node_stays_finite(ctx):
    """cos^(1/8) with a width whose π/2·node/nCAP rounds past π/2 at the edge node (nCAP=13):
    the mask there is exactly 0 on both sides and the propagation stays finite."""
    n, dx = 128, 0.2
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_COS18, nCAP=13, absorber_tau=0.02)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    psi0 = (np.exp(-0.5 * (X * X + Y * Y)) + 0j).ravel()
    f = [T.field_desc(T.AXIS_XY, 0.1, 0.5, 1)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.02, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    dev, ref, pd, po, rd, ro = run_both(ctx, d, pdesc, f, psi0, 20)
    assert np.all(np.isfinite(rd[:, T.REC_NORM])) and np.all(np.isfinite(dev[0].psi()))
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    # the mask of the main region is 0 on the edge rows and columns
    md = T.grid_desc(2, n, dx, absorber=T.ABS_COS18, nCAP=13, absorber_tau=0.02, multigrid=1)
    from qsf_amd import hip
    gs = hip.GridSet(ctx, md)
    gs[0].set_psi(np.ones(n * n, np.complex128))
    gs[0].mask_region()
    m = gs[0].psi().real.reshape(n, n)
    assert np.all(m[0, :] == 0.0) and np.all(m[:, -1] == 0.0) and np.isfinite(m).all()
    gs.close()


# ------------------------------------------------------------------------------------ 5*2^k lengths
@pytest.mark.parametrize("prec", [T.FP64, T.FP32_MIXED])
@pytest.mark.parametrize("multigrid", [0, 1])
def test_propagation_radix5_lengths(ctx, prec, multigrid):
    # 160 = 5 * 32: a radix-5 first stage, then radix 4 and 2 (the fused passes, and for the
    # multigrid set the fused region transfers, at that length)
    n, dx = 160, 0.25
    d = T.grid_desc(2, n, dx, precision=prec, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5,
                    Esoft=0.33, absorber=T.ABS_CAP, nCAP=16, multigrid=multigrid)
    d64 = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                      absorber=T.ABS_CAP, nCAP=16, multigrid=multigrid)
    psi0 = crand(n * n) * 0.01
    if prec == T.FP32_MIXED:
        psi0 = psi0.astype(np.complex64).astype(np.complex128)
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.02, gauge=T.GAUGE_VELOCITY,
                        outputs=T.OUT_NORM | T.OUT_FLUX, borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    from qsf_amd import hip
    print(f"[case] radix-5 n={n} prec={prec} multigrid={multigrid}", flush=True)
    dev = hip.GridSet(ctx, d)
    ref = O.GridSet(d64)
    for r in range(len(ref)):
        z = psi0 if r == 0 else np.zeros_like(psi0)
        dev[r].set_psi(z)
        ref[r].set_psi(z)
    rd = hip.Propagator(dev, pdesc, f).run(8)
    ro = O.Propagator(ref, pdesc, f).run(8)
    tol = FP64_TOL if prec == T.FP64 else 1e-4
    for r in range(len(ref)):
        assert rel(dev[r].psi(), ref[r].psi()) < tol, f"region {r}"
    assert np.allclose(rd[:, 2], ro[:, 2], rtol=1e-11 if prec == T.FP64 else 1e-5)
