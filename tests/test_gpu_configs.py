"""The workload configurations in their own form, on the device against the FP64 oracle.

* 1D single electron, N = 4096, dx = 0.1: imaginary-time ground state with
  OPERATION<Normalize> from gaussian(0, 1) (wf/preset.h:5-9), then real time with Strang, length
  gauge, 800 nm sin^2 pulse E0 = 0.0534, cos^(1/8) over the outer 410 nodes and FLUX<BOX<1500>>
  (x = +-149.95) plus norm and energies every step.
* 2D two-electron model with everything of its workload at 96^2 and 128^2: Eberly interaction,
  velocity gauge, MultiProductSplit order 2, cos^(1/8) over the outer 10%, multigrid regions
  NO/X/Y/XY with their masks and transfers, FLUX<ZOA_FLUX_2D>.
* Batched FP32-mixed solves of uneven sizes including 3*2^k (64, 96, 128, 192) replayed as a
  hipGraph: each solve against the oracle in FP64 and against the same solve advanced alone in
  FP32-mixed.

Bars: FP64 L2(psi) <= 1e-10 relative, records to 1e-10 (norm, energies) and 1e-8 of the
largest flux; FP32-mixed batch solves within 1e-4 (psi) and 1e-6 (norm) of the oracle and
1e-6 (psi) of the single FP32 solve.
"""
import numpy as np
import pytest

from oracle import oracle as O
from qsf_amd import qsf_types as T

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    from qsf_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def c1_desc():
    return T.grid_desc(1, 4096, 0.1, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1,
                       absorber=T.ABS_COS18, nCAP=410, absorber_tau=0.05)


def test_c1_ground_state_then_pulse(ctx):
    from qsf_amd import hip
    d = c1_desc()
    n, dx = 4096, 0.1
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    psi0 = np.exp(-x * x) + 0j              # gaussian(0, 1) of wf/preset.h:5-9 (unnormalised)
    # stops when |dE| <= 1e-11 (about step 370; dE shrinks ~4% per step there, so a 1e-15
    # difference of the energies cannot move the stopping step) or after 2000 steps
    im = T.prop_desc(T.MODE_IM, 1, 0.05, max_steps=2000, state_accuracy=1e-11,
                     outputs=T.OUT_NORMALIZE | T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_EDIFF)
    print("[case] C1 imaginary time (at most 2000 steps)", flush=True)
    dev, ref = hip.GridSet(ctx, d), O.GridSet(d)
    dev[0].set_psi(psi0)
    ref[0].set_psi(psi0)
    rd = hip.Propagator(dev, im, []).run()
    ro = O.Propagator(ref, im, []).run()
    assert rd.shape == ro.shape
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    np.testing.assert_allclose(rd[:, 2:7], ro[:, 2:7], rtol=1e-10, atol=1e-13)
    assert abs(rd[-1, T.REC_ETOT] - (-0.6698)) < 2e-3

    print("[case] C1 real time 600 steps, BOX<1500>", flush=True)
    ground = ref[0].psi()
    f = [T.field_desc(T.AXIS_X, 0.0534, 0.057, 4)]
    re = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH,
                     outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_FLUX,
                     borders=[(T.BORDER_BOX, 1500)])
    dev, ref = hip.GridSet(ctx, d), O.GridSet(d)
    dev[0].set_psi(ground)
    ref[0].set_psi(ground)
    pd = hip.Propagator(dev, re, f)
    po = O.Propagator(ref, re, f)
    assert pd.max_steps == po.max_steps == int(2 * np.pi * 4 / 0.057 / 0.05 + 1)
    rd, ro = pd.run(600), po.run(600)
    assert rel(dev[0].psi(), ref[0].psi()) < FP64_TOL
    np.testing.assert_allclose(rd[:, :2], ro[:, :2], rtol=0, atol=1e-12)
    np.testing.assert_allclose(rd[:, 2:6], ro[:, 2:6], rtol=1e-10, atol=1e-13)
    fl = ro[:, T.REC_FLUX0]
    assert np.abs(rd[:, T.REC_FLUX0] - fl).max() <= 1e-8 * max(np.abs(fl).max(), 1e-300)


@pytest.mark.parametrize("n", [96, 128])
def test_c2_form_multigrid_cos18_order2(ctx, n):
    from qsf_amd import hip
    d = T.grid_desc(2, n, 0.2, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_COS18, nCAP=int(round(0.1 * n)), absorber_tau=0.02, multigrid=1)
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    pdesc = T.prop_desc(T.MODE_RE, 2, 0.02, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    rng = np.random.default_rng(n)
    psi0 = (rng.normal(size=n * n) + 1j * rng.normal(size=n * n)) * 0.01
    print(f"[case] C2 form n={n}", flush=True)
    dev, ref = hip.GridSet(ctx, d), O.GridSet(d)
    for r in range(4):
        dev[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
        ref[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
    rd = hip.Propagator(dev, pdesc, f).run(20)
    ro = O.Propagator(ref, pdesc, f).run(20)
    for r in range(4):
        assert rel(dev[r].psi(), ref[r].psi()) < FP64_TOL, f"region {r}"
    np.testing.assert_allclose(rd[:, T.REC_NORM], ro[:, T.REC_NORM], rtol=1e-11)
    fl = ro[:, T.REC_FLUX0:]
    assert np.abs(rd[:, T.REC_FLUX0:] - fl).max() <= 1e-8 * np.abs(fl).max()
    # the regions received something: the transfers ran
    assert all(np.linalg.norm(ref[r].psi()) > 0 for r in range(1, 4))


def test_fp32_batch_uneven_sizes_graph(ctx):
    from qsf_amd import hip
    base = dict(potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                absorber=T.ABS_COS18, absorber_tau=0.02)
    sizes = [64, 96, 128, 192, 96, 64]
    descs32 = [T.grid_desc(2, n, 0.3, precision=T.FP32_MIXED, nCAP=n // 10, **base) for n in sizes]
    descs64 = [T.grid_desc(2, n, 0.3, nCAP=n // 10, **base) for n in sizes]
    fields = [T.field_desc(T.AXIS_XY, 0.05 + 0.02 * i, 2.0, 1 + (i % 2), phase_in_pi=0.3 * i)
              for i in range(len(sizes))]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    rng = np.random.default_rng(3)
    psis = [((rng.normal(size=n * n) + 1j * rng.normal(size=n * n)) * 0.01) for n in sizes]
    print("[case] FP32 batch 64/96/128/192, graph replay", flush=True)
    b = hip.Batch(ctx, descs32, fields, pdesc)
    for s, p in enumerate(psis):
        b.set_psi(s, p)
    b.run_all(use_graph=True)
    for s, n in enumerate(sizes):
        print(f"[case] FP32 batch solve {s} n={n}", flush=True)
        ref = O.GridSet(descs64[s])
        ref[0].set_psi(psis[s])
        ro = O.Propagator(ref, pdesc, [fields[s]]).run()
        rb = b.records(s)
        assert rb.shape[0] == ro.shape[0]
        assert rel(b.psi(s), ref[0].psi()) < 1e-4
        np.testing.assert_allclose(rb[:, 2], ro[:, T.REC_NORM], rtol=1e-6)
        fl = ro[:, T.REC_FLUX0:]
        assert np.abs(rb[:, 3:] - fl).max() <= 1e-4 * np.abs(fl).max()
        # the same solve alone in FP32-mixed: same kernels, per-solve tables
        gs = hip.GridSet(ctx, descs32[s])
        gs[0].set_psi(psis[s])
        rs = hip.Propagator(gs, pdesc, [fields[s]]).run()
        assert rel(b.psi(s), gs[0].psi()) < 1e-6
        np.testing.assert_allclose(rb[:, 2], rs[:, T.REC_NORM], rtol=1e-7)
        gs.close()
    b.close()
