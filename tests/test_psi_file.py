"""`.psi` files (qsf_amd/psi_file.py): the byte layout the reference writes and its Mathematica
reader reads (wf/grid.h:1025-1071, mathematica/QSF/Kernel/wf.wl:204-222), round trips, block
densities, and on the GPU the centred P representation."""
import os
import struct

import numpy as np
import pytest

from qsf_amd import psi_file as PF
from qsf_amd import qsf_types as T


def test_file_names():
    assert PF.file_name("out/gs") == "out/gs_repX.psi"
    assert PF.file_name("gs", T.REP_P) == "gs_repP.psi"
    assert PF.file_name("gs", T.REP_X, downscale=4) == "gs_scale4_repX.psi"
    assert PF.file_name("backup", evo_backup=True) == "backup.psi"


def test_header_layout_is_the_readers(tmp_path):
    n, dx = [6, 4], [0.2, 0.3]
    psi = (np.arange(24) + 1j * np.arange(24)[::-1]).astype(np.complex128)
    name = PF.write(str(tmp_path / "a"), psi, n, dx, bounded=[True, False])
    raw = open(name, "rb").read()
    # UnsignedInteger8 dim, UnsignedInteger8 rep, Byte isComplex, Integer64 downscale,
    # Integer64 ns[dim], Real64 dxs[dim], Byte bounded[dim]   (wf.wl:204-212)
    assert raw[0] == 2 and raw[1] == T.REP_X and raw[2] == 1
    assert struct.unpack("<q", raw[3:11])[0] == 1
    assert list(struct.unpack("<2q", raw[11:27])) == n
    assert list(struct.unpack("<2d", raw[27:43])) == dx
    assert list(raw[43:45]) == [1, 0]
    assert len(raw) == 45 + 24 * 16
    assert np.array_equal(np.frombuffer(raw[45:], np.complex128), psi)
    hd, data = PF.read(name)
    assert hd["ns"] == n and hd["dxs"] == dx and hd["bounded"] == [True, False] and hd["isComplex"]
    assert np.array_equal(data.ravel(), psi)


def test_downscaled_density_blocks(tmp_path):
    rng = np.random.default_rng(3)
    n = [8, 6]
    psi = rng.normal(size=48) + 1j * rng.normal(size=48)
    name = PF.write(str(tmp_path / "b"), psi, n, [0.1, 0.1], downscale=2)
    hd, data = PF.read(name)
    assert not hd["isComplex"] and hd["downscale"] == 2 and data.shape == (4, 3)
    a = np.abs(psi.reshape(8, 6)) ** 2
    assert np.allclose(data[1, 2], a[2:4, 4:6].sum(), rtol=1e-15)
    # blocks are sums, not means (downscale_div in wf/grid.h:962 is computed but unused), so
    # the data total is the total |psi|^2 of the grid
    assert np.isclose(data.sum(), a.sum(), rtol=1e-14)


@pytest.mark.gpu
def test_device_grid_save_load_and_centred_p(tmp_path):
    from qsf_amd import hip
    ctx = hip.Context(0)
    n, dx = 32, 0.25
    d = T.grid_desc(2, n, dx)
    gs = hip.GridSet(ctx, d)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    psi = (np.exp(-(X - 0.5) ** 2 - 0.5 * Y ** 2) * np.exp(1j * 0.8 * X)).ravel()
    gs[0].set_psi(psi)
    nx = PF.save_grid(gs, str(tmp_path / "g"))
    npn = PF.save_grid(gs, str(tmp_path / "g"), rep=T.REP_P)
    assert np.allclose(gs[0].psi(), psi, rtol=0, atol=1e-15)      # grid left as it was
    hd, data = PF.read(nx)
    assert np.array_equal(data.ravel(), psi)
    hp, pdata = PF.read(npn)
    assert hp["rep"] == T.REP_P
    # centred momentum grid: numpy's forward FFT of the phase-shifted state, scaled, stored in
    # the reference's transposed P order (axes swapped for D = 2)
    want = np.fft.fftn(psi.reshape(n, n) * PF.centred_shift([n, n]).reshape(n, n)) * dx * dx / (2 * np.pi)
    assert np.allclose(pdata, want.T, rtol=1e-12, atol=1e-14 * np.abs(want).max())
    # the centred momentum density integrates to the X norm (Parseval, dp = 2 pi / (n dx))
    dp = 2 * np.pi / (n * dx)
    assert np.isclose(np.sum(np.abs(pdata) ** 2) * dp * dp, np.sum(np.abs(psi) ** 2) * dx * dx, rtol=1e-12)
    gs2 = hip.GridSet(ctx, d)
    PF.load_grid(gs2, nx)
    assert np.array_equal(gs2[0].psi(), psi)
    gs2.close()
    gs.close()
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("multigrid", [0, 1])
def test_backup_and_resume_equal_uninterrupted_run(tmp_path, multigrid):
    # _backup at a step boundary, then run(restore=true) on fresh objects: makeFakeStep to the
    # saved step, load every region, continue. Strang keeps one chain, so the resumed run is the
    # uninterrupted one (propagator.h:425-449).
    import struct
    from qsf_amd import hip
    n, dx, k1, k2 = 64, 0.3, 6, 12
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_CAP, nCAP=8, multigrid=multigrid)
    rng = np.random.default_rng(11)
    psi0 = (rng.normal(size=n * n) + 1j * rng.normal(size=n * n)) * 0.01
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.02, gauge=T.GAUGE_VELOCITY,
                        outputs=T.OUT_NORM | T.OUT_FLUX, borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    print(f"[case] backup/resume multigrid={multigrid}", flush=True)
    ctx = hip.Context(0)

    def loaded():
        gs = hip.GridSet(ctx, d)
        for r in range(len(gs.grids)):
            gs[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
        return gs

    ref = loaded()
    pr = hip.Propagator(ref, pdesc, f)
    pr.start()
    for _ in range(k2):
        pr.step()
    rows_ref = pr.records()

    first = loaded()
    p1 = hip.Propagator(first, pdesc, f)
    p1.start()
    for _ in range(k1):
        p1.step()
    names = PF.backup(first, p1.step_count, tmp_path)
    assert struct.unpack("<q", open(tmp_path / "backup.info", "rb").read())[0] == k1
    assert [os.path.basename(x) for x in names] == (["backup.psi"] if not multigrid else
                                                    [f"backup.psi{r}" for r in range(4)])
    hd, _ = PF.read(names[-1])
    assert hd["bounded"] == ([True, True] if not multigrid else [False, False])
    p1.close()
    first.close()

    second = hip.GridSet(ctx, d)
    p2 = hip.Propagator(second, pdesc, f)
    assert PF.restore(second, p2, tmp_path) == k1
    assert p2.step_count == k1 and abs(p2.time - k1 * 0.02) < 1e-15
    for _ in range(k2 - k1):
        p2.step()
    rows = p2.records()
    assert rows.shape[0] == k2 - k1          # no AT_START row on resume
    np.testing.assert_allclose(rows, rows_ref[k1 + 1:], rtol=1e-12,
                               atol=1e-14 * np.nanmax(np.abs(rows_ref)))
    for r in range(len(ref.grids)):
        a, b = second[r].psi(), ref[r].psi()
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    PF.remove_backups(second, tmp_path)
    assert not os.listdir(tmp_path)
    for o in (p2, pr):
        o.close()
    second.close()
    ref.close()
    ctx.close()


@pytest.mark.gpu
def test_fast_forward_clock_follows_fake_steps():
    # makeFakeStep advances the clock by every chain's P-substep sum with no chain restore:
    # Strang dt per step, MultiProductSplit order 2 (chains 1 + 1) 2 dt per step
    from qsf_amd import hip
    d = T.grid_desc(1, 64, 0.2, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1)
    f = [T.field_desc(T.AXIS_X, 0.05, 0.057, 2)]
    ctx = hip.Context(0)
    for order, per_step in ((1, 0.05), (2, 0.1)):
        gs = hip.GridSet(ctx, d)
        p = hip.Propagator(gs, T.prop_desc(T.MODE_RE, order, 0.05, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM), f)
        p.fast_forward(5)
        assert p.step_count == 5
        assert abs(p.time - 5 * per_step) < 1e-14
        with pytest.raises(Exception):
            p.fast_forward(7)                 # only on a fresh propagator
        p.close()
        gs.close()
    ctx.close()
