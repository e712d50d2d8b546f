"""`.dat` observable streams (qsf_amd/dat_file.py) against the line bookkeeping of
BufferedOutputs (wf/buffer.h:64-97, 281-397), worked through by hand for small cadences."""
import numpy as np
import pytest

from qsf_amd import dat_file as DF
from qsf_amd import qsf_types as T


def fake_records(last_step, width=T.REC_FLUX0 + 7):
    rec = np.zeros((last_step + 1, width))
    rec[:, T.REC_STEP] = np.arange(last_step + 1)
    rec[:, T.REC_TIME] = 0.05 * np.arange(last_step + 1)
    rec[:, 2:] = np.arange(last_step + 1)[:, None] * 10.0 + np.arange(2, width)[None, :]
    return rec


@pytest.mark.parametrize("last,want", [
    # comp_interval 2, log_interval 4 (two buffer lines). AT_START stores step 0 on the last line
    # and writes it; steps 2 and 4 advance/flush; the AT_END step is written with the lines
    # stored since the last flush.
    (8, [0, 2, 4, 6, 8]),
    (7, [0, 2, 4, 6, 7]),
    (9, [0, 2, 4, 6, 8, 9]),
])
def test_line_bookkeeping_by_hand(last, want):
    rec = fake_records(last)
    rows = DF.lines(rec, comp_interval=2, log_interval=4, finished=True)
    assert [int(rec[r, T.REC_STEP]) for r in rows] == want


def test_every_step_and_unfinished_runs():
    rec = fake_records(6)
    assert DF.lines(rec, 1, 1) == list(range(7))
    # comp 1, log 3: steps 3 and 6 flush three lines each; an unfinished run (no AT_END) leaves
    # the lines after the last flush unwritten
    assert DF.lines(rec, 1, 3) == list(range(7))
    assert DF.lines(rec[:6], 1, 3, finished=False) == [0, 1, 2, 3]
    with pytest.raises(ValueError):
        DF.lines(rec, 2, 3)


def test_layout_plain_columns_before_reduced():
    plain, red = DF.layout(["ENERGY_DIFFERENCE", "AVG<Identity>", "VALUE<Step, Time>", "ZOA_FLUX_2D",
                            "ENERGY_TOTAL"])
    assert plain == [T.REC_EDIFF, T.REC_STEP, T.REC_TIME]
    assert red == [T.REC_NORM] + [T.REC_FLUX0 + i for i in range(7)] + [T.REC_ETOT]
    plain, red = DF.layout([("FLUX", 2), "AVG<PotentialEnergy>", ("FLUX", 1)])
    assert plain == [] and red == [T.REC_FLUX0, T.REC_FLUX0 + 1, T.REC_EPOT, T.REC_FLUX0 + 2]
    with pytest.raises(ValueError):
        DF.layout(["AVG<Momentum>"])


def test_binary_and_text_files(tmp_path):
    rec = fake_records(8)
    comps = ["VALUE<Time>", "AVG<Identity>", "ZOA_FLUX_2D"]
    f = DF.write(tmp_path / "run", rec, comps, comp_interval=2, log_interval=4, binary=True)
    assert f.endswith("run.dat")
    got = DF.read(f, 9, binary=True)
    want = rec[[0, 2, 4, 6, 8]][:, [T.REC_TIME, T.REC_NORM] + list(range(T.REC_FLUX0, T.REC_FLUX0 + 7))]
    assert np.array_equal(got, want)
    import os
    assert os.path.getsize(f) == 5 * 9 * 8          # raw doubles, no header
    f = DF.write(tmp_path / "txt", rec, comps, comp_interval=2, log_interval=4, binary=False)
    text = open(f).read().splitlines()
    assert len(text) == 5 and text[1].startswith("|%13.5g " % 0.1)
    assert np.allclose(DF.read(f, 9, binary=False), want, rtol=1e-4)


@pytest.mark.gpu
def test_dat_from_device_records(tmp_path):
    from qsf_amd import hip
    n, dx = 512, 0.1
    d = T.grid_desc(1, n, dx, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1, Nsoft=1,
                    absorber=T.ABS_CAP, nCAP=51)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    f = [T.field_desc(T.AXIS_X, 0.0534, 0.057, 1)]
    pdesc = T.prop_desc(T.MODE_RE, 1, 0.5, gauge=T.GAUGE_LENGTH,
                        outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_EDIFF | T.OUT_FLUX,
                        borders=[(T.BORDER_BOX, 150)])
    print("[case] dat from 1D device records", flush=True)
    ctx = hip.Context(0)
    gs = hip.GridSet(ctx, d)
    gs[0].set_psi((np.exp(-x * x) + 0j) / np.sqrt(np.sqrt(np.pi / 2)))
    rec = hip.Propagator(gs, pdesc, f).run()
    comps = ["VALUE<Step, Time>", "AVG<Identity>", "ENERGY_TOTAL", "ENERGY_DIFFERENCE", ("FLUX", 1)]
    name = DF.write(tmp_path / "he1d", rec, comps, comp_interval=5, log_interval=10)
    got = DF.read(name, 6)
    steps = got[:, 0].astype(int)
    # every multiple of 5, then the AT_END step
    assert list(steps[:-1]) == list(range(0, int(rec[-2, T.REC_STEP]) + 1, 5))[:len(steps) - 1]
    assert steps[-1] == int(rec[-1, T.REC_STEP])
    for row in got:
        r = rec[int(row[0])] if rec[int(row[0]), T.REC_STEP] == row[0] else None
        assert r is not None
        assert np.array_equal(row, r[[T.REC_STEP, T.REC_TIME, T.REC_EDIFF, T.REC_NORM, T.REC_ETOT,
                                      T.REC_FLUX0]])
    gs.close()
    ctx.close()
