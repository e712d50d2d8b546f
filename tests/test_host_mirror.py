"""The C++ host mirror (qsf_amd/host/qsf/*.h): reference type names over the C ABI.

CPU: the example program compiles and links against libqsf_hip.so with plain g++.
GPU: its records equal the oracle's for the same reference-style configuration.
"""
import os
import subprocess

import numpy as np
import pytest

from qsf_amd import qsf_types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "qsf_amd", "host")
EXE = os.path.join(HOST, "examples", "he2d")


def test_host_mirror_builds_with_gxx():
    if not os.path.exists(os.path.join(ROOT, "qsf_amd", "libqsf_hip.so")):
        pytest.skip("libqsf_hip.so not built")
    subprocess.run(["make", "-s", "-C", HOST], check=True)
    assert os.access(EXE, os.X_OK)


def he2d_oracle(n, steps, multigrid, with_state=False):
    """The configuration he2d.cpp writes with reference type names, run on the oracle."""
    from oracle import oracle as O
    dx = 0.3
    d = T.grid_desc(2, n, dx, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1, Nsoft=0.5, Esoft=0.33,
                    absorber=T.ABS_CAP, nCAP=n // 8, multigrid=multigrid)
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    psi0 = (np.exp(-0.5 * ((X - 1) ** 2 + (Y + 0.5) ** 2)) * np.exp(0.3j * X)).ravel()
    f = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]
    p = T.prop_desc(T.MODE_RE, 2, 0.02, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_FLUX,
                    borders=[(b, 0) for b in T.ZOA_FLUX_2D])
    gs = O.GridSet(d)
    for r in range(len(gs)):
        gs[r].set_psi(psi0 if r == 0 else np.zeros_like(psi0))
    prop = O.Propagator(gs, p, f)
    rec = prop.run(steps)
    return (rec, gs[0].psi()) if with_state else rec


@pytest.mark.gpu
@pytest.mark.parametrize("multigrid", [0, 1])
def test_host_mirror_program_matches_oracle(multigrid):
    n, steps = 64, 10
    print(f"[case] he2d n={n} steps={steps} multigrid={multigrid}", flush=True)
    import tempfile
    from qsf_amd import psi_file as PF
    save = os.path.join(tempfile.mkdtemp(), "he2d")
    out = subprocess.run([EXE, str(n), str(steps), str(multigrid), save], check=True, capture_output=True,
                         text=True, timeout=120).stdout
    rec = np.array([[float(v) for v in line.split(",")] for line in out.strip().splitlines()])
    ref, psi_ref = he2d_oracle(n, steps, multigrid, with_state=True)
    assert rec.shape == ref.shape
    # the state the program saved as a .psi file (main region)
    hd, psi = PF.read(save + "_repX.psi")
    assert hd["dim"] == 2 and hd["ns"] == [n, n] and hd["isComplex"]
    assert np.linalg.norm(psi.ravel() - psi_ref) / np.linalg.norm(psi_ref) < 1e-10
    np.testing.assert_allclose(rec[:, :2], ref[:, :2], rtol=0, atol=1e-12)
    np.testing.assert_allclose(rec[:, 2:4], ref[:, 2:4], rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(rec[:, 7:], ref[:, 7:], rtol=1e-8, atol=1e-12 * np.abs(ref[:, 7:]).max())
