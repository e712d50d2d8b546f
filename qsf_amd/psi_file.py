"""`.psi` wavefunction files: QSF's on-disk format for states (ground-state handoff from imaginary
to real time, snapshots), written from and read into device grids.

Layout (wf/grid.h:1025-1071 writes it; mathematica/QSF/Kernel/wf.wl:204-222 reads it):
  u8 dim, u8 rep (REP::X = 1, REP::P = 2), u8 complex, i64 downscale,
  i64 n[dim], f64 dx[dim], u8 bounded[dim],
  then row-major data: complex128 per node (complex files, downscale 1), or float64 |psi|^2
  summed over downscale^dim blocks (wf/grid.h:955-992; sums, not means — the reader's WFNorm,
  wf.wl:229, then comes out downscale^dim too large, as it does for the reference's files).
P-representation files hold the centred momentum grid: psi * (-1)^(i0+i1+...) transformed
forward (wf/grid.h:1073-1107, phaseShiftTrick) and scaled by dV / (2 pi)^(dim/2) (|.|^2 by its
square). File names get "_scale<k>" when downscaled and "_repX" / "_repP", then ".psi"
(wf/grid.h:1040-1050).
"""
import os
import struct

import numpy as np

from . import qsf_types as T

EXT = ".psi"


def file_name(path, rep=T.REP_X, downscale=1, evo_backup=False):
    """The name _save gives `path` (wf/grid.h:1040-1050)."""
    path = str(path)
    if not evo_backup:
        if downscale > 1:
            path += f"_scale{int(downscale)}"
        path += "_repX" if rep == T.REP_X else "_repP"
    return path + EXT


def header_bytes(dim, rep, is_complex, downscale, n, dx, bounded):
    b = struct.pack("<BBBq", dim, rep, 1 if is_complex else 0, int(downscale))
    b += struct.pack(f"<{dim}q", *[int(v) for v in n[:dim]])
    b += struct.pack(f"<{dim}d", *[float(v) for v in dx[:dim]])
    b += struct.pack(f"<{dim}B", *[1 if v else 0 for v in bounded[:dim]])
    return b


def coarse_density(psi, n, downscale):
    """|psi|^2 summed over downscale^dim blocks, row-major (wf/grid.h:955-992)."""
    a = np.abs(np.asarray(psi).reshape(n)) ** 2
    k = int(downscale)
    if k == 1:
        return a
    for ax in range(a.ndim):
        shape = list(a.shape)
        shape[ax:ax + 1] = [shape[ax] // k, k]
        a = a.reshape(shape).sum(axis=ax + 1)
    return a


def write(path, psi, n, dx, bounded=None, rep=T.REP_X, is_complex=True, downscale=1, evo_backup=False):
    """Write an X- or (already centred and scaled) P-representation array; returns the file name."""
    dim = len(n)
    bounded = [True] * dim if bounded is None else list(bounded)
    if downscale > 1:
        is_complex = False
    name = file_name(path, rep, downscale, evo_backup)
    with open(name, "wb") as f:
        f.write(header_bytes(dim, rep, is_complex, downscale, n, dx, bounded))
        if is_complex:
            f.write(np.ascontiguousarray(psi, dtype=np.complex128).tobytes())
        else:
            f.write(np.ascontiguousarray(coarse_density(psi, n, downscale), dtype=np.float64).tobytes())
    return name


def read(name):
    """Header (dict) and data (array shaped n/downscale) of a `.psi` file."""
    with open(name, "rb") as f:
        dim, rep, is_complex, downscale = struct.unpack("<BBBq", f.read(11))
        n = list(struct.unpack(f"<{dim}q", f.read(8 * dim)))
        dx = list(struct.unpack(f"<{dim}d", f.read(8 * dim)))
        bounded = [bool(v) for v in struct.unpack(f"<{dim}B", f.read(dim))]
        data = np.frombuffer(f.read(), dtype=np.complex128 if is_complex else np.float64)
    hd = dict(dim=dim, rep=rep, isComplex=bool(is_complex), downscale=downscale, ns=n, dxs=dx,
              bounded=bounded)
    return hd, data.reshape([v // downscale for v in n])


def centred_shift(n):
    """(-1)^(i0 + i1 + ...) over the grid, row-major (phaseShiftTrick, wf/grid.h:1073-1087)."""
    idx = np.indices(n).sum(axis=0)
    return np.where(idx % 2 == 0, 1.0, -1.0).ravel()


def save_grid(gset, path, region=0, rep=T.REP_X, is_complex=True, downscale=1, bounded=None):
    """`save` of a device grid held in X rep (wf/grid.h:1108-1114). X: the state as is. P: the
    device transforms the phase-shifted state forward; the file then holds the reference's P
    memory order (FFTW's transposed layout for D > 1, which `coarse` walks with X offsets,
    wf/grid.h:960-968) times dV / (2 pi)^(dim/2). The grid is left as it was."""
    d = gset.desc
    grid = gset[region]
    dim = d.dim
    n = [int(d.n[i]) for i in range(dim)]
    dx = [float(d.dx[i]) for i in range(dim)]
    if grid.rep != T.REP_X:
        raise ValueError("save_grid expects the grid in the X representation")
    psi = grid.psi()
    if rep == T.REP_P:
        grid.set_psi(psi * centred_shift(n))
        grid.fourier(T.REP_P)
        data = grid.psi() * (float(np.prod(dx)) / (2 * np.pi) ** (dim / 2))
        grid.set_psi(psi)                  # back to the state as it was
    else:
        data = psi
    return write(path, data, n, dx, bounded, rep, is_complex, downscale)


def load_grid(gset, name, region=0):
    """Load a complex X-representation `.psi` file into a device grid of the same shape."""
    hd, data = read(name)
    d = gset.desc
    if hd["rep"] != T.REP_X or not hd["isComplex"] or hd["downscale"] != 1:
        raise ValueError("only complex X-representation files load into a grid")
    if hd["ns"] != [int(d.n[i]) for i in range(d.dim)]:
        raise ValueError(f"grid shape {hd['ns']} does not match the file")
    gset[region].set_psi(np.asarray(data).ravel())
    return hd


# ---- evolution backups (wf/grid.h:1115-1132, multigrid.h:370; resume: propagator.h:425-436) ----
BACKUP_INFO = "backup.info"


def backup_names(gset, directory="."):
    """backup.psi for a single grid; backup.psi<r> per region of a multigrid set."""
    base = os.path.join(str(directory), "backup")
    if len(gset.grids) == 1:
        return [base + EXT]
    return [base + EXT + str(r) for r in range(len(gset.grids))]


def backup(gset, step, directory="."):
    """_backup(step): backup.info holds the step as one i64; each region's X state goes to an
    evolution-backup file (same header, no rep suffix in the name). Taken between steps, the
    grid entry points hand back the step-boundary state."""
    with open(os.path.join(str(directory), BACKUP_INFO), "wb") as f:
        f.write(struct.pack("<q", int(step)))
    d = gset.desc
    n = [int(d.n[i]) for i in range(d.dim)]
    dx = [float(d.dx[i]) for i in range(d.dim)]
    names = backup_names(gset, directory)
    for r, name in enumerate(names):
        # the region's bounded flags (mcomm.bounded, wf/grid.h:1035); multigrid regions run
        # NO, X, Y, XY (mpi_logic.h:256-259), so region r frees the axes of mask r
        free = r if len(names) > 1 else 0
        bounded = [not (free >> i) & 1 for i in range(d.dim)]
        with open(name, "wb") as f:
            f.write(header_bytes(d.dim, T.REP_X, True, 1, n, dx, bounded))
            f.write(np.ascontiguousarray(gset[r].psi(), dtype=np.complex128).tobytes())
    return names


def restore(gset, prop, directory="."):
    """run(restore=true): read backup.info, fast-forward the propagator to that step
    (makeFakeStep: coupling and clock only), then load every region's backup. Returns the step."""
    with open(os.path.join(str(directory), BACKUP_INFO), "rb") as f:
        (step,) = struct.unpack("<q", f.read(8))
    prop.fast_forward(step)
    for r, name in enumerate(backup_names(gset, directory)):
        _, data = read(name)
        gset[r].set_psi(np.asarray(data).ravel())
    return step


def remove_backups(gset, directory="."):
    """remove_backups + the backup.info removal at the end of run (propagator.h:456-461)."""
    for name in backup_names(gset, directory) + [os.path.join(str(directory), BACKUP_INFO)]:
        if os.path.exists(name):
            os.remove(name)


def orthogonalize_with(gset, name, region=0):
    """orthogonalizeWith(input_path) (wf/grid.h:909-946): project the state stored in a complex
    X-representation `.psi` file out of a device grid; returns <phi|psi> dV. The data start after
    the 11 + 17*dim byte header that `save` writes; the reference reads from 8 + 17*dim bytes
    (its offset counts no i64 downscale and a 4-byte int), three bytes early."""
    hd, data = read(name)
    if hd["rep"] != T.REP_X or not hd["isComplex"] or hd["downscale"] != 1:
        raise ValueError("orthogonalizeWith needs a complex X-representation file")
    return gset[region].orthogonalize(np.asarray(data).ravel())
