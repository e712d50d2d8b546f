"""`.dat` observable streams: the file `BufferedOutputs` writes beside a run (wf/buffer.h), built
from the device's record rows (one FP64 row per computeEach: step, time, norm, epot, ekin, etot,
ediff, flux...).

What reaches the file (wf/buffer.h:64-97 init, 281-345 logAll, 379-397 logOrPass):
* Each computation lands in one of two line buffers: the reduce buffer (AVG<...>, FLUX<...>,
  SUM of reduced terms such as ENERGY_TOTAL) or the plain buffer (VALUE<...>, CHANGE<...> such as
  ENERGY_DIFFERENCE) (computations.h:18-104, propagator.h:26-27). A written line is the plain
  columns, then the reduce columns, each in declaration order.
* Buffers hold log_interval / comp_interval lines and start on the last line. computeEach runs
  after every step and stores into the current line; the line advances only on steps divisible
  by comp_interval, so a line keeps the values of the step that advanced it.
* On steps divisible by log_interval, and at AT_END, lines are written and the position returns
  to line 0: at AT_START the last line only, during the run every line, at AT_END the lines up to
  the current one.
* Binary files hold raw doubles with no header (the header writer is commented out,
  wf/buffer.h:259-279, 304); text files print each value as "|%13.5g " and end a line with "\\n"
  (basics/constants.h:15,26). The name is <name>.dat (basics/io.h:33).
"""
import numpy as np

from . import qsf_types as T

# computation -> (reduce buffer?, record column); FLUX-type entries take flux columns in order
_SINGLE = {
    "VALUE<Step>": (False, T.REC_STEP),
    "VALUE<Time>": (False, T.REC_TIME),
    "AVG<Identity>": (True, T.REC_NORM),
    "AVG<PotentialEnergy>": (True, T.REC_EPOT),
    "AVG<KineticEnergy>": (True, T.REC_EKIN),
    "ENERGY_TOTAL": (True, T.REC_ETOT),
    "ENERGY_DIFFERENCE": (False, T.REC_EDIFF),
}
_FLUX = {"ZOA_FLUX_2D": 7, "ZOA_FLUX_3D": 8, "ZOA_FLUX_2D_SIMPLE": 5}


def layout(computations):
    """(plain columns, reduce columns) of the record rows for a BufferedOutputs list, given as
    names ("AVG<Identity>", "ENERGY_TOTAL", "ZOA_FLUX_2D", "VALUE<Step, Time>", ...) or
    ("FLUX", k) for a FLUX of k borders. Flux computations take the record's flux columns in the
    order the borders were passed to the propagator."""
    plain, red = [], []
    flux = 0
    for c in computations:
        if isinstance(c, tuple) and c[0] == "FLUX":
            k = int(c[1])
        elif c in _FLUX:
            k = _FLUX[c]
        elif isinstance(c, str) and c.startswith("VALUE<") and "," in c:
            for op in c[len("VALUE<"):-1].split(","):
                plain.append(_SINGLE[f"VALUE<{op.strip()}>"][1])
            continue
        elif c in _SINGLE:
            is_red, col = _SINGLE[c]
            (red if is_red else plain).append(col)
            continue
        else:
            raise ValueError(f"computation {c!r} has no record column")
        red.extend(T.REC_FLUX0 + flux + i for i in range(k))
        flux += k
    return plain, red


def lines(records, comp_interval=1, log_interval=1, finished=True):
    """Indices of the record rows in the order the reference writes them. Row 0 is the AT_START
    computeEach; with `finished` the last row is AT_END, the others run DURING."""
    if comp_interval <= 0 or log_interval < comp_interval or log_interval % comp_interval:
        raise ValueError("log_interval must be a positive multiple of comp_interval "
                         "(BufferedOutputs::test, wf/buffer.h:399-410)")
    rows = len(records)
    height = log_interval // comp_interval
    buf = [None] * height
    cur = height - 1
    out = []
    for r in range(rows):
        step = int(records[r][T.REC_STEP])
        when = "start" if r == 0 else ("end" if finished and r == rows - 1 else "during")
        buf[cur] = r
        if (log_interval > 0 and step % log_interval == 0) or when == "end":
            lo, hi = {"start": (height - 1, height), "during": (0, height), "end": (0, cur + 1)}[when]
            out.extend(buf[i] for i in range(lo, hi) if buf[i] is not None)
            cur = 0
        elif step % comp_interval == 0:
            cur += 1
    return out


def write(name, records, computations, comp_interval=1, log_interval=1, binary=True, finished=True):
    """Write `<name>.dat` from a propagator's record rows; returns the file name."""
    rec = np.asarray(records, dtype=np.float64)
    plain, red = layout(computations)
    cols = plain + red
    fname = str(name) + ".dat"
    with open(fname, "wb" if binary else "w") as f:
        for r in lines(rec, comp_interval, log_interval, finished):
            row = rec[r, cols]
            if binary:
                f.write(row.tobytes())
            else:
                f.write("".join("|%13.5g " % v for v in row) + "\n")
    return fname


def read(fname, ncols, binary=True):
    """The rows of a `.dat` file with `ncols` values per line."""
    if binary:
        return np.fromfile(fname, dtype=np.float64).reshape(-1, ncols)
    out = []
    with open(fname) as f:
        for line in f:
            out.append([float(v) for v in line.replace("|", " ").split()])
    return np.array(out)
