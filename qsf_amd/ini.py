"""`project.ini` runs as descriptors: the reference's INI schema read into the C-ABI descriptors,
so a configured run and a serving request come from one parser.

Parsing follows the INI reader QSF vendors (inipp, basics/inipp.h) as `Config` drives it
(basics/config.h:14-31): parse, strip trailing comments, merge [DEFAULT] into every section,
interpolate.
* Lines are trimmed; a line starting with ';', '#' or '/' is a comment; "[name]" opens a
  section; "key = value" splits at the first '=' (key right-trimmed, value left-trimmed); the
  first occurrence of a key in a section wins (inipp.h:199-236).
* Trailing comments: a value is cut at its first ';', '#' or '/', then right-trimmed
  (inipp.h:267-276) — so values cannot hold those characters.
* [DEFAULT] keys are added to every section that lacks them (inipp.h:260-265).
* "${key}" means "${section:key}" of the same section; "${section:key}" is replaced by that
  value, repeated up to 10 rounds (inipp.h:243-258).
* A number is read only when the whole value parses as one (inipp.h:88-103): "2048.0" is no
  integer, and the field keeps its default.

Keys (per section): dt, max_steps, state_accuracy (propagator.h:144-148); n, dx
(coords.h:99-100); nCAP (absorbers.h:35); Ncharge, Echarge, Nsoft, Esoft (potential.h:29-32);
F::<ax>, omega::<ax>, ncycles::<ax>, FWHM::<ax>, phase::<ax>, delay::<ax> with <ax> the field's
axis name X, Y, Z, XY, XZ, YZ, XYZ (field_p.h:19-28, field.h:50, basics/enums.h:68-82);
comp_interval, log_interval (buffer.h:101-102). The template arguments of a QSF program (dims,
potential, absorber, gauge, split order, outputs) are not in the file: they are arguments here.
"""
import re

from . import qsf_types as T

_COMMENT = ";#/"
_AXIS_NAMES = {T.AXIS_X: "X", T.AXIS_Y: "Y", T.AXIS_Z: "Z", T.AXIS_XY: "XY", T.AXIS_XZ: "XZ",
               T.AXIS_YZ: "YZ", T.AXIS_XYZ: "XYZ"}


def parse(text):
    """Sections of an INI text after parse, strip_trailing_comments, [DEFAULT] merge and
    interpolation. Returns ({section: {key: value}}, [lines that did not parse])."""
    sections, errors = {}, []
    section = ""
    for raw in text.splitlines():
        line = raw.strip()
        if not line:
            continue
        eq = line.find("=")
        if line[0] in _COMMENT:
            continue
        if line[0] == "[":
            if line[-1] == "]":
                section = line[1:-1]
            else:
                errors.append(line)
        elif eq > 0:
            key, value = line[:eq].rstrip(), line[eq + 1:].lstrip()
            sec = sections.setdefault(section, {})
            if key in sec:
                errors.append(line)
            else:
                sec[key] = value
        else:
            errors.append(line)
    for sec in sections.values():                       # strip_trailing_comments
        for k, v in sec.items():
            cut = min([i for i, c in enumerate(v) if c in _COMMENT], default=len(v))
            sec[k] = v[:cut].rstrip()
    default = dict(sections.get("DEFAULT", {}))         # default_section
    for sec in sections.values():
        for k, v in default.items():
            sec.setdefault(k, v)
    for name, sec in sections.items():                  # ${key} -> ${section:key}
        for k in list(sec):
            for key in list(sec):
                sec[k] = sec[k].replace("${" + key + "}", "${" + name + ":" + key + "}")
    for _ in range(11):                                 # ${section:key} -> value
        changed = False
        table = {"${%s:%s}" % (n, k): v for n, s in sections.items() for k, v in s.items()}
        for sec in sections.values():
            for k, v in sec.items():
                nv = v
                for sym, val in table.items():
                    nv = nv.replace(sym, val)
                if nv != v:
                    sec[k], changed = nv, True
        if not changed:
            break
    return sections, errors


_INT = re.compile(r"^[+-]?\d+$")
_FLOAT = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$")


def get(sec, key, kind, default=None):
    """inipp::get_value: the value if the whole of it reads as `kind` (int or float)."""
    v = sec.get(key)
    if v is None:
        return default
    s = v.strip()
    if kind is int:
        return int(s) if _INT.match(s) else default
    if not _FLOAT.match(s):          # the forms operator>> reads whole (no inf/nan, no '_')
        return default
    return float(s)


def grid_desc(sec, dim, potential=T.POT_NONE, absorber=T.ABS_NONE, multigrid=0,
              precision=T.FP64, absorber_tau=None):
    """CartesianGrid<dim>(dx, n) + CAP/cos^(1/8)(nCAP) + the potential's charges from a section
    (charges absent from the section keep the descriptor's defaults)."""
    n = get(sec, "n", int)
    dx = get(sec, "dx", float)
    if n is None or dx is None:
        raise KeyError("a grid needs n and dx")
    kw = {k: get(sec, k, float) for k in ("Ncharge", "Echarge", "Nsoft", "Esoft")}
    kw = {k: v for k, v in kw.items() if v is not None}
    if absorber != T.ABS_NONE:
        kw["nCAP"] = get(sec, "nCAP", int, 0)
    if absorber_tau is not None:
        kw["absorber_tau"] = absorber_tau
    return T.grid_desc(dim, n, dx, precision=precision, potential=potential, absorber=absorber,
                       multigrid=multigrid, **kw)


def fields(sec, axis, pulses=1, pulse=T.PULSE_SIN, envelope=T.ENV_SIN2):
    """Field<axis, Pulse...>(settings): every pulse of the field reads the same F::<ax> ...
    keys (field.h:50); missing keys leave 0 (_FieldConfig value-initialised)."""
    ax = "::" + _AXIS_NAMES[axis]
    vals = {k: get(sec, k + ax, float, 0.0) for k in ("F", "omega", "ncycles", "FWHM", "phase", "delay")}
    f = T.field_desc(axis, vals["F"], vals["omega"], vals["ncycles"], pulse=pulse, envelope=envelope,
                     fwhm_percent=vals["FWHM"], phase_in_pi=vals["phase"], delay_in_cycles=vals["delay"])
    return [f] * pulses


def prop_desc(sec, mode, order=1, gauge=T.GAUGE_VELOCITY, outputs=0, borders=()):
    """SplitPropagator(settings): dt, and in imaginary time max_steps and state_accuracy
    (propagator.h:144-148); comp_interval from the outputs' keys."""
    dt = get(sec, "dt", float)
    if dt is None:
        raise KeyError("a propagator needs dt")
    kw = {}
    if mode == T.MODE_IM:
        kw = dict(max_steps=get(sec, "max_steps", int, 0), state_accuracy=get(sec, "state_accuracy", float, 0.0))
    return T.prop_desc(mode, order, dt, gauge=gauge, outputs=outputs,
                       comp_interval=get(sec, "comp_interval", int, 1), borders=borders, **kw)


def intervals(sec):
    """(comp_interval, log_interval) of BufferedOutputs(settings) (buffer.h:101-102)."""
    return get(sec, "comp_interval", int, 1), get(sec, "log_interval", int, 1)


def request(sec, rid, arrival, psi0, dim, field_axis, potential=T.POT_NONE, absorber=T.ABS_NONE,
            precision=T.FP64, absorber_tau=None, order=1, gauge=T.GAUGE_VELOCITY, outputs=0,
            borders=()):
    """A serving request from the same section a project.ini run reads: real-time propagation of
    `psi0` on the section's grid under its field."""
    from .serve import Request
    d = grid_desc(sec, dim, potential=potential, absorber=absorber, precision=precision,
                  absorber_tau=absorber_tau)
    return Request(rid=rid, arrival=arrival, kind=f"{dim}d", desc=d,
                   pdesc=prop_desc(sec, T.MODE_RE, order=order, gauge=gauge, outputs=outputs, borders=borders),
                   fields=fields(sec, field_axis), psi0=psi0)
