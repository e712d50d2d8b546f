"""ctypes binding of libqsf_hip.so (C ABI in include/qsf_hip.h).

The library is the product: every numeric operation below runs in its HIP kernels. Loading
fails loudly when the library is missing; there is no CPU fallback.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from . import qsf_types as T

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("QSF_HIP_LIB") or os.path.join(_HERE, "libqsf_hip.so")
_LIB = None

# (restype, argtypes) of every exported entry point
_P, _D, _I, _I64 = C.c_void_p, C.c_double, C.c_int, C.c_int64
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "qsf_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "qsf_ctx_destroy": (_I, [_P]),
    "qsf_last_error": (C.c_char_p, [_P]),
    "qsf_ctx_synchronize": (_I, [_P]),
    "qsf_ctx_query": (_I, [_P, C.POINTER(_I)]),
    "qsf_prop_fast_forward": (_I, [_P, _I64]),
    "qsf_antisymmetrize": (_I, [_P, _I]),
    "qsf_cross_out": (_I, [_P, C.c_double]),
    "qsf_orthogonalize": (_I, [_P, _dp, _I64, _dp]),
    "qsf_timer_mark": (_I, [_P]),
    "qsf_timer_elapsed_ms": (_I, [_P, _dp]),
    "qsf_set_create": (_I, [_P, C.POINTER(T.GridDesc), C.POINTER(_P)]),
    "qsf_set_destroy": (_I, [_P]),
    "qsf_set_count": (_I, [_P, C.POINTER(_I)]),
    "qsf_set_grid": (_I, [_P, _I, C.POINTER(_P)]),
    "qsf_set_post_compute": (_I, [_P]),
    "qsf_grid_size": (_I, [_P, C.POINTER(_I64)]),
    "qsf_grid_rep": (_I, [_P, C.POINTER(_I)]),
    "qsf_grid_upload": (_I, [_P, _dp, _I64, _I]),
    "qsf_grid_download": (_I, [_P, _dp, _I64]),
    "qsf_fourier": (_I, [_P, _I]),
    "qsf_evolve": (_I, [_P, _I, _I, _D, _dp, _I]),
    "qsf_average": (_I, [_P, _I, _I, _dp]),
    "qsf_average_noalloc": (_I, [_P, _I, _I, _dp]),
    "qsf_normalize": (_I, [_P]),
    "qsf_flux": (_I, [_P, _dp, C.POINTER(T.BorderDesc), _I, _dp]),
    "qsf_mask_region": (_I, [_P]),
    "qsf_prop_create": (_I, [_P, C.POINTER(T.PropDesc), C.POINTER(T.FieldDesc), _I, C.POINTER(_P)]),
    "qsf_prop_destroy": (_I, [_P]),
    "qsf_prop_max_steps": (_I, [_P, C.POINTER(_I64)]),
    "qsf_prop_start": (_I, [_P]),
    "qsf_prop_should_continue": (_I, [_P, C.POINTER(_I)]),
    "qsf_prop_step": (_I, [_P]),
    "qsf_prop_run_steps": (_I, [_P, _I64, C.POINTER(_I64)]),
    "qsf_prop_finish": (_I, [_P]),
    "qsf_prop_step_count": (_I, [_P, C.POINTER(_I64)]),
    "qsf_prop_time": (_I, [_P, _dp]),
    "qsf_prop_coupling": (_I, [_P, _dp]),
    "qsf_prop_record_width": (_I, [_P, C.POINTER(_I)]),
    "qsf_prop_nrecords": (_I, [_P, C.POINTER(_I)]),
    "qsf_prop_records": (_I, [_P, _dp, _I]),
    "qsf_prop_flux_integral": (_I, [_P, _dp]),
    "qsf_prop_set_exec": (_I, [_P, _I]),
    "qsf_prof_begin": (_I, [_P]),
    "qsf_prof_end": (_I, [_P, C.POINTER(_I)]),
    "qsf_prof_entry": (_I, [_P, _I, C.c_char_p, _I, C.POINTER(_I64), _dp, _dp]),
    "qsf_autotune": (_I, [_P, _I]),
    "qsf_tune_count": (_I, [_P, C.POINTER(_I)]),
    "qsf_tune_entry": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I64), C.POINTER(_I64),
                            C.POINTER(_I), _dp]),
    "qsf_tune_set": (_I, [_P, _I, _I, _I64, _I64, _I]),
    "qsf_batch_create": (_I, [_P, C.POINTER(T.GridDesc), C.POINTER(T.FieldDesc), _I, C.POINTER(T.PropDesc),
                              C.POINTER(_P)]),
    "qsf_batch_destroy": (_I, [_P]),
    "qsf_batch_upload": (_I, [_P, _I, _dp]),
    "qsf_batch_download": (_I, [_P, _I, _dp]),
    "qsf_batch_start": (_I, [_P]),
    "qsf_batch_run": (_I, [_P, _I64, _I, C.POINTER(_I)]),
    "qsf_batch_finish": (_I, [_P]),
    "qsf_batch_info": (_I, [_P, _I, C.POINTER(_I64), C.POINTER(_I64)]),
    "qsf_batch_records": (_I, [_P, _I, _dp, _I, C.POINTER(_I)]),
    "qsf_batch_flux_integral": (_I, [_P, _I, _dp]),
}


def load():
    """Load libqsf_hip.so and bind its C ABI (raises if it is missing)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libqsf_hip.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _LIB = L
    return _LIB


class QSFError(RuntimeError):
    pass


_EXITING = [False]


@atexit.register
def _mark_exit():
    # at interpreter exit the HIP runtime may already be gone: leave handles to the OS
    _EXITING[0] = True


def _dptr(a):
    return a.ctypes.data_as(_dp)


class Context:
    """One device and one HIP stream (qsf_ctx)."""

    def __init__(self, device=0):
        L = load()
        h = _P()
        rc = L.qsf_ctx_create(device, C.byref(h))
        if rc != 0:
            raise QSFError(f"qsf_ctx_create: {L.qsf_last_error(None).decode()}")
        self.h = h
        self._children = weakref.WeakSet()

    def adopt(self, obj):
        self._children.add(obj)

    def check(self, rc, what):
        if rc != 0:
            msg = load().qsf_last_error(self.h).decode()
            raise QSFError(f"{what} failed ({rc}): {msg}")

    def synchronize(self):
        self.check(load().qsf_ctx_synchronize(self.h), "synchronize")

    def idle(self):
        """True when everything enqueued on this context's stream has finished (no wait)."""
        v = C.c_int()
        self.check(load().qsf_ctx_query(self.h, C.byref(v)), "ctx_query")
        return bool(v.value)

    def mark(self):
        self.check(load().qsf_timer_mark(self.h), "timer_mark")

    def profile_begin(self):
        """Record HIP events around every axis-pass launch from now on."""
        self.check(load().qsf_prof_begin(self.h), "prof_begin")

    def profile_end(self):
        """Stop recording; return {kernel: (launches, total_ms, algorithmic_bytes_per_launch)}."""
        n = C.c_int()
        self.check(load().qsf_prof_end(self.h, C.byref(n)), "prof_end")
        out = {}
        buf = C.create_string_buffer(256)
        for i in range(n.value):
            k, ms, by = C.c_int64(), C.c_double(), C.c_double()
            self.check(load().qsf_prof_entry(self.h, i, buf, 256, C.byref(k), C.byref(ms), C.byref(by)),
                       "prof_entry")
            out[buf.value.decode()] = (k.value, ms.value, by.value)
        return out

    def tune_table(self):
        """[(precision, row, n, grid_size, variant, best_ms)] chosen by the autotuner; `row` is the
        pass class: 1 row pass, 0 column pass of an inner axis, 2 column pass of the outermost
        axis."""
        n = C.c_int()
        self.check(load().qsf_tune_count(self.h, C.byref(n)), "tune_count")
        out = []
        for i in range(n.value):
            pr, row, v = C.c_int(), C.c_int(), C.c_int()
            nn, m, ms = C.c_int64(), C.c_int64(), C.c_double()
            self.check(load().qsf_tune_entry(self.h, i, C.byref(pr), C.byref(row), C.byref(nn), C.byref(m),
                                             C.byref(v), C.byref(ms)), "tune_entry")
            out.append((pr.value, row.value, nn.value, m.value, v.value, ms.value))
        return out

    def tune_set(self, precision, row, n, grid_size, variant):
        self.check(load().qsf_tune_set(self.h, precision, row, n, grid_size, variant), "tune_set")

    def elapsed_ms(self):
        v = C.c_double()
        self.check(load().qsf_timer_elapsed_ms(self.h, C.byref(v)), "timer_elapsed")
        return v.value

    def close(self):
        if getattr(self, "h", None) and not _EXITING[0]:
            for ch in [c for c in self._children if isinstance(c, Propagator)]:
                ch.close()
            for ch in list(self._children):
                ch.close()
            load().qsf_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuning_cache_path():
    base = os.environ.get("QSF_TUNE_CACHE") or os.path.join(
        os.environ.get("XDG_CACHE_HOME", os.path.expanduser("~/.cache")), "qsf_amd", "autotune.json")
    return base


def load_tuning(ctx, path=None):
    """Apply kernel-variant choices saved by save_tuning (missing file: nothing to do)."""
    import json
    path = path or tuning_cache_path()
    try:
        with open(path) as f:
            entries = json.load(f)
    except (OSError, ValueError):
        return 0
    for e in entries:
        ctx.tune_set(e["precision"], e["row"], e["n"], e["grid_size"], e["variant"])
    return len(entries)


def save_tuning(ctx, path=None):
    import json
    path = path or tuning_cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    rows = [dict(precision=p, row=r, n=n, grid_size=m, variant=v, best_ms=ms)
            for (p, r, n, m, v, ms) in ctx.tune_table()]
    with open(path, "w") as f:
        json.dump(rows, f, indent=1)
    return path


class Grid:
    """One device grid of a GridSet (the HamWF member surface)."""

    def __init__(self, owner, handle):
        self._owner, self.h = owner, handle
        self.ctx = owner.ctx
        m = C.c_int64()
        self.ctx.check(load().qsf_grid_size(handle, C.byref(m)), "grid_size")
        self.m = m.value

    @property
    def rep(self):
        r = C.c_int()
        self.ctx.check(load().qsf_grid_rep(self.h, C.byref(r)), "grid_rep")
        return r.value

    def set_psi(self, psi, rep=T.REP_X):
        a = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128)).ravel()
        self.ctx.check(load().qsf_grid_upload(self.h, _dptr(a), a.size, rep), "upload")

    def psi(self):
        a = np.empty(self.m, np.complex128)
        self.ctx.check(load().qsf_grid_download(self.h, _dptr(a), self.m), "download")
        return a

    def fourier(self, rep):
        self.ctx.check(load().qsf_fourier(self.h, rep), "fourier")

    def antisymmetrize(self, particles):
        """operation<AntiSymmetrize<particles>> on the X state (wf/grid.h:611-723)."""
        self.ctx.check(load().qsf_antisymmetrize(self.h, particles), "antisymmetrize")

    def cross_out(self, au_distance=50.0):
        """croossOut (wf/grid.h:890-907) on the X state."""
        self.ctx.check(load().qsf_cross_out(self.h, float(au_distance)), "cross_out")

    def orthogonalize(self, phi):
        """orthogonalizeWith (wf/grid.h:909-946) against the state `phi`; returns <phi|psi> dV."""
        a = np.ascontiguousarray(np.asarray(phi, dtype=np.complex128)).ravel()
        out = np.zeros(2)
        self.ctx.check(load().qsf_orthogonalize(self.h, _dptr(a), a.size, _dptr(out)), "orthogonalize")
        return complex(out[0], out[1])

    def evolve(self, mode, rep, delta, coupling=None, couples_in=None):
        c = None
        if coupling is not None:
            self._c = np.ascontiguousarray(coupling, dtype=np.float64)
            c = _dptr(self._c)
        ci = rep if (couples_in is None and coupling is not None) else (couples_in or 0)
        self.ctx.check(load().qsf_evolve(self.h, mode, rep, delta, c, ci), "evolve")

    def average(self, rep, op):
        v = C.c_double()
        self.ctx.check(load().qsf_average(self.h, rep, op, C.byref(v)), "average")
        return v.value

    def average_noalloc(self, rep, op):
        """average() through the entry point that allocates no device memory."""
        v = C.c_double()
        self.ctx.check(load().qsf_average_noalloc(self.h, rep, op, C.byref(v)), "average_noalloc")
        return v.value

    def normalize(self):
        self.ctx.check(load().qsf_normalize(self.h), "normalize")

    def flux(self, borders, vg_coupling=None):
        arr = (T.BorderDesc * len(borders))()
        for i, b in enumerate(borders):
            kind, size = b if isinstance(b, tuple) else (b, 0)
            arr[i].kind, arr[i].size = kind, size
        out = np.zeros(len(borders))
        vg = None
        if vg_coupling is not None:
            self._vg = np.ascontiguousarray(vg_coupling, dtype=np.float64)
            vg = _dptr(self._vg)
        self.ctx.check(load().qsf_flux(self.h, vg, arr, len(borders), _dptr(out)), "flux")
        return out

    def mask_region(self):
        self.ctx.check(load().qsf_mask_region(self.h), "mask_region")

    def autotune(self, reps=5):
        """Pick the fastest compiled kernel variant of each pass this grid's step uses."""
        self.ctx.check(load().qsf_autotune(self.h, reps), "autotune")


class GridSet:
    """One grid (multigrid=0) or the 2^D multigrid regions (multigrid=1) on the device."""

    def __init__(self, ctx, desc):
        self.ctx, self.desc = ctx, desc
        h = _P()
        ctx.check(load().qsf_set_create(ctx.h, C.byref(desc), C.byref(h)), "set_create")
        self.h = h
        ctx.adopt(self)
        self._props = weakref.WeakSet()
        n = C.c_int()
        ctx.check(load().qsf_set_count(h, C.byref(n)), "set_count")
        self.grids = []
        for r in range(n.value):
            g = _P()
            ctx.check(load().qsf_set_grid(h, r, C.byref(g)), "set_grid")
            self.grids.append(Grid(self, g))

    def __getitem__(self, r):
        return self.grids[r]

    def __len__(self):
        return len(self.grids)

    def post_compute(self):
        self.ctx.check(load().qsf_set_post_compute(self.h), "post_compute")

    def close(self):
        # propagators hold the set: they go first
        for p in list(getattr(self, "_props", ())):
            p.close()
        if getattr(self, "h", None) and not _EXITING[0] and getattr(self.ctx, "h", None):
            load().qsf_set_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Propagator:
    """SplitPropagator over a device GridSet (fused axis passes by default)."""

    def __init__(self, gset, pdesc, fields, exec_mode=0):
        self.gset, self.ctx = gset, gset.ctx
        self._fields = T.field_array(fields)
        self._pd = pdesc
        h = _P()
        self.ctx.check(load().qsf_prop_create(gset.h, C.byref(pdesc), self._fields, len(fields),
                                              C.byref(h)), "prop_create")
        self.h = h
        self.ctx.adopt(self)
        gset._props.add(self)
        if exec_mode:
            self.ctx.check(load().qsf_prop_set_exec(h, exec_mode), "prop_set_exec")

    def close(self):
        # valid before or after the set's own destroy (the library detaches a propagator whose
        # set went first); only the context must still be alive
        if getattr(self, "h", None) and not _EXITING[0] and getattr(self.ctx, "h", None):
            load().qsf_prop_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _i64(self, fn):
        v = C.c_int64()
        self.ctx.check(fn(self.h, C.byref(v)), fn.__name__)
        return v.value

    @property
    def max_steps(self):
        return self._i64(load().qsf_prop_max_steps)

    @property
    def step_count(self):
        return self._i64(load().qsf_prop_step_count)

    @property
    def time(self):
        v = C.c_double()
        self.ctx.check(load().qsf_prop_time(self.h, C.byref(v)), "prop_time")
        return v.value

    def coupling(self):
        out = np.zeros(3)
        self.ctx.check(load().qsf_prop_coupling(self.h, _dptr(out)), "prop_coupling")
        return out

    def start(self):
        self.ctx.check(load().qsf_prop_start(self.h), "prop_start")

    def fast_forward(self, at_step):
        """Resume in place of start(): makeFakeStep until step == at_step (propagator.h:425-436)."""
        self.ctx.check(load().qsf_prop_fast_forward(self.h, int(at_step)), "prop_fast_forward")

    def should_continue(self):
        v = C.c_int()
        self.ctx.check(load().qsf_prop_should_continue(self.h, C.byref(v)), "should_continue")
        return bool(v.value)

    def step(self):
        self.ctx.check(load().qsf_prop_step(self.h), "prop_step")

    def run_steps(self, k):
        done = C.c_int64()
        self.ctx.check(load().qsf_prop_run_steps(self.h, k, C.byref(done)), "prop_run_steps")
        return done.value

    def finish(self):
        self.ctx.check(load().qsf_prop_finish(self.h), "prop_finish")

    def records(self):
        n, w = C.c_int(), C.c_int()
        self.ctx.check(load().qsf_prop_nrecords(self.h, C.byref(n)), "nrecords")
        self.ctx.check(load().qsf_prop_record_width(self.h, C.byref(w)), "record_width")
        out = np.zeros(n.value * w.value)
        if n.value:
            self.ctx.check(load().qsf_prop_records(self.h, _dptr(out), n.value), "records")
        return out.reshape(n.value, w.value)

    def flux_integral(self):
        out = np.zeros(max(1, self._pd.nborders))
        self.ctx.check(load().qsf_prop_flux_integral(self.h, _dptr(out)), "flux_integral")
        return out[: self._pd.nborders]

    def run(self, limit=None):
        """run() (propagator.h:420-462); `limit` caps the loop iterations for small tests."""
        self.start()
        k = 0
        while self.should_continue() and (limit is None or k < limit):
            self.step()
            k += 1
        if limit is None or k < limit:
            self.finish()
        return self.records()


class Batch:
    """Many single-region 2D solves advanced together (batched launches, hipGraph replay)."""

    def __init__(self, ctx, descs, fields, pdesc):
        self.ctx, self._pd = ctx, pdesc
        n = len(descs)
        self._descs = (T.GridDesc * n)(*descs)
        self._fields = (T.FieldDesc * n)(*fields)
        h = _P()
        ctx.check(load().qsf_batch_create(ctx.h, self._descs, self._fields, n, C.byref(pdesc), C.byref(h)),
                  "batch_create")
        self.h, self.n = h, n
        ctx.adopt(self)
        self.sizes = []
        for s in range(n):
            ms, m = C.c_int64(), C.c_int64()
            ctx.check(load().qsf_batch_info(h, s, C.byref(ms), C.byref(m)), "batch_info")
            self.sizes.append((ms.value, m.value))

    def close(self):
        if getattr(self, "h", None) and not _EXITING[0] and getattr(self.ctx, "h", None):
            load().qsf_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_psi(self, s, psi):
        a = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128)).ravel()
        assert a.size == self.sizes[s][1]
        self.ctx.check(load().qsf_batch_upload(self.h, s, _dptr(a)), "batch_upload")

    def psi(self, s):
        a = np.empty(self.sizes[s][1], np.complex128)
        self.ctx.check(load().qsf_batch_download(self.h, s, _dptr(a)), "batch_download")
        return a

    def start(self):
        self.ctx.check(load().qsf_batch_start(self.h), "batch_start")

    def run(self, steps, use_graph=True):
        done = C.c_int()
        self.ctx.check(load().qsf_batch_run(self.h, steps, 1 if use_graph else 0, C.byref(done)), "batch_run")
        return bool(done.value)

    def finish(self):
        self.ctx.check(load().qsf_batch_finish(self.h), "batch_finish")

    def records(self, s):
        rows = C.c_int()
        self.ctx.check(load().qsf_batch_records(self.h, s, None, 0, C.byref(rows)), "batch_records")
        w = 3 + self._pd.nborders
        out = np.zeros(rows.value * w)
        if rows.value:
            self.ctx.check(load().qsf_batch_records(self.h, s, _dptr(out), rows.value, C.byref(rows)),
                           "batch_records")
        return out.reshape(rows.value, w)

    def flux_integral(self, s):
        out = np.zeros(max(1, self._pd.nborders))
        self.ctx.check(load().qsf_batch_flux_integral(self.h, s, _dptr(out)), "batch_flux_integral")
        return out[: self._pd.nborders]

    def run_all(self, use_graph=True, chunk=1 << 30):
        self.start()
        while not self.run(chunk, use_graph):
            pass
        self.finish()


def autotune_batch(ctx, descs, fields, pdesc, state_of, steps=8):
    """Pick the kernel variants of every size group of a batch by timing that group's batched step
    (graph replay) once per compiled variant: first the column pass, then the row pass with the
    column's choice. The picks go into the context's tuning table under the keys the batch looks
    up (precision, pass class, axis length, grid size). `state_of(desc)` gives a start state.
    Returns {(precision, class, length, grid size): (variant, ms per step)}."""
    import time
    groups = {}
    for i, d in enumerate(descs):
        shape = tuple(int(d.n[a]) for a in range(d.dim))
        groups.setdefault((d.dim, shape, d.precision), []).append(i)
    chosen = {}
    for (dim, shape, prec), idx in groups.items():
        m = int(np.prod(shape))
        sub_d = [descs[i] for i in idx]
        sub_f = [fields[i] for i in idx]
        # the outermost axis of a 2D solve is its column pass (class 2); the last axis its row pass
        passes = ([(2, shape[0])] if dim == 2 else []) + [(1, shape[-1])]
        current = {cls: v for (pr, cls, n, mm, v, _ms) in ctx.tune_table()
                   for c2, n2 in passes if pr == prec and mm == m and cls == c2 and n == n2}

        def step_ms():
            b = Batch(ctx, sub_d, sub_f, pdesc)
            try:
                for s, d in enumerate(sub_d):
                    b.set_psi(s, state_of(d))
                b.start()
                b.run(2, use_graph=True)
                ctx.synchronize()
                t0 = time.perf_counter()
                b.run(steps, use_graph=True)
                ctx.synchronize()
                return (time.perf_counter() - t0) / steps * 1e3
            finally:
                b.close()

        for cls, n in passes:
            for c2, n2 in passes:   # the other pass at its current pick
                ctx.tune_set(prec, c2, n2, m, current.get(c2, 0))
            best = None
            for v in range(6):
                try:
                    ctx.tune_set(prec, cls, n, m, v)
                except QSFError:
                    break           # lengths without tuned variants have variant 0 only
                ms = step_ms()
                if best is None or ms < best[1]:
                    best = (v, ms)
            current[cls] = best[0]
            ctx.tune_set(prec, cls, n, m, best[0])
            chosen[(prec, cls, n, m)] = best
    return chosen
