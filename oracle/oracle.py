"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline. The product never imports it.
"""
import ctypes as C
import os

import numpy as np

from qsf_amd import qsf_types as T

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        P, D, I, I64 = C.c_void_p, C.c_double, C.c_int, C.c_int64
        dp = C.POINTER(C.c_double)
        sig = {
            "orc_fft1d": (None, [I64, I, dp, dp]),
            "orc_dft_bruteforce": (None, [I64, I, dp, dp]),
            "orc_set_threads": (I, [I]),
            "orc_set_create": (P, [C.POINTER(T.GridDesc)]),
            "orc_set_destroy": (None, [P]),
            "orc_set_count": (I, [P]),
            "orc_set_grid": (P, [P, I]),
            "orc_grid_size": (I64, [P]),
            "orc_grid_rep": (I, [P]),
            "orc_grid_set_psi": (None, [P, dp]),
            "orc_grid_get_psi": (None, [P, dp]),
            "orc_fourier": (None, [P, I]),
            "orc_evolve": (None, [P, I, I, D, dp]),
            "orc_average": (D, [P, I, I]),
            "orc_normalize": (None, [P]),
            "orc_flux": (None, [P, dp, C.POINTER(T.BorderDesc), I, dp]),
            "orc_current_map": (None, [P, dp, dp]),
            "orc_mask_region": (None, [P]),
            "orc_set_transfer": (None, [P]),
            "orc_set_post_compute": (None, [P]),
            "orc_set_transfer_slice": (I, [P, I, I, dp]),
            "orc_absorber_mask": (D, [P, I, C.POINTER(C.c_int64)]),
            "orc_absorber_absorb": (D, [P, D, I, C.POINTER(C.c_int64)]),
            "orc_border_vec": (None, [I, I, I, C.POINTER(C.c_int64), dp]),
            "orc_potential": (D, [P, dp]),
            "orc_potential_partial": (D, [P, I, dp]),
            "orc_coupling_sequence": (None, [C.POINTER(T.FieldDesc), I, I, dp, I, dp]),
            "orc_max_pulse_duration": (D, [C.POINTER(T.FieldDesc), I]),
            "orc_prop_create": (P, [P, C.POINTER(T.PropDesc), C.POINTER(T.FieldDesc), I]),
            "orc_prop_destroy": (None, [P]),
            "orc_prop_max_steps": (I64, [P]),
            "orc_prop_start": (None, [P]),
            "orc_prop_should_continue": (I, [P]),
            "orc_prop_step": (None, [P]),
            "orc_prop_finish": (None, [P]),
            "orc_prop_make_step": (None, [P]),
            "orc_prop_step_count": (I64, [P]),
            "orc_prop_time": (D, [P]),
            "orc_prop_nrecords": (I, [P]),
            "orc_prop_record_width": (I, [P]),
            "orc_prop_records": (None, [P, dp]),
            "orc_prop_coupling": (None, [P, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _cx(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def fft1d(x, sign=-1):
    x = _cx(x)
    out = np.empty_like(x)
    lib().orc_fft1d(x.size, sign, _dp(x), _dp(out))
    return out


def set_threads(k):
    """OpenMP threads of the oracle's independent loops (FFT lines, pointwise phases); sums stay
    serial, so every result is the same at any thread count. Returns the count in use."""
    return lib().orc_set_threads(int(k))


def dft_bruteforce(x, sign=-1):
    x = _cx(x)
    out = np.empty_like(x)
    lib().orc_dft_bruteforce(x.size, sign, _dp(x), _dp(out))
    return out


def border_vec_d(kind, size, dim, idx):
    idx = list(idx) + [0] * (3 - len(idx))
    a = (C.c_int64 * 3)(*idx)
    out = np.zeros(3)
    lib().orc_border_vec(kind, size, dim, a, _dp(out))
    return out[:dim]


def coupling_sequence(fields, gauge, times):
    arr = T.field_array(fields)
    t = np.ascontiguousarray(times, dtype=np.float64)
    out = np.zeros((t.size, 3))
    lib().orc_coupling_sequence(arr, len(fields), gauge, _dp(t), t.size, _dp(out))
    return out


def max_pulse_duration(fields):
    return lib().orc_max_pulse_duration(T.field_array(fields), len(fields))


class Grid:
    """One oracle grid (borrowed from a GridSet)."""

    def __init__(self, owner, handle):
        self._owner, self.h = owner, handle
        self.m = lib().orc_grid_size(handle)

    @property
    def rep(self):
        return lib().orc_grid_rep(self.h)

    def set_psi(self, psi):
        a = _cx(psi).ravel()
        assert a.size == self.m
        lib().orc_grid_set_psi(self.h, _dp(a))

    def psi(self):
        a = np.empty(self.m, np.complex128)
        lib().orc_grid_get_psi(self.h, _dp(a))
        return a

    def fourier(self, rep):
        lib().orc_fourier(self.h, rep)

    def evolve(self, mode, rep, delta, coupling=None, couples_in=None):
        # the coupling acts only in the representation it couples in (hamiltonian.h:44-50)
        if couples_in is not None and couples_in != rep:
            coupling = None
        c = None
        if coupling is not None:
            self._keep = np.ascontiguousarray(coupling, dtype=np.float64)
            c = _dp(self._keep)
        lib().orc_evolve(self.h, mode, rep, delta, c)

    def average(self, rep, op):
        return lib().orc_average(self.h, rep, op)

    def normalize(self):
        lib().orc_normalize(self.h)

    def flux(self, borders, vg_coupling=None):
        arr = (T.BorderDesc * len(borders))()
        for i, b in enumerate(borders):
            kind, size = b if isinstance(b, tuple) else (b, 0)
            arr[i].kind, arr[i].size = kind, size
        out = np.zeros(len(borders))
        vg = None
        if vg_coupling is not None:
            self._vg = np.ascontiguousarray(vg_coupling, dtype=np.float64)
            vg = _dp(self._vg)
        lib().orc_flux(self.h, vg, arr, len(borders), _dp(out))
        return out

    def current_map(self, dim, vg_coupling=None):
        out = np.zeros(self.m * dim)
        vg = None
        if vg_coupling is not None:
            self._vg = np.ascontiguousarray(vg_coupling, dtype=np.float64)
            vg = _dp(self._vg)
        lib().orc_current_map(self.h, vg, _dp(out))
        return out.reshape(self.m, dim)

    def mask_region(self):
        lib().orc_mask_region(self.h)

    def absorber_mask(self, nodes):
        a = (C.c_int64 * 3)(*(list(nodes) + [0] * (3 - len(nodes))))
        return lib().orc_absorber_mask(self.h, len(nodes), a)

    def absorber_absorb(self, delta, nodes):
        a = (C.c_int64 * 3)(*(list(nodes) + [0] * (3 - len(nodes))))
        return lib().orc_absorber_absorb(self.h, delta, len(nodes), a)

    def potential(self, coords):
        c = np.zeros(3)
        c[:len(coords)] = coords
        return lib().orc_potential(self.h, _dp(c))

    def potential_partial(self, free_axes, coords):
        c = np.zeros(3)
        c[:len(coords)] = coords
        return lib().orc_potential_partial(self.h, free_axes, _dp(c))


class GridSet:
    """A single grid (multigrid=0) or the 2^D regions of a multigrid (multigrid=1)."""

    def __init__(self, desc):
        self.desc = desc
        self.h = lib().orc_set_create(C.byref(desc))
        self.grids = [Grid(self, lib().orc_set_grid(self.h, r))
                      for r in range(lib().orc_set_count(self.h))]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_set_destroy(self.h)
            self.h = None

    def __getitem__(self, r):
        return self.grids[r]

    def __len__(self):
        return len(self.grids)

    def transfer(self):
        lib().orc_set_transfer(self.h)

    def post_compute(self):
        lib().orc_set_post_compute(self.h)

    def transfer_slice(self, region, source_index):
        m = self.grids[region].m
        out = np.zeros(m, np.complex128)
        k = lib().orc_set_transfer_slice(self.h, region, source_index, _dp(out))
        if k < 0:
            raise IndexError("no such source")
        return out[:k]


class Propagator:
    """SplitPropagator (propagator.h:51-499) over a GridSet."""

    def __init__(self, gset, pdesc, fields):
        self.gset = gset
        self._fields = T.field_array(fields)
        self._pd = pdesc
        self.h = lib().orc_prop_create(gset.h, C.byref(pdesc), self._fields, len(fields))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_prop_destroy(self.h)
            self.h = None

    @property
    def max_steps(self):
        return lib().orc_prop_max_steps(self.h)

    @property
    def step_count(self):
        return lib().orc_prop_step_count(self.h)

    @property
    def time(self):
        return lib().orc_prop_time(self.h)

    def start(self):
        lib().orc_prop_start(self.h)

    def should_continue(self):
        return bool(lib().orc_prop_should_continue(self.h))

    def step(self):
        lib().orc_prop_step(self.h)

    def finish(self):
        lib().orc_prop_finish(self.h)

    def make_step(self):
        lib().orc_prop_make_step(self.h)

    def coupling(self):
        out = np.zeros(3)
        lib().orc_prop_coupling(self.h, _dp(out))
        return out

    def records(self):
        n = lib().orc_prop_nrecords(self.h)
        w = lib().orc_prop_record_width(self.h)
        out = np.zeros(n * w)
        if n:
            lib().orc_prop_records(self.h, _dp(out))
        return out.reshape(n, w)

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
