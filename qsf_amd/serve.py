"""Serving loop for laser-parameter solves on one MI355X.

Requests (a grid, a propagation setting and a field) arrive over time. The server

  * admits a request when its HBM footprint fits what is left of the device budget (288 GB by
    default), first come first served, and keeps its state resident on the device until the
    result is read back;
  * advances every admitted solve together: each solve lives on one of several HIP streams
    (one library context each, small and large solves on separate streams) and gets a quantum
    of steps (about the same grid-point-steps for every solve) whenever its stream has drained,
    so kernels of different solves overlap on the device; solves join after admission and
    leave as soon as their last step is done (continuous batching at step boundaries);
  * reports each request's time to result (arrival -> result on the host), throughput and
    p50/p99.

Results do not depend on what else is in flight: every solve has its own state, tables and
reductions (fixed-order trees), and every context applies the same kernel-variant table, so a
solve computes bit-identical numbers alone and in the mix.

A request runs SplitPropagator::run (propagator.h:420-462) through the library's fused step; the
ionisation yield reported is 1 - norm at the end of the run (population taken by the absorber)
and the FP64 flux integrals through the requested borders.
"""
import dataclasses
import time

import numpy as np

from . import qsf_types as T

HBM_BYTES = 288 * 10**9
SMALL = 1 << 16   # grid points up to which a solve counts as small (its own streams)


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the stream starts
    kind: str                      # "1d", "2d", "3d"
    desc: object                   # qsf_types.GridDesc
    pdesc: object                  # qsf_types.PropDesc
    fields: list
    psi0: np.ndarray               # initial state (reference layout, complex128)
    shadow_of: int = -1            # >= 0: FP64 twin of that request (verification sample)


@dataclasses.dataclass
class Result:
    rid: int
    kind: str
    points: int
    steps: int
    arrival: float
    admitted: float
    finished: float
    records: np.ndarray
    flux_integral: np.ndarray
    final_norm: float
    shadow_of: int = -1

    @property
    def latency(self):
        return self.finished - self.arrival

    @property
    def ionization_yield(self):
        return 1.0 - self.final_norm


def footprint_bytes(desc, pdesc):
    """Device bytes one solve holds while resident: state and its work copy per region grid,
    the per-substep table ring, pair tables, reduction partials, the flux gather and records."""
    dim = desc.dim
    n = [int(desc.n[i]) for i in range(dim)]
    m = int(np.prod(n))
    cb = 16 if desc.precision == T.FP64 else 8
    regions = 2 ** dim if desc.multigrid else 1
    maxn = max(n)
    per_grid = 2 * m * cb + 48 * 3 * maxn * 16 + 2 * maxn * cb
    if desc.potential == T.POT_ECKHARDT_SACHA:
        per_grid += maxn * maxn * 16
    elif desc.potential == T.POT_EBERLY:
        per_grid += maxn * 16
    per_grid += (m // maxn + 1) * 3 * 8               # block partials
    edges = 4 * 2 * int(desc.nCAP) * maxn * cb if desc.multigrid else 0
    gather = (m // maxn) * 64 * cb if pdesc.nborders else 0
    records = 20000 * (8 + int(pdesc.nborders)) * 8
    return int(1.05 * (regions * per_grid + edges + gather + records)) + (1 << 20)


# ------------------------------------------------------------------------------ request streams
def _laser(rng):
    """Intensity log-U[5e13, 1e15] W/cm^2, CEP U[0, 2pi), 2-6 cycles at 800 nm."""
    intensity = 10.0 ** rng.uniform(np.log10(5e13), np.log10(1e15))
    return float(np.sqrt(intensity / 3.50944e16)), float(rng.uniform(0.0, 2.0)), int(rng.integers(2, 7))


def request_shapes(kind, rng, sides=(512, 768, 1024, 1536, 2048), n1=4096, n3=256,
                   precision=T.FP32_MIXED, omega=0.057, cycles=None):
    """Grid, propagation and field of one request of the mixed stream."""
    e0, cep, ncyc = _laser(rng)
    if cycles is not None:
        ncyc = cycles
    if kind == "1d":      # single electron, soft-core Coulomb, length gauge, Strang, dt 0.05
        d = T.grid_desc(1, n1, 0.1, precision=precision, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1,
                        Nsoft=1, absorber=T.ABS_COS18, nCAP=int(round(0.1 * n1)), absorber_tau=0.05)
        p = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=[(T.BORDER_BOX, int(0.35 * n1))])
        f = [T.field_desc(T.AXIS_X, e0, omega, ncyc, phase_in_pi=cep)]
    elif kind == "2d":    # two-electron reduced model, velocity gauge, Strang, dt 0.02
        n = int(rng.choice(sides))
        d = T.grid_desc(2, n, 0.2, precision=precision, potential=T.POT_EBERLY, Ncharge=2, Echarge=-1,
                        Nsoft=0.5, Esoft=0.33, absorber=T.ABS_COS18, nCAP=int(round(0.1 * n)),
                        absorber_tau=0.02)
        p = T.prop_desc(T.MODE_RE, 1, 0.02, gauge=T.GAUGE_VELOCITY, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=[(b, 0) for b in T.ZOA_FLUX_2D])
        f = [T.field_desc(T.AXIS_XY, e0, omega, ncyc, phase_in_pi=cep)]
    else:                 # three-electron restricted geometry, length gauge, Strang, dt 0.05
        d = T.grid_desc(3, n3, 0.3, precision=precision, potential=T.POT_ECKHARDT_SACHA, Ncharge=3,
                        Echarge=-1, Nsoft=1.0, Esoft=1.0, absorber=T.ABS_CAP, nCAP=max(2, n3 // 16))
        p = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH, outputs=T.OUT_NORM | T.OUT_FLUX,
                        borders=[(T.BORDER_BOX, int(0.4 * n3))])
        f = [T.field_desc(T.AXIS_XYZ, e0, omega, ncyc, phase_in_pi=cep)]
    return d, p, f


def bound_state(desc):
    """A normalised bound-like start state of the grid: product of soft-core 1s-shaped orbitals
    (same for every request of one shape, as a cached ground state would be)."""
    dim = desc.dim
    parts = []
    for a in range(dim):
        n, dx = int(desc.n[a]), float(desc.dx[a])
        x = -0.5 * dx * (n - 1) + dx * np.arange(n)
        parts.append(np.exp(-np.sqrt(x * x + 1.0)))
    psi = parts[0]
    for p in parts[1:]:
        psi = np.multiply.outer(psi, p)
    psi = psi.ravel().astype(np.complex128)
    dv = float(np.prod([desc.dx[a] for a in range(dim)]))
    return psi / np.sqrt(np.sum(np.abs(psi) ** 2) * dv)


def poisson_stream(count, rate=4.0, seed=5, mix=(0.60, 0.35, 0.05), verify_every=10, **shape_kw):
    """Seeded Poisson arrivals of mixed requests; every `verify_every`-th FP32 request gets
    an FP64 twin (same arrival) whose result checks it. The kinds hold the mix exactly (rounded
    counts in a seeded random order), so a short stream still carries its share of 3D solves."""
    rng = np.random.default_rng(seed)
    counts = [int(round(count * m)) for m in mix]
    counts[int(np.argmax(counts))] += count - sum(counts)
    kinds = np.repeat(np.array(["1d", "2d", "3d"]), counts)
    rng.shuffle(kinds)
    t = 0.0
    out = []
    cache = {}
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        kind = kinds[i]
        d, p, f = request_shapes(kind, rng, **shape_kw)
        key = (d.dim, tuple(d.n), d.dx[0])
        if key not in cache:
            cache[key] = bound_state(d)
        out.append(Request(len(out), t, str(kind), d, p, f, cache[key]))
        if verify_every and i % verify_every == 0:
            d64 = T.GridDesc.from_buffer_copy(d)
            d64.precision = T.FP64
            out.append(Request(len(out), t, str(kind), d64, p, f, cache[key], shadow_of=out[-1].rid))
    return out


# ------------------------------------------------------------------------------ the server
def admissible(pending, now, used, budget):
    """How many requests at the head of the arrival-ordered queue are admitted now: those that
    have arrived and whose footprints fit the budget, in order (a request that does not fit
    blocks the ones behind it, so large solves are not starved)."""
    k, total = 0, used
    for req in pending:
        if req.arrival > now:
            break
        need = footprint_bytes(req.desc, req.pdesc)
        if need > budget:
            raise ValueError(f"request {req.rid} needs {need} B, more than the device budget {budget} B")
        if total + need > budget:
            break
        total += need
        k += 1
    return k


class _Active:
    __slots__ = ("req", "ctx", "gset", "prop", "admitted", "remaining", "bytes", "quantum")


class Server:
    """Admits requests by footprint and advances all admitted solves tick by tick."""

    def __init__(self, streams=4, hbm_bytes=HBM_BYTES, reserve=0.05, pts_per_tick=4e8, device=0,
                 tuning=None):
        from . import hip
        self.hip = hip
        self.ctxs = [hip.Context(device) for _ in range(streams)]
        self.budget = int(hbm_bytes * (1.0 - reserve))
        self.pts_per_tick = float(pts_per_tick)
        self.tuning = list(tuning or [])
        for c in self.ctxs:
            for (pr, row, n, m, v, _ms) in self.tuning:
                c.tune_set(pr, row, n, m, v)
        self.ticks = 0

    def close(self):
        for c in self.ctxs:
            c.close()
        self.ctxs = []

    def _start(self, req, ctx, now):
        a = _Active()
        a.req, a.ctx, a.admitted = req, ctx, now
        a.gset = self.hip.GridSet(ctx, req.desc)
        a.gset[0].set_psi(req.psi0)
        for r in range(1, len(a.gset)):
            a.gset[r].set_psi(np.zeros_like(req.psi0))
        a.prop = self.hip.Propagator(a.gset, req.pdesc, req.fields)
        a.prop.start()
        a.remaining = a.prop.max_steps + 2          # RE: max_steps + 2 loop bodies
        m = int(np.prod([req.desc.n[i] for i in range(req.desc.dim)]))
        a.quantum = max(1, int(self.pts_per_tick / m))
        a.bytes = footprint_bytes(req.desc, req.pdesc)
        return a

    def _harvest(self, a, now):
        a.prop.finish()
        rec = a.prop.records()
        fi = a.prop.flux_integral()
        res = Result(a.req.rid, a.req.kind, int(np.prod([a.req.desc.n[i] for i in range(a.req.desc.dim)])),
                     int(a.prop.step_count), a.req.arrival, a.admitted, now, rec, fi,
                     float(rec[-1, T.REC_NORM]), a.req.shadow_of)
        a.prop.close()
        a.gset.close()
        return res

    def run(self, requests, log=None):
        """Serve `requests` (arrival times relative to now); return results in completion order.

        Each stream is refilled as soon as it drains (non-blocking query), so a stream of short
        solves does not wait for a stream of long ones; small solves (<= SMALL points) have
        streams of their own (half of them, with two or more streams)."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        S = len(self.ctxs)
        n_small = S // 2 if S >= 2 else S
        pools = (list(range(n_small)), list(range(n_small, S)) if S >= 2 else list(range(S)))
        active = [[] for _ in range(S)]
        results = []
        used = 0
        load = [0] * S
        t0 = time.perf_counter()
        while pending or any(active):
            now = time.perf_counter() - t0
            # admission: first come first served, by footprint against the HBM budget
            for _ in range(admissible(pending, now, used, self.budget)):
                req = pending.pop(0)
                m = int(np.prod([req.desc.n[i] for i in range(req.desc.dim)]))
                pool = pools[0] if m <= SMALL else pools[1]
                k = min(pool, key=lambda j: load[j])   # the stream holding the fewest bytes
                a = self._start(req, self.ctxs[k], now)
                load[k] += a.bytes
                used += a.bytes
                active[k].append(a)
            progressed = False
            for k in range(S):
                if not active[k] or not self.ctxs[k].idle():
                    continue
                # stream k drained: hand back its finished solves, give the rest their quantum
                keep = []
                for a in active[k]:
                    if a.remaining <= 1:
                        results.append(self._harvest(a, time.perf_counter() - t0))
                        load[k] -= a.bytes
                        used -= a.bytes
                        if log:
                            r = results[-1]
                            log(f"request {r.rid} ({r.kind}, {r.points} pts, {r.steps} steps) done: "
                                f"latency {r.latency:.3f} s, yield {r.ionization_yield:.6e}")
                    else:
                        keep.append(a)
                active[k] = keep
                for a in keep:
                    steps = min(a.quantum, a.remaining - 1)  # the last loop body runs in finish()
                    done = a.prop.run_steps(steps)
                    a.remaining = a.remaining - done if done == steps else 1
                self.ticks += 1
                progressed = True
            if not progressed:
                time.sleep(0.0005 if any(active) else max(0.0, min(0.05, pending[0].arrival - now)))
        return results


def summarize(results, wall):
    """Throughput and time-to-result percentiles of the served (non-shadow) requests, plus the
    FP32-mixed vs FP64 error of every verification pair."""
    served = [r for r in results if r.shadow_of < 0]
    lat = np.array([r.latency for r in served])
    pts = sum(r.points * r.steps for r in results)
    by_rid = {r.rid: r for r in results}
    checks = []
    for r in results:
        if r.shadow_of >= 0 and r.shadow_of in by_rid:
            a = by_rid[r.shadow_of]
            y64, y32 = r.ionization_yield, a.ionization_yield
            fl = np.abs(a.flux_integral - r.flux_integral) / np.maximum(np.abs(r.flux_integral), 1e-300)
            checks.append({"rid": a.rid, "kind": a.kind, "points": a.points, "yield_fp64": y64,
                           "yield_rel_err": abs(y32 - y64) / max(abs(y64), 1e-300),
                           "norm_abs_err": abs(a.final_norm - r.final_norm),
                           "flux_integral_max_rel_err": float(fl.max()) if fl.size else 0.0})
    return {
        "requests": len(served), "verification_pairs": len(checks),
        "throughput_grid_point_steps_per_s": pts / wall if wall > 0 else 0.0,
        "requests_per_s": len(served) / wall if wall > 0 else 0.0,
        "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
        "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
        "max_yield_rel_err": max((c["yield_rel_err"] for c in checks), default=None),
        "checks": checks,
    }
