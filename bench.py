"""Benchmark of the split-operator step on one MI355X.

Default workload (the largest single-GPU configuration, where the FP64 roofline target sits):
the 3D three-electron restricted-geometry model, 512^3, dx=0.3, Eckhardt-Sacha interaction
(Z=3, soft-core 1.0 / 1.0), length gauge, Strang splitting, dt=0.05, 2-cycle 800 nm sin^2 pulse
with E0 drawn from U[0.05, 0.15] (seed 3), CAP absorber on the outer 32 nodes, FLUX<BOX<224>>
(six faces) plus the norm every step. Synthetic initial state (seeded complex Gaussian noise).

One "step" = one loop body of SplitPropagator::run (makeStep + computeEach). value =
grid-point-steps/s in FP64 with the state resident in HBM. The same line carries the FP32-mixed
form of the same workload (with its error against FP64 after the same steps), the batched
FP32-mixed step of 64 uneven 2D solves against the plain FP64 step of the same solves, and the
CPU baseline: the FP64 oracle on the host cores (one whole step of the same workload on all
cores, and a one-core sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from qsf_amd import qsf_types as T  # noqa: E402

HBM_PEAK_GBS = 8000.0


def c2_desc(precision, multigrid=1, n=2048):
    return T.grid_desc(2, n, 0.2, precision=precision, potential=T.POT_EBERLY, Ncharge=2,
                       Echarge=-1, Nsoft=0.5, Esoft=0.33, absorber=T.ABS_COS18,
                       nCAP=int(round(0.1 * n)), absorber_tau=0.02, multigrid=multigrid)


def c2_prop():
    return T.prop_desc(T.MODE_RE, 2, 0.02, gauge=T.GAUGE_VELOCITY,
                       outputs=T.OUT_NORM | T.OUT_FLUX, borders=[(b, 0) for b in T.ZOA_FLUX_2D])


C2_FIELDS = [T.field_desc(T.AXIS_XY, 0.12, 0.057, 3)]


def c1_desc(precision=T.FP64, n=4096):
    return T.grid_desc(1, n, 0.1, precision=precision, potential=T.POT_COULOMB, Ncharge=1, Echarge=-1,
                       Nsoft=1.0, absorber=T.ABS_COS18, nCAP=int(round(0.1 * n)), absorber_tau=0.05)


C1_WORKLOAD = ("1D single electron N=4096 dx=0.1, soft-core Coulomb, cos^(1/8) on 10% of each edge: "
               "imaginary-time ground state (2000 steps, dt=0.05, Normalize + energies every step), "
               "then the real-time Strang run over the 4-cycle 800 nm sin^2 pulse (E0=0.0534, length "
               "gauge, dt=0.05) with norm, energies and flux through |x|=150 a.u. every step")


def run_c1(oracle_too=True):
    """C1: the ground state by imaginary time, then the pulse in real time, each timed whole."""
    from qsf_amd import hip
    x = -0.5 * 0.1 * 4095 + 0.1 * np.arange(4096)
    guess = (np.exp(-np.abs(x)) + 0j) / np.sqrt(np.sum(np.exp(-2 * np.abs(x))) * 0.1)
    im = T.prop_desc(T.MODE_IM, 1, 0.05, max_steps=2000,
                     outputs=T.OUT_NORMALIZE | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_EDIFF)
    re = T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH,
                     outputs=T.OUT_NORM | T.OUT_EPOT | T.OUT_EKIN | T.OUT_ETOT | T.OUT_FLUX,
                     borders=[(T.BORDER_BOX, 1500)])
    fields = [T.field_desc(T.AXIS_X, 0.0534, 0.057, 4)]
    ctx = hip.Context(0)
    gs = hip.GridSet(ctx, c1_desc())
    gs[0].autotune()               # kernel variants of the 4096-point pass, before the timed runs
    gs[0].set_psi(guess)
    t0 = time.perf_counter()
    p = hip.Propagator(gs, im, [T.field_desc(T.AXIS_X, 0.0, 0.057, 4)])
    rim = p.run()
    ctx.synchronize()
    t_im = time.perf_counter() - t0
    p.close()
    t0 = time.perf_counter()
    p = hip.Propagator(gs, re, fields)
    rre = p.run()
    ctx.synchronize()
    t_re = time.perf_counter() - t0
    flux = p.flux_integral()
    p.close()
    gs.close()
    ctx.close()
    out = dict(im_steps=int(rim.shape[0]), im_s=t_im, e0=float(rim[-1, T.REC_ETOT]),
               re_steps=int(rre.shape[0]), re_s=t_re, yield_norm=float(1 - rre[-1, T.REC_NORM]),
               flux_integral=float(flux[0]))
    if oracle_too:
        from oracle import oracle as O
        progress("c1: the same two runs in the oracle (one host thread)")
        O.set_threads(1)
        g = O.GridSet(c1_desc())
        g[0].set_psi(guess)
        t0 = time.perf_counter()
        oi = O.Propagator(g, im, [T.field_desc(T.AXIS_X, 0.0, 0.057, 4)]).run()
        t_oim = time.perf_counter() - t0
        t0 = time.perf_counter()
        po = O.Propagator(g, re, fields)
        orr = po.run()
        t_ore = time.perf_counter() - t0
        out.update(oracle_im_s=t_oim, oracle_re_s=t_ore, oracle_e0=float(oi[-1, T.REC_ETOT]),
                   e0_abs_diff=abs(float(oi[-1, T.REC_ETOT]) - out["e0"]),
                   yield_rel_diff=abs(float(1 - orr[-1, T.REC_NORM]) - out["yield_norm"]) /
                   max(abs(float(1 - orr[-1, T.REC_NORM])), 1e-300),
                   oracle_flux_integral=float(po.flux_integral()[0]) if hasattr(po, "flux_integral") else None)
    return out


def c3_desc(precision, n=512):
    return T.grid_desc(3, n, 0.3, precision=precision, potential=T.POT_ECKHARDT_SACHA, Ncharge=3,
                       Echarge=-1, Nsoft=1.0, Esoft=1.0, absorber=T.ABS_CAP, nCAP=32)


def c3_prop():
    return T.prop_desc(T.MODE_RE, 1, 0.05, gauge=T.GAUGE_LENGTH,
                       outputs=T.OUT_NORM | T.OUT_FLUX, borders=[(T.BORDER_BOX, 224)])


def c3_fields():
    E0 = float(np.random.default_rng(3).uniform(0.05, 0.15))
    return [T.field_desc(T.AXIS_XYZ, E0, 0.057, 2)]


C4_SIDES = (512, 768, 1024, 1536, 2048)


def c4_solves(precision, count=64, seed=4):
    """64 independent 2D two-electron solves: sides U{512..2048}, intensity log-U[5e13, 1e15]
    W/cm^2 (E0 = sqrt(I / 3.50944e16)), CEP U[0, 2pi), 2-6 cycles at 800 nm; the 2048^2
    workload's physics without multigrid (Strang, dt=0.02, velocity gauge, cos^(1/8))."""
    rng = np.random.default_rng(seed)
    sides = rng.choice(C4_SIDES, size=count)
    intensity = 10.0 ** rng.uniform(np.log10(5e13), np.log10(1e15), size=count)
    e0 = np.sqrt(intensity / 3.50944e16)
    cep = rng.uniform(0.0, 2.0, size=count)           # phase in units of pi
    cycles = rng.integers(2, 7, size=count)
    descs = [c2_desc(precision, multigrid=0, n=int(n)) for n in sides]
    fields = [T.field_desc(T.AXIS_XY, float(e), 0.057, int(c), phase_in_pi=float(p))
              for e, c, p in zip(e0, cycles, cep)]
    return descs, fields


C4_WORKLOAD = ("64 independent 2D two-electron solves, sides U{512..2048}, I log-U[5e13,1e15] W/cm^2, "
               "CEP U[0,2pi), 2-6 cycles, Strang dt=0.02, velocity gauge, cos^(1/8), norm + ZOA_FLUX_2D "
               "every step; batched FP32-mixed step replayed as a hipGraph")


def c4_prop(order=1):
    return T.prop_desc(T.MODE_RE, order, 0.02, gauge=T.GAUGE_VELOCITY,
                       outputs=T.OUT_NORM | T.OUT_FLUX, borders=[(b, 0) for b in T.ZOA_FLUX_2D])


def ground_like_state(n, dx=0.2):
    """A bound two-electron-like packet (product of soft-core 1s-shaped orbitals)."""
    x = -0.5 * dx * (n - 1) + dx * np.arange(n)
    f = np.exp(-np.sqrt(x * x + 0.5))
    psi = np.outer(f, f).ravel().astype(np.complex128)
    return psi / np.sqrt(np.sum(np.abs(psi) ** 2) * dx * dx)


def run_c4(steps, warmup):
    """The batched FP32-mixed step of 64 uneven solves (hipGraph replay) beside the plain FP64
    step of the same solves run one after another."""
    from qsf_amd import hip
    ctx = hip.Context(0)
    descs32, fields = c4_solves(T.FP32_MIXED)
    pd = c4_prop()
    # autotune every size once per precision (single grids share the tuning table)
    for n in C4_SIDES:
        for prec in (T.FP32_MIXED, T.FP32_STATE, T.FP64):
            gs = hip.GridSet(ctx, c2_desc(prec, multigrid=0, n=n))
            gs[0].set_psi(ground_like_state(n))
            gs[0].autotune(5)
            gs.close()
    # then each size group's passes on the batched step itself: a single grid's timings need not
    # rank the variants as the batch does, and the variants of the 2048^2 group differ by up to
    # a quarter
    for prec in (T.FP32_MIXED, T.FP32_STATE):
        descs_p, _ = c4_solves(prec)
        hip.autotune_batch(ctx, descs_p, fields, pd, lambda d: ground_like_state(int(d.n[0])))
    progress("autotuned; building the batches")
    pts = sum(int(d.n[0]) * int(d.n[1]) for d in descs32)

    def timed_batch(prec):
        descs, _ = c4_solves(prec)
        b = hip.Batch(ctx, descs, fields, pd)
        for s, d in enumerate(descs):
            b.set_psi(s, ground_like_state(d.n[0]))
        b.start()
        b.run(warmup, use_graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        b.run(steps, use_graph=True)
        ctx.synchronize()
        return b, time.perf_counter() - t0

    b, wall = timed_batch(T.FP32_MIXED)
    batched = pts * steps / wall
    bs, wall_s = timed_batch(T.FP32_STATE)
    # every solve checked against the same solve advanced in FP64 for the same steps
    progress("checking each solve against an FP64 batch")
    descs64, _ = c4_solves(T.FP64)
    b64 = hip.Batch(ctx, descs64, fields, pd)
    for s, d in enumerate(descs64):
        b64.set_psi(s, ground_like_state(d.n[0]))
    b64.start()
    b64.run(warmup + steps, use_graph=True)
    ctx.synchronize()

    # a batch state downloads only after its whole pulse; mid-run, every observed record row
    # (norm, fluxes) and the FP64 flux time integrals are compared at the same step
    def worst_of(bb):
        worst = dict(norm_rel_err=0.0, flux_integral_max_rel_err=0.0)
        for s in range(len(descs32)):
            r32, r64 = bb.records(s), b64.records(s)
            k = min(len(r32), len(r64))
            worst["norm_rel_err"] = max(worst["norm_rel_err"],
                                        float(np.max(np.abs(r32[:k, T.REC_NORM] / r64[:k, T.REC_NORM] - 1))))
            f32, f64 = bb.flux_integral(s), b64.flux_integral(s)
            scale = np.abs(f64).max()
            if scale > 0:
                worst["flux_integral_max_rel_err"] = max(worst["flux_integral_max_rel_err"],
                                                         float(np.abs(f32 - f64).max() / scale))
        return dict(worst, steps=warmup + steps, solves=len(descs32))

    worst, worst_s = worst_of(b), worst_of(bs)
    b64.close()
    bs.close()
    b.close()
    # the plain FP64 step: one propagator per solve, per-size step time
    per_size = {}
    for n in C4_SIDES:
        gs = hip.GridSet(ctx, c2_desc(T.FP64, multigrid=0, n=n))
        gs[0].set_psi(ground_like_state(n))
        p = hip.Propagator(gs, pd, [fields[0]])
        p.start()
        p.run_steps(warmup)
        ctx.synchronize()
        t1 = time.perf_counter()
        p.run_steps(steps)
        ctx.synchronize()
        per_size[n] = (time.perf_counter() - t1) / steps
        p.close()
        gs.close()
    plain_s = sum(per_size[int(d.n[0])] for d in descs32)
    plain = pts / plain_s
    ctx.close()
    return dict(value=batched, ms_per_step=wall / steps * 1e3, pts=pts, plain=plain,
                plain_ms_per_step=plain_s * 1e3, per_size_ms={k: round(v * 1e3, 4) for k, v in per_size.items()},
                sides={int(n): int(sum(1 for d in descs32 if d.n[0] == n)) for n in C4_SIDES},
                vs_fp64=worst,
                fp32_state=dict(value=pts * steps / wall_s, ms_per_step=wall_s / steps * 1e3,
                                speedup_vs_plain_fp64=round(pts * steps / wall_s / plain, 3), vs_fp64=worst_s))


def run_c5(count, streams=4, precision=T.FP32_MIXED):
    """The serving loop on a seeded Poisson stream (mean 4 requests/s) of mixed solves: 60% 1D
    N=4096, 35% 2D sides 512-2048, 5% 3D 256^3, FP32-mixed (or FP32 state with FP64
    arithmetic), every 10th with an FP64 twin."""
    from qsf_amd import hip, serve
    reqs = serve.poisson_stream(count, rate=4.0, seed=5, precision=precision)
    # one autotuning pass per distinct grid shape and precision; every stream applies its table
    ctx = hip.Context(0)
    seen = set()
    for r in reqs:
        key = (r.desc.dim, tuple(r.desc.n), r.desc.precision)
        if key in seen:
            continue
        seen.add(key)
        gs = hip.GridSet(ctx, r.desc)
        gs[0].set_psi(r.psi0)
        gs[0].autotune(3)
        gs.close()
    table = ctx.tune_table()
    ctx.close()
    progress(f"autotuned {len(seen)} shapes; serving {len(reqs)} requests")
    srv = serve.Server(streams=streams, tuning=table)
    t0 = time.perf_counter()
    res = srv.run(reqs, log=progress)
    wall = time.perf_counter() - t0
    srv.close()
    summ = serve.summarize(res, wall)
    summ["wall_s"] = wall
    summ["ticks"] = srv.ticks
    summ["kinds"] = {k: sum(1 for r in reqs if r.kind == k and r.shadow_of < 0) for k in ("1d", "2d", "3d")}
    return summ


def initial_state(m, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=m) + 1j * rng.normal(size=m)) * 1e-3


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_device(desc, pdesc, fields, steps, warmup, profile=True, keep_state=False):
    from qsf_amd import hip
    progress(f"device run: precision {desc.precision}, {steps} steps after {warmup} warm-up")
    ctx = hip.Context(0)
    gset = hip.GridSet(ctx, desc)
    gset[0].set_psi(initial_state(gset[0].m))
    for g in gset.grids:   # pick the fastest compiled kernel variant of every pass (untimed)
        g.autotune(5)
    progress("autotuned; warming up")
    prop = hip.Propagator(gset, pdesc, fields)
    prop.start()
    prop.run_steps(warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.mark()
    done = prop.run_steps(steps)
    ctx.mark()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = ctx.elapsed_ms()
    kernels = {}
    if profile:
        # live per-kernel timing over the same number of steps (HIP events on the stream)
        ctx.profile_begin()
        prop.run_steps(steps)
        kernels = ctx.profile_end()
    else:
        prop.run_steps(steps)   # same step count as a profiled run, so the two can be compared
    rec = prop.records()
    psi = gset[0].psi() if keep_state else None
    m_main = gset[0].m
    m_all = sum(g.m for g in gset.grids)
    prop.close()
    gset.close()
    ctx.close()
    return dict(done=done, wall=wall, dev_ms=dev_ms, m_main=m_main, m_all=m_all, kernels=kernels,
                records=rec, psi=psi)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summaries
    (FETCH_SIZE and WRITE_SIZE passes scaled by a plain copy's readings), or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_pmc_traffic.txt"))):
        take = False
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                take = True
                continue
            parts = line.split()
            if take and len(parts) == 2 and parts[0] == kernel:
                return int(parts[1]), os.path.basename(path)
    return None, None


def roofline_of(kernels):
    if not kernels:
        return None
    name, (k, ms, by) = max(kernels.items(), key=lambda kv: kv[1][1])
    avg_ms = ms / k
    achieved = by / (avg_ms * 1e-3) / 1e9
    traffic, source = measured_traffic(name)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": source,
            "kernel": name,
            "avg_launch_us": round(avg_ms * 1e3, 2), "algorithmic_bytes_per_launch": int(by),
            "launches": int(k)}


def copy_calibration(steps):
    """A plain device-to-device copy of a known size (256 MiB). Its FETCH_SIZE / WRITE_SIZE
    readings under rocprofv3 --pmc give the scale that turns those counters into HBM bytes."""
    import torch
    nbytes = 256 << 20
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    a.normal_()
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return {"metric": "copy GB/s", "value": round(2 * nbytes / (ms * 1e-3) / 1e9, 1), "unit": "GB/s",
            "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
            "ms_per_launch": round(ms, 4), "launches": steps + 3}


def cpu_baseline(desc, pdesc, fields, budget_s=20.0, threads=0):
    """The FP64 oracle on whole steps of `desc`, at least one, then until budget_s / 2 (at most
    50); `threads` OpenMP threads for its independent loops (0: the runtime default, which is
    OMP_NUM_THREADS where set)."""
    from oracle import oracle as O
    used = O.set_threads(threads)
    progress(f"cpu baseline: oracle on {used} thread(s), grid {list(desc.n)[:desc.dim]}")
    gs = O.GridSet(desc)
    gs[0].set_psi(initial_state(gs[0].m))
    prop = O.Propagator(gs, pdesc, fields)
    prop.start()
    t0 = time.perf_counter()
    steps = 0
    while steps < 1 or (time.perf_counter() - t0 < budget_s / 2 and steps < 50):
        prop.step()
        steps += 1
    dt = time.perf_counter() - t0
    return dict(value=round(gs[0].m * steps / dt, 1), cores=used, steps=steps, seconds=round(dt, 2),
                m=gs[0].m, regions=len(gs))


def host_cores():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"nproc": os.cpu_count(), "available": avail, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline_c3(pdesc, fields):
    """All threads on one whole step of the 512^3 workload, and one thread on one whole step of
    the same physics at 256^3 (a 512^3 step takes about 90 s on one core)."""
    full = cpu_baseline(c3_desc(T.FP64), pdesc, fields, budget_s=0.0, threads=0)
    one = cpu_baseline(c3_desc(T.FP64, n=256), pdesc, fields, budget_s=0.0, threads=1)
    return {"value": full["value"], "unit": "grid-point-steps/s", "cores": full["cores"], "kind": "port",
            "sample": f"{full['steps']} whole step(s) of the 512^3 workload (observables every step) "
                      f"on {full['cores']} host threads in {full['seconds']} s",
            "host": host_cores(),
            "single_core": {"value": one["value"], "unit": "grid-point-steps/s", "cores": 1,
                            "sample": f"{one['steps']} whole step(s) of the same physics at 256^3 on one "
                                      f"host thread in {one['seconds']} s"}}


def cpu_baseline_c2(desc, pdesc, fields):
    full = cpu_baseline(desc, pdesc, fields, threads=0)
    one = cpu_baseline(desc, pdesc, fields, threads=1)
    return {"value": full["value"], "unit": "grid-point-steps/s", "cores": full["cores"], "kind": "port",
            "sample": f"{full['steps']} whole step(s) of the same workload (all {full['regions']} region "
                      f"grids, observables every step) on {full['cores']} host threads in {full['seconds']} s",
            "host": host_cores(),
            "single_core": {"value": one["value"], "unit": "grid-point-steps/s", "cores": 1,
                            "sample": f"{one['steps']} whole step(s) on one host thread in {one['seconds']} s"}}


FP32_BAR = 1e-6


def fp32_vs_fp64(m, r, nb):
    """Errors of an FP32-state run `m` (FP32-mixed or FP32 state with FP64 arithmetic) against
    the FP64 run `r` after the same steps."""
    a, b = m["records"], r["records"]
    fa = a[:, T.REC_FLUX0:T.REC_FLUX0 + nb]
    fb = b[:, T.REC_FLUX0:T.REC_FLUX0 + nb]
    errs = {
        "steps": int(b[-1, T.REC_STEP]),
        "norm_rel_err": float(np.max(np.abs(a[:, T.REC_NORM] / b[:, T.REC_NORM] - 1))),
        "flux_max_rel_err": float(np.abs(fa - fb).max() / max(np.abs(fb).max(), 1e-300)),
        "psi_l2_rel_err": float(np.linalg.norm(m["psi"] - r["psi"]) / np.linalg.norm(r["psi"]))}
    errs["meets_1e-6_bar"] = all(errs[k] <= FP32_BAR for k in ("norm_rel_err", "flux_max_rel_err", "psi_l2_rel_err"))
    return errs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "c5", "copy"], default="c3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="c3: leave out the batched 2D comparison")
    ap.add_argument("--requests", type=int, default=40, help="c5: number of requests in the stream")
    ap.add_argument("--precision", choices=["fp32_mixed", "fp32_state"], default="fp32_mixed",
                    help="c5: arithmetic of the served requests")
    ap.add_argument("--no-variants", action="store_true", help="leave out the FP32-mixed run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every solve lives whole on one MI355X; --gpus {args.gpus} "
              "is not supported in this build", file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy":
        print(json.dumps(copy_calibration(args.steps)))
        return
    if args.workload == "c1":
        r = run_c1(oracle_too=not args.no_cpu)
        steps = r["im_steps"] + r["re_steps"]
        wall = r["im_s"] + r["re_s"]
        out = {"metric": "grid-point-steps/s", "value": round(4096 * steps / wall, 1),
               "unit": "grid-point-steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
               "ms_per_step": round(wall / steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "c128 (f64)", "data": "synthetic",
               "config": {"workload": C1_WORKLOAD}, "c1": r}
        print(json.dumps(out))
        return
    if args.workload == "c5":
        prec = T.FP32_STATE if args.precision == "fp32_state" else T.FP32_MIXED
        r = run_c5(args.requests, precision=prec)
        out = {"metric": "grid-point-steps/s", "value": round(r["throughput_grid_point_steps_per_s"], 1),
               "unit": "grid-point-steps/s", "n_gpus": 1, "steps": None, "warmup": 0,
               "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": ("c64 state, f64 arithmetic" if prec == T.FP32_STATE else
                         "c64 (f32, f64 phases and reductions)") + "; FP64 twins for verification",
               "data": "synthetic",
               "config": {"workload": f"serving: {args.requests} requests, Poisson 4/s (seed 5), mix 60% 1D "
                                      "N=4096 / 35% 2D 512-2048 / 5% 3D 256^3, full pulses, continuous "
                                      "batching over 4 streams, admission by HBM footprint (288 GB)",
                          "requests_per_kind": r["kinds"]},
               "p50_s": r["p50_s"], "p99_s": r["p99_s"], "requests_per_s": r["requests_per_s"],
               "wall_s": r["wall_s"], "ticks": r["ticks"],
               "verification": {"pairs": r["verification_pairs"], "max_yield_rel_err": r["max_yield_rel_err"],
                                "checks": r["checks"]}}
        print(json.dumps(out))
        return
    if args.workload == "c4":
        r = run_c4(args.steps, args.warmup)
        out = {"metric": "grid-point-steps/s", "value": round(r["value"], 1), "unit": "grid-point-steps/s",
               "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
               "ms_per_step": round(r["ms_per_step"], 4), "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "c64 (f32, f64 phases and reductions)", "data": "synthetic",
               "config": {"workload": C4_WORKLOAD, "grid_points_per_step": r["pts"], "solves_per_side": r["sides"]},
               "plain_fp64": {"value": round(r["plain"], 1), "ms_per_step": round(r["plain_ms_per_step"], 4),
                              "per_size_ms_per_step": r["per_size_ms"]},
               "speedup_vs_plain_fp64": round(r["value"] / r["plain"], 3),
               "vs_fp64_same_steps": r["vs_fp64"],
               "fp32_state": r["fp32_state"]}
        print(json.dumps(out))
        return
    if args.workload == "c2":
        desc, pdesc, fields = c2_desc(T.FP64), c2_prop(), C2_FIELDS
        workload = ("2D two-electron reduced model 2048x2048 dx=0.2, Eberly, velocity gauge, "
                    "MultiProductSplit order 2, dt=0.02, 3-cycle 800 nm sin^2 E0=0.12, cos^(1/8) "
                    "absorber, multigrid NO/X/Y/XY, ZOA_FLUX_2D + norm every step")
        pts_bytes = 64 * 2   # two P substeps per step survive (last MPS chain), 2D FP64
        nb = len(T.ZOA_FLUX_2D)
    else:
        desc, pdesc, fields = c3_desc(T.FP64), c3_prop(), c3_fields()
        workload = ("3D three-electron Eckhardt-Sacha 512^3 dx=0.3, Z=3, length gauge, Strang, "
                    "dt=0.05, 2-cycle 800 nm, CAP 32 nodes, BOX<224> flux + norm every step")
        # three sweeps per step (one per axis: forward transform, P phase, inverse), each a read and
        # a write of 16 B per point; the X halves and the observation ride on the last-axis sweep
        pts_bytes = 96
        nb = 1

    r = run_device(desc, pdesc, fields, args.steps, args.warmup, keep_state=True)
    ms_per_step = r["wall"] / r["done"] * 1e3
    value = r["m_main"] * r["done"] / r["wall"]
    out = {
        "metric": "grid-point-steps/s", "value": round(value, 1), "unit": "grid-point-steps/s",
        "n_gpus": 1, "steps": r["done"], "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128 (f64)",
        "data": "synthetic", "config": {"workload": workload, "grid_points_main": r["m_main"],
                                        "grid_points_all_regions": r["m_all"]},
        "device_ms_per_step": round(r["dev_ms"] / r["done"], 4),
        "step_roofline": {"bytes_per_grid_point_step": pts_bytes,
                          "frac": round(value * pts_bytes / (HBM_PEAK_GBS * 1e9), 4)},
        # the same step priced at 128 B/pt (four sweeps per step, as a 512^3 step with a separate
        # inverse and forward transform around the X factor would move): the 60% mark at 3.58 ms
        **({"step_roofline_128B": {"bytes_per_grid_point_step": 128,
                                   "frac": round(value * 128 / (HBM_PEAK_GBS * 1e9), 4)}}
           if args.workload == "c3" else {}),
        "roofline": roofline_of(r["kernels"]),
        "kernels": {k: {"launches": v[0], "avg_us": round(v[1] / v[0] * 1e3, 2)} for k, v in r["kernels"].items()},
    }
    if args.workload == "c2":
        out["value_all_regions"] = round(r["m_all"] * r["done"] / r["wall"], 1)
    r["kernels"] = None
    if not args.no_variants:
        d32 = c2_desc(T.FP32_MIXED) if args.workload == "c2" else c3_desc(T.FP32_MIXED)
        m = run_device(d32, pdesc, fields, args.steps, args.warmup, profile=False, keep_state=True)
        out["fp32_mixed"] = {"value": round(m["m_main"] * m["done"] / m["wall"], 1),
                             "ms_per_step": round(m["wall"] / m["done"] * 1e3, 4),
                             "vs_fp64_same_steps": fp32_vs_fp64(m, r, nb)}
        del m
        if args.workload == "c3":
            # the FP32 state with FP64 arithmetic (multigrid sets take FP64 or FP32-mixed only)
            m = run_device(c3_desc(T.FP32_STATE), pdesc, fields, args.steps, args.warmup, profile=False,
                           keep_state=True)
            out["fp32_state"] = {"value": round(m["m_main"] * m["done"] / m["wall"], 1),
                                 "ms_per_step": round(m["wall"] / m["done"] * 1e3, 4),
                                 "vs_fp64_same_steps": fp32_vs_fp64(m, r, nb)}
            del m
    r["psi"] = None
    if args.workload == "c3" and not args.no_c4:
        c4 = run_c4(args.steps, args.warmup)
        out["batched_fp32_2d"] = {
            "workload": C4_WORKLOAD, "value": round(c4["value"], 1), "unit": "grid-point-steps/s",
            "ms_per_step": round(c4["ms_per_step"], 4),
            "plain_fp64_value": round(c4["plain"], 1), "plain_fp64_ms_per_step": round(c4["plain_ms_per_step"], 4),
            "speedup_vs_plain_fp64": round(c4["value"] / c4["plain"], 3),
            "vs_fp64_same_steps": c4["vs_fp64"],
            "fp32_state": c4["fp32_state"]}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_c3(pdesc, fields) if args.workload == "c3" else \
            cpu_baseline_c2(desc, pdesc, fields)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
