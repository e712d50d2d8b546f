"""The serving loop (qsf_amd/serve.py): admission by footprint, the request stream, and on the
GPU that every result is the same alone and in the mix."""
import numpy as np
import pytest

from qsf_amd import qsf_types as T
from qsf_amd import serve


def small_stream(count=8, seed=11, verify_every=4, precision=T.FP32_MIXED):
    return serve.poisson_stream(count, rate=50.0, seed=seed, mix=(0.4, 0.4, 0.2), verify_every=verify_every,
                                sides=(64, 96, 128), n1=512, n3=32, omega=0.5, cycles=1, precision=precision)


def test_footprint_covers_state_and_scales():
    d, p, _ = serve.request_shapes("2d", np.random.default_rng(0), sides=(1024,))
    fp = serve.footprint_bytes(d, p)
    assert fp > 1024 * 1024 * 8 * 2            # state + work copy in FP32
    d64 = T.GridDesc.from_buffer_copy(d)
    d64.precision = T.FP64
    assert serve.footprint_bytes(d64, p) > fp
    d3, p3, _ = serve.request_shapes("3d", np.random.default_rng(0))
    assert serve.footprint_bytes(d3, p3) > 256 ** 3 * 8 * 2


def test_stream_is_seeded_and_mixed():
    a = serve.poisson_stream(200, seed=5)
    b = serve.poisson_stream(200, seed=5)
    assert [(r.kind, r.arrival, r.desc.n[0]) for r in a] == [(r.kind, r.arrival, r.desc.n[0]) for r in b]
    served = [r for r in a if r.shadow_of < 0]
    frac = {k: sum(r.kind == k for r in served) / len(served) for k in ("1d", "2d", "3d")}
    assert abs(frac["1d"] - 0.60) < 0.1 and abs(frac["2d"] - 0.35) < 0.1 and frac["3d"] < 0.12
    gaps = np.diff([r.arrival for r in served])
    assert 0.15 < gaps.mean() < 0.35                 # mean 1/4 s
    twins = [r for r in a if r.shadow_of >= 0]
    assert twins and all(r.desc.precision == T.FP64 for r in twins)
    assert all(a[r.shadow_of].desc.precision == T.FP32_MIXED for r in twins)


def test_admission_is_fifo_by_footprint():
    reqs = small_stream(6, verify_every=0)
    for i, r in enumerate(reqs):
        r.arrival = 0.1 * i
    need = [serve.footprint_bytes(r.desc, r.pdesc) for r in reqs]
    # everything arrived and fits
    assert serve.admissible(reqs, 1.0, 0, sum(need)) == len(reqs)
    # only the first three have arrived
    assert serve.admissible(reqs, 0.25, 0, sum(need)) == 3
    # the budget holds the first two only: the third waits and blocks those behind it
    assert serve.admissible(reqs, 1.0, 0, need[0] + need[1]) == 2
    assert serve.admissible(reqs[1:], 1.0, need[0], need[0] + need[1]) == 1   # first one resident
    with pytest.raises(ValueError):
        serve.admissible(reqs, 1.0, 0, min(need) - 1)


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix():
    reqs = small_stream()
    print(f"[case] serving mix of {len(reqs)} small requests", flush=True)
    srv = serve.Server(streams=4)
    mixed = {r.rid: r for r in srv.run(reqs)}
    srv.close()
    assert len(mixed) == len(reqs)
    for req in reqs:
        req1 = serve.Request(**{**req.__dict__, "arrival": 0.0})
        srv = serve.Server(streams=1)
        alone = srv.run([req1])[0]
        srv.close()
        m = mixed[req.rid]
        assert m.steps == alone.steps
        # independent state, tables and fixed-order reductions: identical bits (columns of
        # observables the request did not ask for hold NaN in both)
        assert np.array_equal(m.records, alone.records, equal_nan=True), f"request {req.rid} ({req.kind})"
        assert np.array_equal(m.flux_integral, alone.flux_integral, equal_nan=True)


@pytest.mark.gpu
def test_admission_under_a_small_budget_serves_everything():
    reqs = small_stream(6, verify_every=0)
    biggest = max(serve.footprint_bytes(r.desc, r.pdesc) for r in reqs)
    srv = serve.Server(streams=2, hbm_bytes=int(biggest / 0.95) + 1)
    res = srv.run(reqs)
    srv.close()
    assert sorted(r.rid for r in res) == [r.rid for r in reqs]
    # one at a time: a request is admitted only after the previous one left
    res.sort(key=lambda r: r.admitted)
    for a, b in zip(res, res[1:]):
        assert b.admitted >= a.finished - 1e-9 or serve.footprint_bytes(reqs[a.rid].desc, reqs[a.rid].pdesc) + \
            serve.footprint_bytes(reqs[b.rid].desc, reqs[b.rid].pdesc) <= int(biggest / 0.95 * 0.95) + 1


@pytest.mark.gpu
@pytest.mark.parametrize("precision,bar", [(T.FP32_MIXED, 1e-5), (T.FP32_STATE, 1e-6)])
def test_fp32_requests_match_their_fp64_twins(precision, bar):
    # FP32-mixed butterflies, or the FP32 state under FP64 arithmetic (closer to FP64)
    reqs = small_stream(8, verify_every=1, precision=precision)
    srv = serve.Server(streams=4)
    res = srv.run(reqs)
    srv.close()
    summ = serve.summarize(res, 1.0)
    assert summ["verification_pairs"] == 8
    for c in summ["checks"]:
        print(f"[case] precision {precision} request {c['rid']} ({c['kind']}): norm {c['norm_abs_err']:.2e}, "
              f"flux integrals {c['flux_integral_max_rel_err']:.2e}", flush=True)
        assert c["norm_abs_err"] < bar, c
