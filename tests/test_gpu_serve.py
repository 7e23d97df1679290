"""Served PDLP on the GPU (``sovlp_server_*`` through the C ABI): results alone
vs in a mix, warm starts, HBM admission, and a short Poisson stream of blending
LPs checked against HiGHS objectives and closed-form optima."""

import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(__file__))

from golden.make_blending_vectors import build as build_blend  # noqa: E402
from sovlp.core.problem import Status  # noqa: E402
from sovlp.lp.pdlp import PDLPParams  # noqa: E402
from sovlp.lp.serve import (DeviceServer, HbmAdmission, Request, ServeLoop,  # noqa: E402
                            WarmStart, summarize)
from sovlp.workloads import blending_lp, perturbed, uneven_batch_member  # noqa: E402

BLEND = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "blending_highs.json")))


def _params(tol=1e-6, **kw):
    return PDLPParams(eps_abs=tol, eps_rel=tol, feas_cap=tol, device="gpu", **kw)


def _solve_in(srv, waves, steps_between=2, max_steps=200_000):
    """Admit waves of (tag, problem, warm) into the lowest free slots, a wave at
    least ``steps_between`` steps after the previous one; returns {tag: Solution}."""
    out, live = {}, {}
    free = list(range(srv.slots))
    pending = list(waves)
    steps, since = 0, steps_between
    while pending or live:
        if pending and (since >= steps_between or not live):
            wave = pending[0]
            assert len(wave) <= srv.slots
            if len(wave) <= len(free):
                pending.pop(0)
                for tag, p, w in wave:
                    s = free.pop(0)
                    srv.admit(s, p, w)
                    live[s] = tag
                since = 0
        for s in srv.step(1):
            if s in live:
                out[live.pop(s)] = srv.result(s)
                srv.release(s)
                free.append(s)
                free.sort()
        steps += 1
        since += 1
        assert steps < max_steps
    return out


def _same(a, b):
    assert int(a.status) == int(b.status)
    assert a.iterations == b.iterations
    assert a.objective == b.objective or (np.isnan(a.objective) and np.isnan(b.objective))
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert np.array_equal(a.reduced_costs, b.reduced_costs)


@pytest.mark.parametrize("precision,team", [(32, 1), (64, 1), (32, 4), (64, 3)])
def test_served_result_is_identical_alone_and_in_the_mix(precision, team):
    print(f"[gpu] served alone vs mix, precision {precision}, team {team}", flush=True)
    target = build_blend(BLEND["cases"][2])
    others = [uneven_batch_member(5000 + i).problem for i in range(5)] + \
             [build_blend(c) for c in BLEND["cases"][4:8]]
    srv = DeviceServer(4, _params(), precision=precision, phases_per_launch=16, team=team)
    try:
        alone = _solve_in(srv, [[("t", target, None)]])["t"]
        # in the mix: three other LPs take slots 0-2, the target joins slot 3 three
        # launches later while they run, and the rest follow as slots free up
        mix = [[(i, others[i], None) for i in range(3)], [("t", target, None)]] + \
              [[(i, others[i], None)] for i in range(3, len(others))]
        got = _solve_in(srv, mix, steps_between=3)
        mixed = got["t"]
        assert len(got) == len(others) + 1
    finally:
        srv.close()
    assert int(alone.status) == int(Status.OPTIMAL)
    _same(alone, mixed)
    f = BLEND["cases"][2]["objective"]
    assert abs(alone.objective - f) <= 1e-6 * max(1.0, abs(f))


def test_served_blending_lps_and_resolves_match_highs():
    # FP32 + refinement first; an LP it cannot finish goes to an FP64 server,
    # warm-started.  Refined results: objective within 1e-8 of HiGHS; FP64
    # results within the 1e-6 termination criterion.
    cases = BLEND["cases"]
    srv = DeviceServer(8, _params(), precision=32, phases_per_launch=32)
    fb = DeviceServer(4, _params(), precision=64, phases_per_launch=32)
    try:
        reqs = [Request(i, build_blend(c), arrival=0.0) for i, c in enumerate(cases)]
        out = ServeLoop(srv, HbmAdmission(1 << 32), fallback=fb).run(reqs)
    finally:
        srv.close()
        fb.close()
    by = {s.key: s for s in out}
    for i, c in enumerate(cases):
        s = by[i]
        assert int(s.solution.status) == int(Status.OPTIMAL), (c, s.solution.status)
        f = c["objective"]
        tol = 1e-6 if s.escalated else 1e-8
        assert abs(s.solution.objective - f) <= tol * max(1.0, abs(f)), (c, s.solution.objective, f)
    print("escalated:", [s.key for s in out if s.escalated])


def test_warm_start_from_the_optimum_stops_at_the_first_check():
    # FP64 engine, tolerance 1e-6: started from a 1e-9-accurate (x, y) and its
    # primal weight, the first check (k = 64) already meets the criterion.
    # A wrong unit or sign in the warm-start conversion would not.
    p = build_blend(BLEND["cases"][0])
    srv = DeviceServer(2, _params(1e-9), precision=64)
    try:
        cold = _solve_in(srv, [[(0, p, None)]])[0]
        assert int(cold.status) == int(Status.OPTIMAL)
    finally:
        srv.close()
    srv = DeviceServer(2, _params(1e-6), precision=64)
    try:
        cold6 = _solve_in(srv, [[(0, p, None)]])[0]
        w = WarmStart(cold.x, cold.y, cold.info.primal_weight)
        warm = _solve_in(srv, [[(1, p, w)]])[1]
    finally:
        srv.close()
    assert int(warm.status) == int(Status.OPTIMAL)
    assert warm.iterations == 64 < cold6.iterations
    f = BLEND["cases"][0]["objective"]
    assert abs(warm.objective - f) <= 1e-6 * max(1.0, abs(f))


def test_warm_started_resolve_is_correct():
    # a 1 % resolve warm-started from the base model's solution reaches the
    # resolve's HiGHS objective (FP32 + refinement)
    base, res = BLEND["cases"][4], BLEND["cases"][5]
    srv = DeviceServer(2, _params(), precision=32)
    try:
        s0 = _solve_in(srv, [[(0, build_blend(base), None)]])[0]
        w = WarmStart(s0.x, s0.y, s0.info.primal_weight)
        s1 = _solve_in(srv, [[(1, build_blend(res), w)]])[1]
    finally:
        srv.close()
    assert int(s1.status) == int(Status.OPTIMAL)
    f = res["objective"]
    assert abs(s1.objective - f) <= 1e-8 * max(1.0, abs(f))


def test_serve_loop_under_a_tight_hbm_budget():
    probs = [uneven_batch_member(7000 + i) for i in range(10)]
    srv = DeviceServer(8, _params(), precision=32)
    fb = DeviceServer(2, _params(), precision=64)
    try:
        need = max(srv.request_bytes(k.problem, True) for k in probs)
        adm = HbmAdmission(int(2.5 * need))      # at most two LPs resident at once
        reqs = [Request(i, k.problem, arrival=0.0) for i, k in enumerate(probs)]
        out = ServeLoop(srv, adm, fallback=fb).run(reqs)
    finally:
        srv.close()
        fb.close()
    assert adm.peak <= adm.budget and adm.used == 0
    by = {s.key: s for s in out}
    for i, k in enumerate(probs):
        s = by[i].solution
        assert int(s.status) == int(Status.OPTIMAL)
        tol = 1e-6 if by[i].escalated else 1e-8
        assert abs(s.objective - k.f_star) <= tol * max(1.0, abs(k.f_star))


def test_poisson_stream_with_warm_started_resolves():
    rng = np.random.default_rng(3)
    reqs, models = [], {}
    t = 0.0
    for key in range(40):
        t += rng.exponential(1 / 100.0)
        if models and rng.random() < 0.3:
            mk = int(rng.choice(list(models)))
            p = perturbed(models[mk], int(rng.integers(1 << 30)), 0.01)
            reqs.append(Request(key, p, t, model=mk, resolve=True))
        else:
            K, P, T = int(rng.integers(4, 9)), int(rng.integers(2, 4)), int(rng.integers(4, 13))
            p = blending_lp(int(rng.integers(1 << 30)), K=K, P=P, T=T)
            models[key] = p
            reqs.append(Request(key, p, t, model=key))
    srv = DeviceServer(16, _params(), precision=32, phases_per_launch=32)
    fb = DeviceServer(4, _params(), precision=64, phases_per_launch=32)
    try:
        out = ServeLoop(srv, HbmAdmission(1 << 32), fallback=fb).run(reqs)
    finally:
        srv.close()
        fb.close()
    st = summarize(out)
    print(st)
    assert st["finished"] == 40 and st["optimal"] == 40
    assert st["latency_p99_s"] >= st["latency_p50_s"] > 0
    # every served objective agrees with the oracle's FP64 PDLP
    from oracle import oracle
    for s in out[:: 4]:
        p = next(r.problem for r in reqs if r.key == s.key)
        r = oracle.solve(p, eps_abs=1e-9, eps_rel=1e-9)
        assert r.status == 0
        assert abs(s.solution.objective - r.objective) <= 1e-6 * max(1.0, abs(r.objective))


def test_served_history_matches_the_single_lp_engine():
    """A served FP64 LP on one workgroup logs the same check history as the
    single-LP engine on one workgroup (the same arithmetic): one row per check,
    at iterations 64, 128, ..., read through sovlp_server_history."""
    from sovlp.lp.pdlp import solve_pdlp
    p = build_blend(BLEND["cases"][1])
    srv = DeviceServer(2, _params(), precision=64, history_cap=4096)
    try:
        s = _solve_in(srv, [[(0, p, None)]])[0]
    finally:
        srv.close()
    ref = solve_pdlp(p, PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6, device="gpu",
                                   team="block"))
    assert s.iterations == ref.iterations
    h, r = s.info.history, ref.info.history
    assert len(h) == len(r) > 0
    assert [row[0] for row in h] == [64 * (i + 1) for i in range(len(h))]
    for a, b in zip(h, r):
        assert a == b
