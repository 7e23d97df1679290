"""Every workload of the benchmark definition at its stated size, on the GPU.

* The 200k x 1M skewed LP (Zipf column degrees, 7.97M nonzeros) in FP64 at
  relative KKT 1e-6, and in FP32 + iterative refinement with the objective
  within 1e-8 of the closed form and the per-round record checked.
* The 4096 uneven LPs of the batch in one launch sequence, with the bench's
  settings: the OPTIMAL fraction, and every OPTIMAL objective within 1e-8 of
  its closed form.
* Multi-period blending LPs inside the served workload's ranges (K >= 20,
  P >= 5, T in the hundreds) and a warm-started 1 % resolve, on the FP64 team
  server and on FP32 + refinement with an FP64 fallback, against HiGHS
  objectives (``tests/golden/make_blending_vectors.py --spec``).

Reference: the solve they run is ``src/sovopt/lp/pdlp.py:238-494``.
"""

import json
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(__file__))

from golden.make_blending_vectors import build as build_blend  # noqa: E402
from sovlp.core.problem import Status  # noqa: E402
from sovlp.lp.pdlp import PDLPParams, solve_pdlp  # noqa: E402

SPEC = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "blending_spec_highs.json")))


def _rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.fixture(scope="module")
def large_lp():
    from sovlp.workloads import skewed_lp
    t = time.perf_counter()
    ko = skewed_lp(1)
    print(f"[gpu] 200k x 1M LP built in {time.perf_counter() - t:.1f} s "
          f"({ko.problem.A.nnz} nonzeros)", flush=True)
    assert (ko.problem.m, ko.problem.n) == (200_000, 1_000_000)
    assert 7_900_000 <= ko.problem.A.nnz <= 8_000_000
    return ko


def test_large_skewed_lp_fp64(large_lp):
    ko = large_lp
    par = PDLPParams(device="gpu", eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6, precision=64)
    print("[gpu] 200k x 1M FP64 solve", flush=True)
    s = solve_pdlp(ko.problem, par)
    err = _rel(s.objective, ko.f_star)
    print(f"[gpu] FP64: {s.status.name} it {s.iterations} restarts {s.info.restarts} "
          f"device {s.info.device_seconds:.2f} s W {s.info.workgroups} rel err {err:.1e}",
          flush=True)
    assert s.status == Status.OPTIMAL
    assert err <= 1e-6
    rv, bv, _ = ko.problem.violation(s.x)
    assert rv <= 1e-6 and bv <= 1e-9


def test_large_skewed_lp_fp32_refined(large_lp):
    ko = large_lp
    par = PDLPParams(device="gpu", eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6, precision=32)
    print("[gpu] 200k x 1M FP32 + refinement solve", flush=True)
    s = solve_pdlp(ko.problem, par)
    err = _rel(s.objective, ko.f_star)
    rows = s.info.refinement
    print(f"[gpu] FP32: {s.status.name} it {s.iterations} device {s.info.device_seconds:.2f} s "
          f"rel err {err:.1e}; rounds (it, pres, gap, err): "
          f"{[(r[0], f'{r[1]:.1e}', f'{r[3]:.1e}', f'{_rel(r[4], ko.f_star):.1e}') for r in rows]}",
          flush=True)
    assert s.status == Status.OPTIMAL
    assert err <= 1e-8
    # the record: one row per evaluated round, iterations increasing, the
    # primal residual and objective error falling from the first round
    assert len(rows) >= 2
    assert all(b[0] > a[0] for a, b in zip(rows, rows[1:]))
    assert rows[-1][1] < rows[0][1]
    assert _rel(rows[-1][4], ko.f_star) < _rel(rows[0][4], ko.f_star)
    rv, bv, _ = ko.problem.violation(s.x)
    assert rv <= 1e-6 and bv <= 1e-6


def test_full_uneven_batch_with_bench_settings():
    from sovlp.lp.autotune import Autotuner, packaged_table
    from sovlp.lp.batch import BatchEngine
    from sovlp.workloads import uneven_batch
    kos = uneven_batch(4096, base_seed=1000)
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6, device="gpu",
                     refine_stall_open=1024)
    tuner = Autotuner.load(packaged_table(32))
    print("[gpu] 4096-LP batch, FP32 + refinement, bench settings", flush=True)
    eng = BatchEngine([k.problem for k in kos], par, precision=32, tuner=tuner)
    try:
        t = time.perf_counter()
        graphs = eng.run()
        wall = time.perf_counter() - t
        sols, lat = eng.solutions(keep_outputs=True)
    finally:
        eng.close()
    opt = [i for i, s in enumerate(sols) if s.status == Status.OPTIMAL]
    errs = np.array([_rel(sols[i].objective, kos[i].f_star) for i in opt])
    print(f"[gpu] batch: {len(opt)} of 4096 OPTIMAL in {wall:.2f} s ({graphs} replays), "
          f"max rel err {errs.max():.1e}, p99 {np.percentile(errs, 99):.1e}, latency p50 "
          f"{np.percentile(lat, 50):.3f} s p99 {np.percentile(lat, 99):.3f} s", flush=True)
    assert len(opt) >= int(0.98 * 4096)
    assert errs.max() <= 1e-8
    for i in opt[:: 64]:
        rv, bv, _ = kos[i].problem.violation(sols[i].x)
        assert rv <= 1e-6 * (1 + np.abs(kos[i].problem.row_ub).max()) and bv <= 1e-6
    # every LP that did not finish ends with a reported status, not a hang
    assert all(s.status in (Status.OPTIMAL, Status.NUMERICAL, Status.ITERATION_LIMIT)
               for s in sols)


def _spec_requests():
    from sovlp.lp.serve import Request
    reqs = []
    for i, c in enumerate(SPEC["cases"]):
        p = build_blend(c)
        resolve = c["perturb_seed"] is not None
        base = next(j for j, d in enumerate(SPEC["cases"])
                    if d["seed"] == c["seed"] and d["perturb_seed"] is None)
        reqs.append(Request(i, p, arrival=0.05 * i, model=base, resolve=resolve))
    return reqs


@pytest.mark.parametrize("mode", ["fp64_team", "fp32_with_fp64_fallback"])
def test_spec_sized_blending_lps_served(mode):
    """The new models are served side by side; the 1 % resolve arrives after
    its base model's result and is warm-started from it."""
    from sovlp.lp.serve import DeviceServer, HbmAdmission, ServeLoop, summarize
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6, device="gpu")
    reqs = _spec_requests()
    print(f"[gpu] spec-sized blending LPs, {mode}: "
          f"{[(r.problem.m, r.problem.n, r.problem.A.nnz) for r in reqs]}", flush=True)
    if mode == "fp64_team":
        srv = DeviceServer(len(reqs), par, precision=64, team=8, phases_per_launch=64)
        fb = None
    else:
        srv = DeviceServer(len(reqs), par, precision=32, team=8, phases_per_launch=64)
        fb = DeviceServer(len(reqs), par, precision=64, team=8, phases_per_launch=64)
    try:
        loop = ServeLoop(srv, HbmAdmission(1 << 34), fallback=fb)
        t = time.perf_counter()
        out = loop.run([r for r in reqs if not r.resolve])
        out += loop.run([r for r in reqs if r.resolve])
        wall = time.perf_counter() - t
    finally:
        srv.close()
        if fb is not None:
            fb.close()
    st = summarize(out)
    print(f"[gpu] {mode}: {wall:.1f} s, {st}", flush=True)
    by = {s.key: s for s in out}
    for i, c in enumerate(SPEC["cases"]):
        s = by[i]
        assert s.solution is not None and s.solution.status == Status.OPTIMAL, (c, s)
        refined = mode != "fp64_team" and not s.escalated
        tol = 1e-8 if refined else 1e-6
        err = _rel(s.solution.objective, c["objective"])
        print(f"[gpu]   {c['K']}x{c['P']}x{c['T']} resolve={reqs[i].resolve}: "
              f"it {s.solution.iterations} warm {s.warm} escalated {s.escalated} "
              f"latency {s.latency:.1f} s rel err {err:.1e}", flush=True)
        assert err <= tol
        assert s.warm == reqs[i].resolve
        rv, bv, _ = reqs[i].problem.violation(s.solution.x)
        assert rv <= 1e-6 and bv <= 1e-6
