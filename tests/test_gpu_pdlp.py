"""End-to-end device PDLP against the oracle, the reference's golden runs and the
closed-form optima (reference tests: tests/test_gpu.py:255-272,
tests/test_solvers.py:50-128)."""

import json
import os

import numpy as np
import pytest

from oracle import oracle
from sovlp.core.problem import ObjSense, Problem, Status
from sovlp.core.sparse import SparseMatrix
from sovlp.lp.pdlp import PDLPParams, solve_pdlp
from sovlp.workloads import known_optimum_lp, small_lp

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_pdlp.json")))


def _problem(case):
    A = SparseMatrix.from_triplets(np.array(case["rows"]), np.array(case["cols"]),
                                   np.array(case["vals"], float), case["m"], case["n"])
    return Problem(A=A, c=np.array(case["c"]), row_lb=np.array(case["rl"]),
                   row_ub=np.array(case["ru"]), col_lb=np.array(case["lo"]),
                   col_ub=np.array(case["hi"]), obj_offset=case["offset"],
                   sense=ObjSense(case["sense"]), name=case["name"])


@pytest.mark.parametrize("case", GOLD["cases"], ids=[c["name"] for c in GOLD["cases"]])
def test_block_team_matches_reference_golden(case):
    print(f"[gpu] block team golden {case['name']}", flush=True)
    p = _problem(case)
    e = case["expected"]
    s = solve_pdlp(p, PDLPParams(device="gpu", team="block", **case["params"]))
    assert s.method == "pdlp[gpu]"
    assert int(s.status) == e["status"]
    assert s.info.scale_obj == e["scale_obj"]
    tol = case["params"].get("eps_rel", 1e-8)
    assert abs(s.objective - e["objective"]) <= 10 * tol * max(1.0, abs(e["objective"]))
    if "f_star" in case:
        assert abs(s.objective - case["f_star"]) <= 10 * tol * max(1.0, abs(case["f_star"]))
    # first history row agrees with the oracle's to rounding
    r = oracle.solve(p, **case["params"])
    if s.info.history and r.history:
        a, b = s.info.history[0], r.history[0]
        assert a[0] == b[0]
        for u, v in zip(a[1:], b[1:]):
            assert abs(u - v) <= 1e-8 * max(1.0, abs(v)), (a, b)
    # the reported point is feasible to the termination tolerance
    rv, bv, _ = p.violation(s.x)
    assert rv <= 1e-6 and bv <= 1e-12
    # textbook duals: reduced costs d = c - A^T y in the original sense
    np.testing.assert_allclose(s.reduced_costs, p.c - p.A.rmatvec(s.y), rtol=1e-9, atol=1e-9)


def test_toy_and_testprob_known_answers():
    toy = _problem(next(c for c in GOLD["cases"] if c["name"] == "toy_max"))
    s = solve_pdlp(toy, PDLPParams(device="gpu"))
    assert s.status == Status.OPTIMAL and abs(s.objective - 11.0) < 1e-6
    np.testing.assert_allclose(s.x, [3.0, 1.0], atol=1e-5)
    q = toy.copy()
    q.c = -q.c
    q.sense = ObjSense.MINIMISE
    assert abs(solve_pdlp(q, PDLPParams(device="gpu")).objective + s.objective) < 1e-5
    tp = _problem(next(c for c in GOLD["cases"] if c["name"] == "testprob"))
    s = solve_pdlp(tp, PDLPParams(device="gpu"))
    assert s.status == Status.OPTIMAL and abs(s.objective - 16.0) < 1e-7
    np.testing.assert_allclose(s.x, [0.0, -1.0, 6.0], atol=1e-6)


def test_small_lp_to_1e8_matches_closed_form_and_oracle():
    print("[gpu] small_lp 500x2000 block team", flush=True)
    ko = small_lp(0)
    s = solve_pdlp(ko.problem, PDLPParams(device="gpu", team="block"))
    r = oracle.solve(ko.problem)
    assert s.status == Status.OPTIMAL and r.status == 0
    assert abs(s.objective - ko.f_star) <= 1e-7 * max(1.0, abs(ko.f_star))
    assert abs(s.objective - r.objective) <= 1e-7 * max(1.0, abs(r.objective))
    rv, bv, _ = ko.problem.violation(s.x)
    assert rv <= 1e-6 and bv <= 1e-12


def test_solves_are_bitwise_repeatable():
    ko = known_optimum_lp(50, 120, 4, seed=21, m_eq=25)
    a = solve_pdlp(ko.problem, PDLPParams(device="gpu", team="block"))
    b = solve_pdlp(ko.problem, PDLPParams(device="gpu", team="block"))
    assert a.iterations == b.iterations
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_limits_are_reported_like_the_reference():
    ko = known_optimum_lp(50, 120, 4, seed=22, m_eq=25)
    s = solve_pdlp(ko.problem, PDLPParams(device="gpu", max_iter=100))
    assert s.status == Status.ITERATION_LIMIT and s.iterations == 100
    assert np.isfinite(s.objective)
    s = solve_pdlp(ko.problem, PDLPParams(device="gpu", time_limit=0.0))
    assert s.status == Status.TIME_LIMIT and s.iterations == 0


def test_grid_team_on_a_medium_lp():
    print("[gpu] grid team medium LP 5000x12000", flush=True)
    ko = known_optimum_lp(5000, 12000, 4, seed=31, m_eq=2500)
    par = dict(eps_abs=1e-6, eps_rel=1e-6, max_iter=100_000, time_limit=120.0)
    g = solve_pdlp(ko.problem, PDLPParams(device="gpu", team="grid", **par))
    print(f"[gpu] grid done: {g.status.name} it={g.iterations} W={g.info.workgroups}", flush=True)
    b = solve_pdlp(ko.problem, PDLPParams(device="gpu", team="block", **par))
    print(f"[gpu] block done: {b.status.name} it={b.iterations}", flush=True)
    assert g.status == Status.OPTIMAL and b.status == Status.OPTIMAL
    assert g.info.team == "grid" and g.info.workgroups > 1
    for s in (g, b):
        assert abs(s.objective - ko.f_star) <= 1e-5 * max(1.0, abs(ko.f_star))
        rv, bv, _ = ko.problem.violation(s.x)
        assert rv <= 1e-6 and bv <= 1e-12


def test_fp32_refined_single_lp_matches_fp64_oracle_to_1e8():
    """One LP through the single-LP engine in FP32 with refinement: the
    objective is within 1e-8 of the FP64 oracle and of the closed form, and
    the per-round errors shrink."""
    print("[gpu] small_lp 500x2000 fp32 + refinement", flush=True)
    ko = small_lp(0)
    par = PDLPParams(device="gpu", precision=32, eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6)
    s = solve_pdlp(ko.problem, par)
    r = oracle.solve(ko.problem)                      # FP64 oracle at 1e-8
    assert s.status == Status.OPTIMAL, s.info.refinement
    for ref_obj in (ko.f_star, r.objective):
        assert abs(s.objective - ref_obj) <= 1e-8 * max(1.0, abs(ref_obj))
    rows = s.info.refinement
    assert len(rows) >= 2
    errs = [abs(row[4] - ko.f_star) / max(1.0, abs(ko.f_star)) for row in rows]
    print("[gpu] refinement rounds:", [(row[0], f"{row[1]:.1e}", f"{e:.1e}")
                                       for row, e in zip(rows, errs)], flush=True)
    assert rows[-1][1] < rows[0][1]
    rv, bv, _ = ko.problem.violation(s.x)
    assert rv <= 1e-6 and bv <= 1e-6


def test_fp32_refined_maximisation_and_free_variables():
    for name, expect in (("toy_max", 11.0), ("testprob", 16.0)):
        p = _problem(next(c for c in GOLD["cases"] if c["name"] == name))
        s = solve_pdlp(p, PDLPParams(device="gpu", precision=32, eps_abs=1e-6, eps_rel=1e-6))
        assert s.status == Status.OPTIMAL and abs(s.objective - expect) < 1e-8 * max(1, expect)
        np.testing.assert_allclose(s.reduced_costs, p.c - p.A.rmatvec(s.y), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("precision", [64, 32])
def test_skewed_lp_over_many_workgroups_matches_closed_form(precision):
    """A 20k x 100k LP with Zipf column degrees (the 200k x 1M construction at a
    tenth of the size): every phase over W workgroups; FP64 and FP32 +
    refinement reach the closed-form optimum."""
    from sovlp.workloads import skewed_lp
    print(f"[gpu] skewed 20k x 100k, precision {precision}", flush=True)
    ko = skewed_lp(1, 20_000, 100_000, 800_000)
    par = PDLPParams(device="gpu", eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6,
                     precision=precision, team="grid")
    s = solve_pdlp(ko.problem, par)
    assert s.status == Status.OPTIMAL
    assert s.info.workgroups > 1
    tol = 1e-8 if precision == 32 else 1e-6
    assert abs(s.objective - ko.f_star) <= tol * max(1.0, abs(ko.f_star))
    rv, bv, _ = ko.problem.violation(s.x)
    assert rv <= 1e-6 and bv <= 1e-9


@pytest.mark.parametrize("precision", [64, 32])
def test_columns_in_locality_order_solve_to_the_same_optimum(precision):
    """``column_order="locality"`` solves the LP with its columns relabelled and
    returns x and the reduced costs in the caller's order: same optimum as the
    given order, x feasible for the original rows, d = c - A^T y."""
    from sovlp.workloads import skewed_lp
    print(f"[gpu] skewed 20k x 100k, locality order, precision {precision}", flush=True)
    ko = skewed_lp(1, 20_000, 100_000, 800_000)
    tol = 1e-8 if precision == 32 else 1e-6
    par = PDLPParams(device="gpu", eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6,
                     precision=precision, team="grid", column_order="locality")
    s = solve_pdlp(ko.problem, par)
    print(f"[gpu] locality done: {s.status.name} it={s.iterations}", flush=True)
    assert s.status == Status.OPTIMAL
    assert abs(s.objective - ko.f_star) <= tol * max(1.0, abs(ko.f_star))
    rv, bv, _ = ko.problem.violation(s.x)
    assert rv <= 1e-6 and bv <= 1e-9
    np.testing.assert_allclose(s.reduced_costs, ko.problem.c - ko.problem.A.rmatvec(s.y),
                               rtol=1e-9, atol=1e-9)
    small = small_lp(0)
    eps = dict(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6) if precision == 32 else {}
    g = solve_pdlp(small.problem, PDLPParams(device="gpu", precision=precision, **eps))
    o = solve_pdlp(small.problem, PDLPParams(device="gpu", precision=precision,
                                             column_order="locality", **eps))
    assert g.status == o.status == Status.OPTIMAL
    assert abs(o.objective - g.objective) <= 1e-7 * max(1.0, abs(g.objective))
