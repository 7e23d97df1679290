"""Ranging of an optimal basis, checked against its definition.

The reference's ranging (``src/sovopt/lp/sensitivity.py:154-332``) answers:
over which interval of ``c_j`` does the basis stay optimal, and over which
interval of a binding row's bound does it stay feasible.  These tests do not
restate its formulas: they take the basis the simplex returns, move one cost
or one bound to points inside and outside the reported interval, and decide
optimality and feasibility of that same basis directly with a dense solve
(``numpy.linalg``) -- dual feasibility of the reduced costs for costs, the
basic values within their bounds for right-hand sides -- and check that the
objective moves by the shadow price times the change inside a bound's range.
"""

import numpy as np
import pytest

from sovlp.core.problem import ObjSense, Problem, Status
from sovlp.core.sparse import SparseMatrix
from sovlp.core.tolerances import INF
from sovlp.lp.basis import AT_LOWER, AT_UPPER, BASIC, FIXED, FREE
from sovlp.lp.crossover import crossover
from sovlp.lp.sensitivity import compute_sensitivity
from sovlp.lp.simplex import SimplexParams, solve_simplex
from sovlp.workloads import known_optimum_lp


def dense(p: Problem) -> np.ndarray:
    A = np.zeros((p.m, p.n))
    for j in range(p.n):
        for q in range(p.A.cp[j], p.A.cp[j + 1]):
            A[p.A.ci[q], j] += p.A.cx[q]
    return A


class BasisCheck:
    """The basis of a status vector over [A | -I], evaluated densely."""

    def __init__(self, p: Problem, st: np.ndarray):
        self.p, self.st = p, st
        self.M = np.hstack([dense(p), -np.eye(p.m)])
        self.basic = np.flatnonzero(st == BASIC)
        self.non = np.flatnonzero(st != BASIC)
        self.B = self.M[:, self.basic]

    def point(self, row_lb, row_ub):
        p, st = self.p, self.st
        lo = np.concatenate([p.col_lb, row_lb])
        up = np.concatenate([p.col_ub, row_ub])
        z = np.where(st == AT_UPPER, up, np.where(st == FREE, 0.0, lo))
        zN = z[self.non]
        z[self.basic] = np.linalg.solve(self.B, -self.M[:, self.non] @ zN)
        return z, lo, up

    def primal_feasible(self, row_lb, row_ub, tol=1e-9):
        z, lo, up = self.point(row_lb, row_ub)
        zb = z[self.basic]
        return bool(np.all(zb >= lo[self.basic] - tol * (1 + np.abs(lo[self.basic]))) and
                    np.all(zb <= up[self.basic] + tol * (1 + np.abs(up[self.basic]))))

    def dual_feasible(self, c_min, tol=1e-9):
        full = np.concatenate([c_min, np.zeros(self.p.m)])
        y = np.linalg.solve(self.B.T, full[self.basic])
        d = full - self.M.T @ y
        st = self.st
        t = tol * (1 + np.abs(full))
        return bool(np.all(d[st == AT_LOWER] >= -t[st == AT_LOWER]) and
                    np.all(d[st == AT_UPPER] <= t[st == AT_UPPER]) and
                    np.all(np.abs(d[st == FREE]) <= t[st == FREE]))


def solved(seed: int):
    ko = known_optimum_lp(40, 90, 4, seed, m_eq=15, ub=10.0)
    p = ko.problem
    s = solve_simplex(p, SimplexParams(feas_tol=1e-10, opt_tol=1e-10, sensitivity=True))
    assert s.status == Status.OPTIMAL and s.sensitivity is not None
    return p, s


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cost_ranges_are_where_the_basis_stays_optimal(seed):
    p, s = solved(seed)
    sv = s.sensitivity
    chk = BasisCheck(p, s.basis_status)
    assert chk.dual_feasible(p.c)
    checked = 0
    for j in range(p.n):
        lo, hi = sv.cost_lo[j], sv.cost_hi[j]
        assert lo <= p.c[j] + 1e-9 and p.c[j] <= hi + 1e-9
        w = hi - lo
        if np.isfinite(lo) and np.isfinite(hi):
            inside = [lo + 0.25 * w, lo + 0.75 * w] if w > 1e-6 else []   # zero width: degenerate
        elif np.isinf(lo) and np.isinf(hi):
            inside = [p.c[j] - 1.0, p.c[j] + 1.0]
        elif np.isinf(hi):
            inside = [max(p.c[j], lo) + 1.0]
        else:
            inside = [min(p.c[j], hi) - 1.0]
        for v in inside:
            c = p.c.copy()
            c[j] = v
            assert chk.dual_feasible(c, tol=1e-8), (j, lo, hi, v)
        for v, finite in ((hi + 0.05 * (1 + abs(hi)), np.isfinite(hi)),
                          (lo - 0.05 * (1 + abs(lo)), np.isfinite(lo))):
            if finite:
                c = p.c.copy()
                c[j] = v
                assert not chk.dual_feasible(c, tol=1e-12), (j, lo, hi, v)
                checked += 1
    assert checked >= p.n // 2


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rhs_ranges_are_where_the_basis_stays_feasible(seed):
    p, s = solved(seed)
    sv = s.sensitivity
    st = s.basis_status
    chk = BasisCheck(p, st)
    z0, _, _ = chk.point(p.row_lb, p.row_ub)
    obj0 = float(p.c @ z0[: p.n])
    assert obj0 == pytest.approx(s.objective - p.obj_offset, rel=1e-9, abs=1e-9)
    np.testing.assert_allclose(sv.activity, z0[p.n:], rtol=1e-9, atol=1e-9)
    checked = 0
    for i in np.flatnonzero(sv.binding):
        b, lo, hi = sv.rhs[i], sv.rhs_lo[i], sv.rhs_hi[i]
        assert lo <= b + 1e-9 <= hi + 2e-9
        logical = st[p.n + i]

        def moved(v):
            rl, ru = p.row_lb.copy(), p.row_ub.copy()
            if logical == FIXED:
                rl[i] = ru[i] = v
            elif logical == AT_LOWER:
                rl[i] = v
            else:
                ru[i] = v
            return rl, ru

        w = hi - lo
        for v in ([lo + 0.25 * w, lo + 0.75 * w] if np.isfinite(w) and w > 1e-6 else []):
            rl, ru = moved(v)
            assert chk.primal_feasible(rl, ru, tol=1e-8), (i, lo, hi, v)
            z, _, _ = chk.point(rl, ru)
            # inside the range the objective moves by the shadow price
            assert float(p.c @ z[: p.n]) - obj0 == pytest.approx(sv.dual[i] * (v - b),
                                                                 rel=1e-7, abs=1e-7)
        for v, finite in ((hi + 0.05 * (1 + abs(hi)), np.isfinite(hi)),
                          (lo - 0.05 * (1 + abs(lo)), np.isfinite(lo))):
            if finite:
                rl, ru = moved(v)
                if np.all(rl <= ru):
                    assert not chk.primal_feasible(rl, ru, tol=1e-12), (i, lo, hi, v)
                    checked += 1
    assert checked > 0
    # a non-binding row's bound may travel to its activity
    for i in np.flatnonzero(~sv.binding):
        assert sv.dual[i] == pytest.approx(0.0, abs=1e-9)
        assert sv.activity[i] in (sv.rhs_lo[i], sv.rhs_hi[i])


def test_maximisation_ranges_in_the_models_sense():
    """max 3 x0 + 2 x1, x0 + x1 <= 4, x0 + 3 x1 <= 7, x0 <= 3: optimum (3, 1)
    with x0 at its bound; x1 basic.  Cost of x1 may range over [0, 3] (below
    0 it leaves; above 3 x0 would rather decrease); the first row's shadow
    price is 2 and its bound ranges over [3, 13/3] (x1 = b - 3 stays >= 0 and
    the second row, 3 + 3 x1 <= 7, holds)."""
    A = SparseMatrix.from_dense(np.array([[1., 1.], [1., 3.]]))
    p = Problem(A=A, c=np.array([3., 2.]), row_lb=np.array([-INF, -INF]),
                row_ub=np.array([4., 7.]), col_lb=np.zeros(2), col_ub=np.array([3., INF]),
                sense=ObjSense.MAXIMISE)
    s = solve_simplex(p, SimplexParams(sensitivity=True))
    assert s.status == Status.OPTIMAL and abs(s.objective - 11.0) < 1e-12
    sv = s.sensitivity
    np.testing.assert_allclose(sv.value, [3.0, 1.0])
    assert sv.cost_lo[1] == pytest.approx(0.0) and sv.cost_hi[1] == pytest.approx(3.0)
    assert sv.cost_lo[0] == pytest.approx(2.0) and np.isinf(sv.cost_hi[0])
    np.testing.assert_allclose(sv.dual, [2.0, 0.0], atol=1e-12)
    assert sv.binding[0] and not sv.binding[1]
    assert sv.rhs_lo[0] == pytest.approx(3.0) and sv.rhs_hi[0] == pytest.approx(13.0 / 3.0)
    np.testing.assert_allclose(sv.reduced_cost, p.c - p.A.rmatvec(sv.dual), atol=1e-12)
    rep = sv.report()
    assert "Shadow prices" in rep and "R0" in rep and "C1" in rep


def test_crossover_hands_over_ranging():
    """The reference's claim for crossover: an interior point has no ranging;
    crossing over gives a basis and with it the report."""
    from sovlp.io import read_model
    p = read_model(__file__.rsplit("/", 1)[0] + "/golden/models/blend_plan_6h.mps.gz")
    opt = solve_simplex(p, SimplexParams(feas_tol=1e-9, opt_tol=1e-9))
    # a point off the vertex: the optimum blended with a feasible interior-ish point
    rng = np.random.default_rng(0)
    x = opt.x * (1.0 + 1e-7 * rng.standard_normal(p.n))
    xo = crossover(p, np.clip(x, p.col_lb, p.col_ub), opt.y,
                   SimplexParams(feas_tol=1e-9, opt_tol=1e-9, sensitivity=True))
    assert xo.status == Status.OPTIMAL
    assert abs(xo.objective - 6421.520443) <= 5e-7
    sv = xo.sensitivity
    assert sv is not None and int((xo.basis_status == BASIC).sum()) == p.m
    assert sv.binding.sum() > 0 and sv.report().strip()
    assert sv.col_names is not None and sv.col_names == p.col_names
    again = compute_sensitivity(p, xo.basis_status)
    np.testing.assert_array_equal(again.cost_lo, sv.cost_lo)
