"""Crossover from a PDLP point to an optimal basis, and the simplex behind it.

Restates the reference's own crossover tests (``tests/test_crossover.py``):
the identified basis is nonsingular by construction (no repair), a fixed
variable is never made basic, the crossover agrees with a cold simplex, a
maximisation comes back in the caller's sense, and crossing over saves
pivots against a cold start in aggregate.  The objectives are pinned on
HiGHS (``scipy.optimize.linprog``) and on the LPs' closed-form optima, and
every returned basis is checked against the optimality conditions directly:
primal feasibility, ``d = c - Aᵀy``, reduced-cost signs by status, and
strong duality.  The points crossed over from are the LPs' known optima
blended toward the box centre (off every vertex, as a first-order point is);
``tests/test_gpu_crossover.py`` crosses over from device PDLP solves.
"""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

from sovlp import _lib
from sovlp.core.problem import ObjSense, Problem, Status
from sovlp.core.sparse import SparseMatrix
from sovlp.core.tolerances import INF
from sovlp.lp.basis import AT_LOWER, AT_UPPER, BASIC, FIXED, FREE, Basis
from sovlp.lp.crossover import basis_from_point, crossover
from sovlp.lp.simplex import SimplexParams, solve_simplex
from sovlp.workloads import known_optimum_lp, uneven_batch_member


def highs(p: Problem) -> float:
    A = p.A
    M = sp.csc_matrix((A.cx, A.ci, A.cp), shape=(p.m, p.n))
    eq = p.row_lb == p.row_ub
    up = ~eq & (p.row_ub < INF)
    lo = ~eq & (p.row_lb > -INF)
    Aub = sp.vstack([M[up], -M[lo]])
    bub = np.concatenate([p.row_ub[up], -p.row_lb[lo]])
    c = -p.c if p.sense == ObjSense.MAXIMISE else p.c
    bounds = [(a if a > -INF else None, b if b < INF else None)
              for a, b in zip(p.col_lb, p.col_ub)]
    r = linprog(c, A_ub=Aub if Aub.shape[0] else None, b_ub=bub if Aub.shape[0] else None,
                A_eq=M[eq] if eq.any() else None, b_eq=p.row_lb[eq] if eq.any() else None,
                bounds=bounds, method="highs")
    assert r.status == 0
    f = r.fun if p.sense == ObjSense.MINIMISE else -r.fun
    return f + p.obj_offset


def lp(seed: int):
    """A small known-optimum LP (about a third of the seeds are uneven-batch
    members, the rest 60 x 150)."""
    if seed % 3 == 0:
        return uneven_batch_member(2000 + seed)
    return known_optimum_lp(60, 150, 4, seed, m_eq=30, ub=10.0)


def interior_point(ko, w: float = 1e-3):
    """The known optimum blended toward the centre of the column box."""
    p = ko.problem
    mid = 0.5 * (p.col_lb + np.minimum(p.col_ub, 1e3))
    return (1.0 - w) * ko.x_star + w * np.clip(mid, p.col_lb, p.col_ub), ko.y_star


def assert_optimal_basis(p: Problem, s, tol: float = 1e-8):
    """The returned basis satisfies the optimality conditions of the LP in the
    caller's sense: exactly m basic, x within its bounds and the row bounds,
    d = c - Aᵀy, reduced costs signed by status, primal = dual objective."""
    m, n = p.m, p.n
    st = s.basis_status
    assert st.shape == (n + m,) and int((st == BASIC).sum()) == m
    rv, bv, _ = p.violation(s.x)
    scale = 1.0 + max(np.abs(p.row_lb[p.row_lb > -INF]).max(initial=0.0),
                      np.abs(p.row_ub[p.row_ub < INF]).max(initial=0.0))
    assert rv <= tol * scale and bv <= tol * scale
    d = p.c - p.A.rmatvec(s.y)
    np.testing.assert_allclose(s.reduced_costs, d, rtol=1e-9, atol=1e-9)
    sg = -1.0 if p.sense == ObjSense.MAXIMISE else 1.0     # minimisation signs
    dm, ym = sg * d, sg * s.y
    ctol = tol * (1.0 + np.abs(p.c).max())
    xs = st[:n]
    assert np.all(np.abs(dm[xs == BASIC]) <= ctol)
    assert np.all(dm[xs == AT_LOWER] >= -ctol)
    assert np.all(dm[xs == AT_UPPER] <= ctol)
    assert np.all(np.abs(dm[xs == FREE]) <= ctol)
    # logicals: d = +y over [A | -I]
    ls = st[n:]
    assert np.all(np.abs(ym[ls == BASIC]) <= ctol)
    # strong duality: c x = y.(row bound) + d.(column bound) at the basis
    rb = np.where(ls == AT_UPPER, p.row_ub, p.row_lb)
    rb = np.where(ls == BASIC, 0.0, rb)
    cb = np.where(xs == AT_UPPER, p.col_ub, np.where(xs == FREE, 0.0, p.col_lb))
    cb = np.where(xs == BASIC, 0.0, cb)
    dual = float(np.dot(s.y, rb) + np.dot(d, cb)) + p.obj_offset
    assert abs(dual - s.objective) <= 1e-7 * max(1.0, abs(s.objective))


# --------------------------------------------------------------------------- #
# the basis itself                                                             #
# --------------------------------------------------------------------------- #


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_identified_basis_is_nonsingular_without_repair(seed):
    ko = lp(seed)
    p = ko.problem
    x, y = interior_point(ko)
    st = basis_from_point(p, x, y)
    assert st.shape == (p.n + p.m,)
    assert int((st == BASIC).sum()) == p.m
    B = Basis(p)
    B.status[:] = st
    B.factorize()
    assert B.n_repairs == 0


def test_factorisation_counts_the_repairs_of_a_singular_basis():
    """Two copies of one column in a basis: one of them has no pivot."""
    A = SparseMatrix.from_dense(np.array([[1., 1., 2.], [2., 2., 1.]]))
    p = Problem(A=A, c=np.ones(3), row_lb=np.array([1.0, 1.0]), row_ub=np.array([INF, INF]),
                col_lb=np.zeros(3), col_ub=np.full(3, 10.0))
    B = Basis(p)
    B.status[:] = [BASIC, BASIC, AT_LOWER, AT_LOWER, AT_LOWER]
    B.factorize()
    assert B.n_repairs == 1
    B.status[:] = [BASIC, AT_LOWER, BASIC, AT_LOWER, AT_LOWER]
    B.factorize()
    assert B.n_repairs == 0
    # a warm start from the singular basis is repaired and still solves
    s = solve_simplex(p, warm_basis=np.array([BASIC, BASIC, AT_LOWER, AT_LOWER, AT_LOWER],
                                             dtype=np.uint8))
    assert s.status == Status.OPTIMAL and s.info["repairs"] == 1
    assert abs(s.objective - highs(p)) <= 1e-12


def test_a_fixed_variable_is_never_made_basic():
    A = SparseMatrix.from_dense(np.array([[1., 1., 1.], [1., -1., 0.]]))
    p = Problem(A=A, c=np.array([1., 1., -1.]), row_lb=np.array([4.0, -INF]),
                row_ub=np.array([4.0, 3.0]), col_lb=np.array([0.0, 2.0, 0.0]),
                col_ub=np.array([5.0, 2.0, 5.0]), name="pinned")
    x = np.array([1.0, 2.0, 1.0])
    st = basis_from_point(p, x, np.zeros(2))
    assert st[1] == FIXED
    s = crossover(p, x, np.zeros(2))
    assert s.status == Status.OPTIMAL
    assert s.basis_status[1] != BASIC
    assert abs(s.objective - highs(p)) <= 1e-9


def test_free_columns_are_pushed_first():
    """A free structural is always a candidate, ahead of every scored one."""
    ko = lp(1)
    p = ko.problem
    lb = p.col_lb.copy()
    ub = p.col_ub.copy()
    lb[0], ub[0] = -INF, INF
    q = Problem(A=p.A, c=p.c, row_lb=p.row_lb, row_ub=p.row_ub, col_lb=lb, col_ub=ub)
    x, y = interior_point(ko)
    st = basis_from_point(q, x, y)
    assert st[0] == BASIC


# --------------------------------------------------------------------------- #
# it must not change the answer                                                #
# --------------------------------------------------------------------------- #


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5])
def test_agrees_with_a_cold_simplex_highs_and_the_closed_form(seed):
    ko = lp(seed)
    p = ko.problem
    cold = solve_simplex(p)
    x, y = interior_point(ko)
    xo = crossover(p, x, y)
    ref = highs(p)
    assert cold.status == xo.status == Status.OPTIMAL
    for s in (cold, xo):
        assert abs(s.objective - ref) <= 1e-9 * max(1.0, abs(ref))
        assert abs(s.objective - ko.f_star) <= 1e-9 * max(1.0, abs(ko.f_star))
        assert_optimal_basis(p, s)
    assert xo.info["crossover"] and xo.info["repairs"] == 0


def test_the_reference_two_step_crossover_agrees():
    """The reference's own sequence, basis_from_point then solve_simplex from
    those statuses alone, reaches the same optimum (in more pivots)."""
    ko = lp(1)
    x, y = interior_point(ko)
    st = basis_from_point(ko.problem, x, y)
    two = solve_simplex(ko.problem, warm_basis=st)
    one = crossover(ko.problem, x, y)
    assert two.status == one.status == Status.OPTIMAL
    assert abs(two.objective - one.objective) <= 1e-9 * max(1.0, abs(one.objective))
    assert one.iterations <= two.iterations
    assert one.info["superbasics"] > 0


def test_crosses_over_from_a_point_without_duals():
    ko = lp(2)
    x, _ = interior_point(ko)
    xo = crossover(ko.problem, x)
    assert xo.status == Status.OPTIMAL
    assert abs(xo.objective - ko.f_star) <= 1e-9 * max(1.0, abs(ko.f_star))


def test_handles_a_maximisation():
    """max 3 x0 + 2 x1 s.t. x0 + x1 <= 4, x0 + 3 x1 <= 6, x0 <= 3: 11 at (3, 1);
    duals and reduced costs in the caller's (maximising) sense."""
    A = SparseMatrix.from_dense(np.array([[1., 1.], [1., 3.]]))
    p = Problem(A=A, c=np.array([3., 2.]), row_lb=np.array([-INF, -INF]),
                row_ub=np.array([4., 6.]), col_lb=np.zeros(2), col_ub=np.array([3., INF]),
                sense=ObjSense.MAXIMISE, name="toy")
    s = crossover(p, np.array([2.5, 1.0]), np.array([2.0, 0.0]))
    assert s.status == Status.OPTIMAL
    assert abs(s.objective - 11.0) < 1e-9
    np.testing.assert_allclose(s.x, [3.0, 1.0], atol=1e-12)
    np.testing.assert_allclose(s.y, [2.0, 0.0], atol=1e-12)     # the first row binds
    assert_optimal_basis(p, s)
    ko = lp(4)
    q = ko.problem.copy()
    q.c = -q.c
    q.sense = ObjSense.MAXIMISE
    x, y = interior_point(ko)
    t = crossover(q, x, -y)
    assert t.status == Status.OPTIMAL
    assert abs(t.objective + ko.f_star) <= 1e-9 * max(1.0, abs(ko.f_star))
    assert_optimal_basis(q, t)


def test_saves_pivots_against_a_cold_start():
    cold_total = xover_total = 0
    for seed in range(6):
        ko = lp(seed)
        x, y = interior_point(ko)
        cold = solve_simplex(ko.problem)
        xo = crossover(ko.problem, x, y)
        assert abs(xo.objective - cold.objective) <= 1e-9 * max(1.0, abs(cold.objective))
        cold_total += cold.iterations
        xover_total += xo.iterations
    assert xover_total < 0.5 * cold_total, (xover_total, cold_total)


def test_warm_start_from_an_optimal_basis_takes_no_pivots():
    ko = lp(1)
    first = solve_simplex(ko.problem)
    again = solve_simplex(ko.problem, warm_basis=first.basis_status)
    assert again.status == Status.OPTIMAL and again.iterations == 0
    assert again.objective == pytest.approx(first.objective, rel=1e-12)


# --------------------------------------------------------------------------- #
# the simplex's other outcomes and its argument checks                         #
# --------------------------------------------------------------------------- #


def test_infeasible_and_unbounded_are_reported():
    A = SparseMatrix.from_dense(np.array([[1., 1.]]))
    inf = Problem(A=A, c=np.array([1., 1.]), row_lb=np.array([5.0]), row_ub=np.array([INF]),
                  col_lb=np.zeros(2), col_ub=np.array([1.0, 1.0]))
    assert solve_simplex(inf).status == Status.INFEASIBLE
    unb = Problem(A=A, c=np.array([-1., 0.]), row_lb=np.array([1.0]), row_ub=np.array([INF]),
                  col_lb=np.zeros(2), col_ub=np.array([INF, INF]))
    assert solve_simplex(unb).status == Status.UNBOUNDED


def test_arguments_are_checked():
    ko = lp(1)
    p = ko.problem
    with pytest.raises(ValueError, match="warm_basis"):
        solve_simplex(p, warm_basis=np.zeros(3, dtype=np.uint8))
    with pytest.raises(ValueError, match="x has"):
        basis_from_point(p, np.zeros(p.n + 1))
    B = Basis(p)
    B.status[:] = AT_LOWER
    with pytest.raises(ValueError, match="exactly m"):
        B.factorize()
    big = Problem(A=SparseMatrix.from_triplets(np.array([0]), np.array([0]), np.array([1.0]),
                                               8193, 1),
                  c=np.ones(1), row_lb=np.full(8193, -INF), row_ub=np.full(8193, INF),
                  col_lb=np.zeros(1), col_ub=np.ones(1))
    with pytest.raises(_lib.SovlpError, match="8192 rows"):
        solve_simplex(big)


# --------------------------------------------------------------------------- #
# the model fixtures, by the simplex alone                                     #
# --------------------------------------------------------------------------- #

MODELS = __file__.rsplit("/", 1)[0] + "/golden/models/"


@pytest.mark.parametrize("fname", ["testprob.mps", "testprob.lp"])
def test_simplex_solves_testprob(fname):
    """Optimum 16 at x = (0, -1, 6) (reference tests/test_solvers.py:58-66)."""
    from sovlp.solve import solve_file
    s = solve_file(MODELS + fname, method="simplex")
    assert s.status == Status.OPTIMAL
    assert abs(s.objective - 16.0) < 1e-12
    np.testing.assert_allclose(s.x, [0.0, -1.0, 6.0], atol=1e-12)


def test_simplex_solves_the_hourly_blending_plan():
    """blend_plan_6h (654 x 1104, a maximisation): the reference's simplex
    finds 6421.520443 and its test holds it to 1e-6 relative in under 10,000
    iterations (reference tests/test_simplex.py:230-245).  At the default
    tolerances this simplex is within that; at 1e-9 it agrees with the
    reference's figure to every digit the reference prints."""
    from sovlp.io import read_model
    p = read_model(MODELS + "blend_plan_6h.mps.gz")
    s = solve_simplex(p)
    assert s.status == Status.OPTIMAL
    assert abs(s.objective - 6421.520443) <= 1e-6 * 6421.520443
    assert s.iterations < 10_000
    t = solve_simplex(p, SimplexParams(feas_tol=1e-9, opt_tol=1e-9))
    assert t.status == Status.OPTIMAL
    assert abs(t.objective - 6421.520443) <= 5e-7
    assert_optimal_basis(p, t, tol=1e-9)
