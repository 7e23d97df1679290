"""The certificate oracle (oracle/certify_oracle.c), pinned on the CPU.

Pins: the reference's own known answers (tests/test_verify_optimality.py:82-96,
106-132 of the reference), the closed form of the known-optimum LPs (at the
exact optimal duals the Neumaier-Shcherbina bound is the optimum, and for any
other duals it is below it), the sense symmetry of a maximisation, and exact
rational arithmetic for the compensated residual.
"""

from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle
from sovlp.core.problem import ObjSense, Problem
from sovlp.core.sparse import SparseMatrix
from sovlp.core.tolerances import INF
from sovlp.workloads import small_lp, uneven_batch_member

EPS = np.finfo(np.float64).eps


def free_column_lp():
    A = SparseMatrix.from_dense(np.array([[1.0, 1.0]]))
    return Problem(A=A, c=np.array([1.0, 2.0]), row_lb=np.array([1.0]), row_ub=np.array([1.0]),
                   col_lb=np.array([0.0, -INF]), col_ub=np.array([INF, INF]))


def steep_lp():
    A = SparseMatrix.from_triplets(np.array([0]), np.array([0]), np.array([1e6]), 1, 1)
    return Problem(A=A, c=np.array([1e8]), row_lb=np.array([1e-4]), row_ub=np.array([INF]),
                   col_lb=np.zeros(1), col_ub=np.full(1, INF), name="steep")


def test_duals_pointing_at_an_infinite_bound_certify_nothing():
    # d = c - A'y = (-4, -3): both point at an infinite bound; the reference's
    # answer is (-inf, 4.0)
    c = oracle.certify(free_column_lp(), np.array([1.0, 0.0]), np.array([5.0]))
    assert c.bound == -np.inf and c.perturbation == 4.0 and c.vacuous


def test_a_small_reduced_cost_at_an_infinite_bound_is_absorbed_and_reported():
    # min x0 + x1, x0 + x1 = 1, x0 >= 0, x1 free; y = 1 - 2^-40 leaves
    # d = (2^-40, 2^-40), and d1 > 0 points at x1's infinite lower bound
    q = Problem(A=SparseMatrix.from_dense(np.array([[1.0, 1.0]])), c=np.array([1.0, 1.0]),
                row_lb=np.array([1.0]), row_ub=np.array([1.0]),
                col_lb=np.array([0.0, -INF]), col_ub=np.array([INF, INF]))
    y = np.array([1.0 - 2.0 ** -40])
    c = oracle.certify(q, np.array([1.0, 0.0]), y)
    # below dual_tol (1 + |c|_inf): absorbed and reported; the bound is y * rl
    assert not c.vacuous and c.perturbation == 2.0 ** -40
    assert c.bound == pytest.approx(1.0 - 2.0 ** -40, abs=1e-15) and c.bound < 1.0
    # above it: the bound is vacuous, the perturbation still reported
    c2 = oracle.certify(q, np.array([1.0, 0.0]), y, dual_tol=1e-20)
    assert c2.vacuous and c2.bound == -np.inf and c2.perturbation == 2.0 ** -40


def test_the_steep_model_certifies_its_exact_pair_and_sees_the_shifted_point():
    p = steep_lp()
    good = oracle.certify(p, np.array([1e-10]), np.array([100.0]))
    assert good.bound <= 0.01 and 0.01 - good.bound <= 1e-16
    assert good.row_violation == 0.0
    bad = oracle.certify(p, np.array([1e-10 - 5e-13]), np.array([100.0]))
    assert 4e-7 < bad.row_violation <= 1e-6        # inside the 1e-6 feasibility line
    assert (bad.bound - bad.objective) / bad.objective > 1e-5   # but below the bound


@pytest.mark.parametrize("make", [lambda: small_lp(0), lambda: uneven_batch_member(1000),
                                  lambda: uneven_batch_member(1003)])
def test_the_bound_at_the_optimal_duals_is_the_optimum(make):
    k = make()
    c = oracle.certify(k.problem, k.x_star, k.y_star)
    f = k.f_star
    assert not c.vacuous and c.perturbation == 0.0
    assert c.bound <= f
    assert f - c.bound <= 1e-13 * max(1.0, abs(f))
    assert c.col_violation == 0.0 and c.row_violation <= 1e-12
    assert c.objective == pytest.approx(f, rel=1e-14)
    # the reduced costs are the construction's z = c - A'y*
    d_ref = k.problem.c - k.problem.A.rmatvec(k.y_star)
    np.testing.assert_allclose(c.d, d_ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_any_dual_vector_gives_a_valid_bound(seed):
    k = small_lp(seed)
    rng = np.random.default_rng(seed)
    for scale in (1e-8, 1e-3, 1.0):
        y = k.y_star + scale * rng.standard_normal(k.problem.m)
        c = oracle.certify(k.problem, k.x_star, y)
        assert c.bound <= k.f_star + 1e-12 * abs(k.f_star)
    # a random y's bound is far below; a slightly wrong one only a little
    c1 = oracle.certify(k.problem, k.x_star, k.y_star + 1e-8 * rng.standard_normal(k.problem.m))
    assert k.f_star - c1.bound < 1e-4 * abs(k.f_star)


def test_a_maximisation_certifies_an_upper_bound():
    k = uneven_batch_member(1001)
    p = k.problem
    q = Problem(A=p.A, c=-p.c, row_lb=p.row_lb, row_ub=p.row_ub, col_lb=p.col_lb,
                col_ub=p.col_ub, obj_offset=2.5, sense=ObjSense.MAXIMISE)
    # max -c'x + 2.5 = -(min c'x) + 2.5; its textbook duals are -y*
    c = oracle.certify(q, k.x_star, -k.y_star)
    opt = -k.f_star + 2.5
    assert c.bound >= opt and c.bound - opt <= 1e-13 * max(1.0, abs(opt))
    assert c.objective == pytest.approx(opt, rel=1e-14)


def _exact_residual(rp, ri, rx, x, b):
    out = []
    for i in range(len(rp) - 1):
        s = Fraction(b[i])
        for p in range(rp[i], rp[i + 1]):
            s -= Fraction(rx[p]) * Fraction(x[ri[p]])
        out.append(s)
    return out


def test_the_compensated_residual_is_double_double_accurate():
    # a row whose terms cancel: the plain sum is 0 or 2, the residual is -1
    rp = np.array([0, 3], dtype=np.int64)
    ri = np.array([0, 1, 2], dtype=np.int32)
    rx = np.array([1e16, 1.0, -1e16])
    x = np.ones(3)
    r = oracle.residual(rp, ri, rx, x, np.zeros(1))
    assert r[0] == -1.0
    assert float(np.dot(rx, x)) != 1.0                     # the plain dot loses it
    rng = np.random.default_rng(7)
    m, n = 60, 40
    for trial in range(3):
        lens = rng.integers(0, 30, m)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        ri = rng.integers(0, n, rp[-1]).astype(np.int32)
        rx = rng.standard_normal(rp[-1]) * 10.0 ** rng.integers(-8, 9, rp[-1])
        x = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 9, n)
        # b = the rounded A x, so the residuals are tiny against the terms
        b = np.array([sum(rx[p] * x[ri[p]] for p in range(rp[i], rp[i + 1])) for i in range(m)])
        r = oracle.residual(rp, ri, rx, x, b)
        ex = _exact_residual(rp, ri, rx, x, b)
        for i in range(m):
            k = int(rp[i + 1] - rp[i])
            mag = abs(b[i]) + sum(abs(rx[p] * x[ri[p]]) for p in range(rp[i], rp[i + 1]))
            err = abs(Fraction(r[i]) - ex[i])
            assert err <= EPS * abs(ex[i]) + 2 * (k + 2) ** 2 * EPS ** 2 * Fraction(mag), (trial, i)
