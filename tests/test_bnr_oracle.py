"""The batched-node-relaxation oracle (oracle/bnr_oracle.py), pinned on the CPU.

Pins: the reference's own batches (``tests/golden/reference_bnr.json``, written
by running its ``BNREngine.solve_batch`` on the CPU: ``make_bnr_vectors.py``),
reproduced bit for bit with the reference's summation order and to rounding
with the device's; at the exact optimal duals of a known-optimum LP the
batched bound is the optimum (closed form); for any iterate it is below each
node's optimum as HiGHS computes it (scipy); an empty box is flagged; with
textbook-signed duals the bound converges to the optimum; a one-variable case
by hand.
"""

import json
import os

import numpy as np
import pytest
from scipy.optimize import linprog

from oracle import bnr_oracle as bo
from sovlp.core.problem import Problem
from sovlp.core.sparse import SparseMatrix
from sovlp.core.tolerances import INF
from sovlp.workloads import known_optimum_lp


def node_boxes(p, K, seed):
    rng = np.random.default_rng(seed)
    L = np.repeat(p.col_lb[:, None], K, 1).copy()
    U = np.repeat(p.col_ub[:, None], K, 1).copy()
    for q in range(1, K):
        js = rng.choice(p.n, 5, replace=False)
        U[js, q] = np.minimum(U[js, q], rng.uniform(0.0, 5.0, 5))
        ls = rng.choice(p.n, 3, replace=False)
        L[ls, q] = np.minimum(U[ls, q], rng.uniform(0.0, 2.0, 3))
    return L, U


def highs_optima(p, L, U):
    A = p.A.to_dense()
    rl, ru = p.row_lb, p.row_ub
    A_ub = np.vstack([A[ru < INF], -A[rl > -INF]])
    b_ub = np.concatenate([ru[ru < INF], -rl[rl > -INF]])
    out = []
    for q in range(L.shape[1]):
        r = linprog(p.c, A_ub=A_ub, b_ub=b_ub, bounds=list(zip(L[:, q], U[:, q])),
                    method="highs")
        out.append(r.fun if r.status == 0 else np.inf)
    return np.array(out)


REF = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_bnr.json")))


def reference_case(c):
    """(problem, L, U, visits) of a golden case; bounds as floats."""
    A = SparseMatrix.from_triplets(np.array(c["rows"]), np.array(c["cols"]),
                                   np.array(c["vals"], float), c["m"], c["n"])
    p = Problem(A=A, c=np.array(c["c"]), row_lb=np.array(c["rl"]), row_ub=np.array(c["ru"]),
                col_lb=np.zeros(c["n"]), col_ub=np.full(c["n"], 10.0))
    visits = []
    for v in c["visits"]:
        visits.append(dict(X=np.array(v["X"]), Y=np.array(v["Y"]), eta=np.array(v["eta"]),
                           omega=np.array(v["omega"]),
                           bounds=np.array([float(b) for b in v["bounds"]]),
                           empty=np.array(v["empty"])))
    return p, np.array(c["L"]), np.array(c["U"]), visits


@pytest.mark.parametrize("case", REF["cases"], ids=[c["name"] for c in REF["cases"]])
def test_reference_order_reproduces_the_reference_bit_for_bit(case):
    """Cold then warm visits (the second starting from the first's X, Y and step
    sizes, as the engine carries them): X, Y, step sizes and flags bit for bit;
    the bounds too, with an empty node's bound reported as +inf as the engine
    does (bnr.py:285-286)."""
    p, L, U, visits = reference_case(case)
    warm = eta = omega = None
    for v in visits:
        b, X, Y, e, empty = bo.solve_batch(p, L, U, iters=case["iters"], warm=warm, eta=eta,
                                           omega=omega, order="rows")
        assert np.array_equal(X, v["X"]) and np.array_equal(Y, v["Y"])
        assert np.array_equal(e, v["eta"])
        assert np.array_equal(empty, v["empty"])
        assert np.array_equal(np.where(empty, np.inf, b), v["bounds"])
        warm, eta, omega = (v["X"], v["Y"]), v["eta"], v["omega"]


@pytest.mark.parametrize("case", REF["cases"], ids=[c["name"] for c in REF["cases"]])
def test_device_order_agrees_with_the_reference_to_rounding(case):
    """The device's summation order (what the GPU is compared with) moves the
    reference's batches by rounding only."""
    p, L, U, visits = reference_case(case)
    warm = eta = omega = None
    for v in visits:
        b, X, Y, e, empty = bo.solve_batch(p, L, U, iters=case["iters"], warm=warm, eta=eta,
                                           omega=omega, order="device")
        for got, want in ((X, v["X"]), (Y, v["Y"]), (e, v["eta"])):
            assert np.all(np.abs(got - want) <= 1e-12 * np.maximum(1.0, np.abs(want)))
        bb, rb = np.where(empty, np.inf, b), v["bounds"]
        fin = np.isfinite(rb)
        assert np.array_equal(np.isfinite(bb), fin) and np.array_equal(bb[~fin], rb[~fin])
        assert np.all(np.abs(bb[fin] - rb[fin]) <= 1e-12 * np.maximum(1.0, np.abs(rb[fin])))
        warm, eta, omega = (v["X"], v["Y"]), v["eta"], v["omega"]


def test_the_bound_at_the_optimal_duals_is_the_optimum():
    k = known_optimum_lp(30, 60, 3, seed=4, m_eq=15)
    p = k.problem
    b, *_ = bo.solve_batch(p, p.col_lb[:, None], p.col_ub[:, None], iters=0,
                           warm=(k.x_star[:, None], k.y_star[:, None]))
    assert b[0] <= k.f_star + 1e-12 * abs(k.f_star)
    assert k.f_star - b[0] <= 1e-12 * abs(k.f_star)


@pytest.mark.parametrize("m_eq", [30, 15])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_every_bound_is_valid(m_eq, sign):
    k = known_optimum_lp(30, 60, 3, seed=4, m_eq=m_eq)
    p = k.problem
    L, U = node_boxes(p, 6, seed=1)
    opt = highs_optima(p, L, U)
    for iters in (1, 20, 200):
        b, X, Y, eta, empty = bo.solve_batch(p, L, U, iters=iters, dual_sign=sign)
        assert np.all(b <= opt + 1e-9 * np.maximum(1.0, np.abs(opt))), (iters, b, opt)
        assert np.all(X >= L - 1e-12) and np.all(X <= U + 1e-12)


def test_textbook_duals_converge_to_the_optimum_and_the_reference_sign_does_not():
    k = known_optimum_lp(30, 60, 3, seed=4, m_eq=30)
    p = k.problem
    L, U = p.col_lb[:, None].copy(), p.col_ub[:, None].copy()
    tb, *_ = bo.solve_batch(p, L, U, iters=3000, dual_sign=-1.0)
    rb, *_ = bo.solve_batch(p, L, U, iters=3000, dual_sign=1.0)
    assert tb[0] <= k.f_star and k.f_star - tb[0] <= 2e-3 * abs(k.f_star)
    assert k.f_star - rb[0] > 1.0 * abs(k.f_star)        # valid, but far below


def test_an_empty_box_is_flagged_and_one_variable_by_hand():
    k = known_optimum_lp(10, 20, 2, seed=2, m_eq=5)
    p = k.problem
    L, U = p.col_lb[:, None].repeat(2, 1), p.col_ub[:, None].repeat(2, 1)
    L[3, 1], U[3, 1] = 2.0, 1.0
    *_, empty = bo.solve_batch(p, L, U, iters=3)
    assert list(empty) == [False, True]
    # min x over [1, 5], no rows: the start is clamped to 1 and stays; d = 1 > 0 -> lo
    one = Problem(A=SparseMatrix.from_triplets(np.array([], np.int64), np.array([], np.int64),
                                               np.array([]), 0, 1),
                  c=np.array([1.0]), row_lb=np.array([]), row_ub=np.array([]),
                  col_lb=np.array([1.0]), col_ub=np.array([5.0]))
    b, X, *_ = bo.solve_batch(one, np.array([[1.0]]), np.array([[5.0]]), iters=5)
    assert b[0] == 1.0 and X[0, 0] == 1.0
