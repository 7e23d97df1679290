"""Batched node relaxation on the GPU against the oracle.

The shared-A products are the oracle's bit for bit (same per-element order);
the per-node sums (step-size test, bound) run in the device's fixed blocked
order in both; bounds and iterates are compared to 1e-9 relative (the oracle's
NumPy elementwise steps may round differently from the device's), flags
exactly.  Every bound is checked valid against HiGHS.
"""

import numpy as np
import pytest

from oracle import bnr_oracle as bo
from sovlp.core.problem import Problem
from sovlp.core.sparse import SparseMatrix
from sovlp.workloads import known_optimum_lp

from test_bnr_oracle import REF, highs_optima, node_boxes, reference_case

pytestmark = pytest.mark.gpu


def _close(a, b, rel=1e-9):
    a, b = np.asarray(a), np.asarray(b)
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin) and np.array_equal(a[~fin], b[~fin])
    assert np.all(np.abs(a[fin] - b[fin]) <= rel * np.maximum(1.0, np.abs(b[fin]))), (a, b)


@pytest.mark.parametrize("m_eq,K,iters,textbook", [(30, 5, 60, False), (30, 5, 60, True),
                                                   (15, 64, 60, True), (15, 130, 20, False),
                                                   (15, 1, 200, True)])
def test_batches_match_the_oracle(m_eq, K, iters, textbook):
    print(f"[gpu] bnr m_eq={m_eq} K={K} iters={iters} textbook={textbook}", flush=True)
    from sovlp.mip.bnr import BNRConfig, BNREngine
    k = known_optimum_lp(30, 60, 3, seed=4, m_eq=m_eq)
    p = k.problem
    L, U = node_boxes(p, K, seed=3)
    if K > 2:
        L[5, 2], U[5, 2] = 3.0, 1.0                        # an empty node
    sign = -1.0 if textbook else 1.0
    eng = BNREngine(p, config=BNRConfig(batch=K, iters=iters, textbook_duals=textbook))
    r = eng.solve_batch(L, U)
    b, X, Y, eta, empty = bo.solve_batch(p, L, U, iters=iters, dual_sign=sign)
    assert np.array_equal(r.infeasible, empty)
    _close(r.bounds, np.where(empty, np.inf, b))
    _close(r.x, X)
    _close(eng.eta[:K], eta)
    # a second visit, warm from the first, with the step sizes carried over
    Xw, Yw = eng.iterates(K)
    r2 = eng.solve_batch(L, U, warm=(Xw, Yw))
    b2, X2, *_ = bo.solve_batch(p, L, U, iters=iters, warm=(X, Y), eta=eta, dual_sign=sign)
    _close(r2.bounds, np.where(empty, np.inf, b2))
    _close(r2.x, X2)
    opt = highs_optima(p, L, U)
    ok = ~empty
    assert np.all(r2.bounds[ok] <= opt[ok] + 1e-9 * np.maximum(1.0, np.abs(opt[ok])))


@pytest.mark.parametrize("case", REF["cases"], ids=[c["name"] for c in REF["cases"]])
def test_batches_match_the_reference_run(case):
    """The reference's own batches (its BNREngine run on the CPU, cold then warm
    with the step sizes the engine keeps): the device's X, Y, step sizes and
    bounds to 1e-12 relative (the per-node sums run in another order), the
    empty flags exactly."""
    from sovlp.mip.bnr import BNRConfig, BNREngine
    print(f"[gpu] bnr vs reference {case['name']}", flush=True)
    p, L, U, visits = reference_case(case)
    K = case["K"]
    eng = BNREngine(p, config=BNRConfig(batch=K, iters=case["iters"]))
    warm = None
    for v in visits:
        r = eng.solve_batch(L, U, warm=warm)
        X, Y = eng.iterates(K)
        assert np.array_equal(r.infeasible, v["empty"])
        _close(r.bounds, v["bounds"], 1e-12)
        _close(X, v["X"], 1e-12)
        _close(Y, v["Y"], 1e-12)
        _close(eng.eta[:K], v["eta"], 1e-12)
        warm = (X, Y)


def test_a_model_without_rows():
    from sovlp.mip.bnr import BNRConfig, BNREngine
    one = Problem(A=SparseMatrix.from_triplets(np.array([], np.int64), np.array([], np.int64),
                                               np.array([]), 0, 2),
                  c=np.array([1.0, -2.0]), row_lb=np.array([]), row_ub=np.array([]),
                  col_lb=np.array([1.0, 0.0]), col_ub=np.array([5.0, 3.0]))
    r = BNREngine(one, config=BNRConfig(batch=2, iters=5)).solve_batch(
        np.array([[1.0, 2.0], [0.0, 0.0]]), np.array([[5.0, 5.0], [3.0, 1.0]]))
    # min x0 - 2 x1: x0 at its lower bound, x1 at its upper bound
    np.testing.assert_array_equal(r.bounds, [1.0 - 6.0, 2.0 - 2.0])


def test_calls_under_another_stream_match_the_default_stream():
    """Engines built on the default stream and called under another stream queue
    their uploads and kernels on that stream: the same results, bit for bit."""
    import torch
    from sovlp.lp.certify import DeviceCertifier
    from sovlp.mip.bnr import BNRConfig, BNREngine
    k = known_optimum_lp(30, 60, 3, seed=4, m_eq=15)
    p = k.problem
    L, U = node_boxes(p, 8, seed=3)
    a = BNREngine(p, config=BNRConfig(batch=8, iters=60)).solve_batch(L, U)
    eng = BNREngine(p, config=BNRConfig(batch=8, iters=60))
    dc = DeviceCertifier(p)
    ref = dc.certify(k.x_star, k.y_star)
    ref = (ref.bound, ref.perturbation, ref.objective)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b = eng.solve_batch(L, U)
        got = dc.certify(k.x_star, k.y_star)
        got = (got.bound, got.perturbation, got.objective)
    torch.cuda.synchronize()
    assert np.array_equal(a.bounds, b.bounds) and np.array_equal(a.x, b.x)
    assert got == ref


def test_repeated_calls_replay_the_same_launches():
    """The library replays a repeated call as a HIP graph (captured on the
    second call with the same arguments): a first (direct), second (captured)
    and later (replayed) call on the same boxes give the same results bit for
    bit, and a call with other boxes in between does not leak into them."""
    from sovlp.mip.bnr import BNRConfig, BNREngine
    k = known_optimum_lp(30, 60, 3, seed=6, m_eq=15)
    p = k.problem
    L, U = node_boxes(p, 16, seed=4)
    L2, U2 = node_boxes(p, 16, seed=9)
    print("[gpu] bnr: repeated calls", flush=True)
    eng = BNREngine(p, config=BNRConfig(batch=16, iters=40, textbook_duals=True))
    runs = []
    for i in range(5):
        # each visit from the same state: cold, the engine's step sizes reset
        eng.eta[:] = 1.0
        eng.omega[:] = 1.0
        if i == 3:
            other = eng.solve_batch(L2, U2)
            eng.eta[:] = 1.0
            eng.omega[:] = 1.0
        r = eng.solve_batch(L, U)
        runs.append((r.bounds.copy(), r.x.copy(), eng.eta[:16].copy()))
    for b, x, e in runs[1:]:
        assert np.array_equal(b, runs[0][0]) and np.array_equal(x, runs[0][1])
        assert np.array_equal(e, runs[0][2])
    assert not np.array_equal(other.bounds, runs[0][0])
    fresh = BNREngine(p, config=BNRConfig(batch=16, iters=40, textbook_duals=True))
    assert np.array_equal(fresh.solve_batch(L2, U2).bounds, other.bounds)
