"""Device certificate and verifier against the oracle (GPU).

Bit for bit where the device follows the reference's operation order (the
compensated residual and row activities, the reduced costs c - A'y, the
violations and the perturbation); the bound within 1e-14 relative of the
oracle's (double-double sums against the reference's sequential Neumaier sum,
both exact to a few ulps); the verdicts of the reference's verifier tests on
points from this package's own GPU solver.
"""

import numpy as np
import pytest

from oracle import oracle
from sovlp.core.problem import ObjSense, Problem, Status
from sovlp.core.sparse import SparseMatrix
from sovlp.core.tolerances import INF
from sovlp.workloads import skewed_lp_degrees, small_lp, uneven_batch_member

pytestmark = pytest.mark.gpu


def _cert():
    from sovlp.lp import certify
    return certify


def _same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.int64), b.view(np.int64))


def _random_csr(rng, m, n, maxlen, long_rows=0):
    lens = rng.integers(0, maxlen, m)
    if long_rows:
        lens[rng.choice(m, long_rows, replace=False)] = 5000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ri = rng.integers(0, n, rp[-1]).astype(np.int32)
    rx = rng.standard_normal(rp[-1]) * 10.0 ** rng.integers(-6, 7, rp[-1])
    return rp, ri, rx


def test_compensated_residual_is_the_references_bit_for_bit():
    print("[gpu] certify: compensated residual", flush=True)
    cert = _cert()
    rng = np.random.default_rng(3)
    for m, n, maxlen, long_rows in [(1, 3, 4, 0), (500, 300, 40, 0), (2000, 5000, 12, 5)]:
        rp, ri, rx = _random_csr(rng, m, n, maxlen, long_rows)
        x = rng.standard_normal(n) * 10.0 ** rng.integers(-6, 7, n)
        b = rng.standard_normal(m)
        got = cert.compensated_residual(rp, ri, rx, x, b)
        _same_bits(got, oracle.residual(rp, ri, rx, x, b))
    # the cancelling row: exactly -1
    r = cert.compensated_residual(np.array([0, 3]), np.array([0, 1, 2]),
                                  np.array([1e16, 1.0, -1e16]), np.ones(3), np.zeros(1))
    assert r[0] == -1.0
    assert cert.compensated_residual(np.array([0]), np.array([], np.int32), np.array([]),
                                     np.ones(2), np.array([])).size == 0


def _check_against_oracle(prob, x, y, dual_tol=1e-6):
    cert = _cert()
    dc = cert.DeviceCertifier(prob)
    got = dc.certify(x, y, dual_tol)
    ref = oracle.certify(prob, x, y, dual_tol)
    _same_bits(dc.reduced_costs(), ref.d)
    _same_bits(dc.activities(), ref.act)
    assert got.perturbation == ref.perturbation
    assert got.col_violation == ref.col_violation
    assert got.row_violation == ref.row_violation
    assert bool(got.vacuous) == ref.vacuous
    if ref.vacuous:
        assert got.bound == ref.bound
    else:
        assert abs(got.bound - ref.bound) <= 1e-14 * max(1.0, abs(ref.bound))
    assert abs(got.objective - ref.objective) <= 1e-14 * max(1.0, abs(ref.objective))
    return got, ref


@pytest.mark.parametrize("make", [lambda: small_lp(0), lambda: uneven_batch_member(1002),
                                  lambda: uneven_batch_member(1010)])
def test_certificate_matches_the_oracle(make):
    print("[gpu] certify: certificate vs oracle", flush=True)
    k = make()
    rng = np.random.default_rng(11)
    got, _ = _check_against_oracle(k.problem, k.x_star, k.y_star)
    assert got.bound <= k.f_star and k.f_star - got.bound <= 1e-13 * max(1.0, abs(k.f_star))
    for scale in (1e-9, 1e-2):
        y = k.y_star + scale * rng.standard_normal(k.problem.m)
        x = np.clip(k.x_star + scale * rng.standard_normal(k.problem.n), -1.0, None)
        got, _ = _check_against_oracle(k.problem, x, y)
        assert got.bound <= k.f_star + 1e-12 * abs(k.f_star)


def test_certificate_edge_cases_match_the_oracle():
    print("[gpu] certify: vacuous, absorbed, steep, maximise, skewed", flush=True)
    free = Problem(A=SparseMatrix.from_dense(np.array([[1.0, 1.0]])), c=np.array([1.0, 2.0]),
                   row_lb=np.array([1.0]), row_ub=np.array([1.0]),
                   col_lb=np.array([0.0, -INF]), col_ub=np.array([INF, INF]))
    got, _ = _check_against_oracle(free, np.array([1.0, 0.0]), np.array([5.0]))
    assert got.bound == -np.inf and got.perturbation == 4.0 and got.vacuous
    q = Problem(A=SparseMatrix.from_dense(np.array([[1.0, 1.0]])), c=np.array([1.0, 1.0]),
                row_lb=np.array([1.0]), row_ub=np.array([1.0]),
                col_lb=np.array([0.0, -INF]), col_ub=np.array([INF, INF]))
    y = np.array([1.0 - 2.0 ** -40])
    got, _ = _check_against_oracle(q, np.array([1.0, 0.0]), y)
    assert not got.vacuous and got.perturbation == 2.0 ** -40
    got, _ = _check_against_oracle(q, np.array([1.0, 0.0]), y, dual_tol=1e-20)
    assert got.vacuous
    steep = Problem(A=SparseMatrix.from_triplets(np.array([0]), np.array([0]), np.array([1e6]),
                                                 1, 1),
                    c=np.array([1e8]), row_lb=np.array([1e-4]), row_ub=np.array([INF]),
                    col_lb=np.zeros(1), col_ub=np.full(1, INF))
    _check_against_oracle(steep, np.array([1e-10]), np.array([100.0]))
    _check_against_oracle(steep, np.array([1e-10 - 5e-13]), np.array([100.0]))
    k = uneven_batch_member(1001)
    p = k.problem
    mx = Problem(A=p.A, c=-p.c, row_lb=p.row_lb, row_ub=p.row_ub, col_lb=p.col_lb,
                 col_ub=p.col_ub, obj_offset=2.5, sense=ObjSense.MAXIMISE)
    got, _ = _check_against_oracle(mx, k.x_star, -k.y_star)
    assert got.bound >= -k.f_star + 2.5
    # an LP with columns of up to 5000 nonzeros (one lane walks each column)
    rng = np.random.default_rng(5)
    deg = skewed_lp_degrees(n=20_000, total=200_000)
    cols = np.repeat(np.arange(deg.size), deg)
    A = SparseMatrix.from_triplets(rng.integers(0, 4000, cols.size), cols,
                                   rng.standard_normal(cols.size), 4000, deg.size)
    big = Problem(A=A, c=rng.standard_normal(deg.size), row_lb=np.full(4000, -1.0),
                  row_ub=np.ones(4000), col_lb=np.zeros(deg.size), col_ub=np.full(deg.size, 3.0))
    _check_against_oracle(big, rng.uniform(0, 3, deg.size), rng.standard_normal(4000))


def test_verify_accepts_a_gpu_pdlp_optimum_and_separates_the_other_cases():
    print("[gpu] certify: verify verdicts", flush=True)
    from sovlp.lp.pdlp import PDLPParams, solve_pdlp
    cert = _cert()
    k = small_lp(1)
    p = k.problem
    s = solve_pdlp(p, PDLPParams(eps_abs=1e-9, eps_rel=1e-9, device="gpu"))
    assert s.status == Status.OPTIMAL
    v = cert.verify(p, s.x, s.objective, y=s.y, opt_tol=1e-7)
    assert v.ok, v.report()
    word, detail = v.headline()
    assert word == "ACCEPTED" and "certified optimal" in detail
    assert v.certified_bound <= k.f_star + 1e-9 * abs(k.f_star)
    # the same duals and another feasible point, the optimum of a perturbed
    # cost (feasible, and worse for the true cost): not certified
    rng = np.random.default_rng(2)
    q = Problem(A=p.A, c=p.c + 3.0 * rng.standard_normal(p.n) * (1.0 + np.abs(p.c)),
                row_lb=p.row_lb, row_ub=p.row_ub, col_lb=p.col_lb, col_ub=p.col_ub)
    sw = solve_pdlp(q, PDLPParams(eps_abs=1e-9, eps_rel=1e-9, device="gpu"))
    assert sw.status == Status.OPTIMAL
    w = sw.x
    worse = float(p.c @ w)
    assert worse > k.f_star + 1e-6 * abs(k.f_star)
    v = cert.verify(p, w, worse, y=s.y, opt_tol=1e-7)
    assert all(c[1] for c in v.checks if c[0] != "optimality")
    assert not [c for c in v.checks if c[0] == "optimality"][0][1]
    assert v.headline()[0] == "FEASIBLE"
    # pushed off the rows and bounds: rejected
    word, detail = cert.verify(p, s.x - 1e3, y=s.y).headline()
    assert word == "REJECTED" and ("row constraints" in detail or "column bounds" in detail)
    # a point below a valid bound is infeasible at the bound's resolution
    steep = Problem(A=SparseMatrix.from_triplets(np.array([0]), np.array([0]), np.array([1e6]),
                                                 1, 1),
                    c=np.array([1e8]), row_lb=np.array([1e-4]), row_ub=np.array([INF]),
                    col_lb=np.zeros(1), col_ub=np.full(1, INF))
    good = cert.verify(steep, np.array([1e-10]), 1e-2, y=np.array([100.0]))
    assert good.ok, good.report()
    v = cert.verify(steep, np.array([1e-10 - 5e-13]), 1e8 * (1e-10 - 5e-13), y=np.array([100.0]))
    checks = {c[0]: c for c in v.checks}
    assert checks["row constraints"][1] and not checks["optimality"][1]
    assert "BELOW" in checks["optimality"][2]
    # certified_bound, the reference's signature
    b, pert = cert.certified_bound(p, s.x, s.y)
    assert pert == 0.0 and b <= k.f_star + 1e-9 * abs(k.f_star)
    assert k.f_star - b <= 1e-6 * abs(k.f_star)


def test_long_lines_listed_over_chunks_of_many_rounds():
    """More than 262,144 lines (1024 workgroups x 256 lanes), so each workgroup
    of the listing counts and writes its chunk in several rounds: long rows
    (and columns) scattered over the chunks, some adjacent, one at each end.
    The list comes out in index order; the compensated residual and the
    certificate are the oracle's bit for bit."""
    import ctypes

    import torch
    from sovlp import _lib
    print("[gpu] certify: 300k-line listing", flush=True)
    rng = np.random.default_rng(21)
    m, n = 300_000, 4_000
    lens = rng.integers(0, 6, m)
    longs = np.unique(np.concatenate([[0, 1, 2, m - 1], rng.choice(m, 60, replace=False),
                                      np.arange(150_000, 150_005)]))
    lens[longs] = rng.integers(129, 3000, longs.size)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ri = rng.integers(0, n, rp[-1]).astype(np.int32)
    rx = rng.standard_normal(rp[-1]) * 10.0 ** rng.integers(-4, 5, rp[-1])
    x = rng.standard_normal(n)
    b = rng.standard_normal(m)
    dev = [torch.from_numpy(a).cuda() for a in (rp, ri, rx, x, b)]
    out = torch.empty(m, dtype=torch.float64, device="cuda")
    nb = ctypes.c_int64()
    _lib.call("sovlp_residual_workspace_bytes", m, ctypes.byref(nb))
    ws = torch.zeros(int(nb.value) // 4, dtype=torch.int32, device="cuda")
    _lib.call("sovlp_residual_f64", m, *[a.data_ptr() for a in dev], out.data_ptr(),
              ws.data_ptr(), int(nb.value), torch.cuda.current_stream().cuda_stream)
    w = ws.cpu().numpy()
    assert w[0] == longs.size
    assert np.array_equal(w[4: 4 + longs.size], longs)
    _same_bits(out.cpu().numpy(), oracle.residual(rp, ri, rx, x, b))
    # the certificate of an LP whose CSC has 300k columns, long ones scattered
    # (the transpose of the residual's matrix: its long rows become long columns)
    A = SparseMatrix.from_triplets(ri.astype(np.int64), np.repeat(np.arange(m), lens), rx, n, m)
    prob = Problem(A=A, c=rng.standard_normal(m), row_lb=np.full(n, -1.0), row_ub=np.ones(n),
                   col_lb=np.zeros(m), col_ub=np.full(m, 5.0))
    xs = rng.uniform(0.0, 5.0, m)
    ys = rng.standard_normal(n)
    _check_against_oracle(prob, xs, ys)
