"""Write golden vectors of the reference's batched node relaxation.

Run in the development container only (the reference tree is not present on
the GPU machines):

    SOVOPT_FETCH=0 python tests/golden/make_bnr_vectors.py

It imports the reference package from ``/root/reference/src`` and runs
``BNREngine.solve_batch`` (``src/sovopt/mip/bnr.py:176-301``, with
``safe_dual_bound_batch``, ``src/sovopt/mip/safebound.py:194-235``) on its CPU
backend, on tiny LPs built here from seeds with a few node boxes each: a cold
visit, then a warm second visit of the same nodes with the step sizes the
engine kept.  The output, ``reference_bnr.json``, holds data only: each case's
problem in triplet form, the node boxes, and the reference's X, Y, step sizes,
bounds and empty flags after each visit.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def _cases():
    """(name, problem arrays, node boxes, iterations per visit)."""
    sys.path.insert(0, os.path.join(ROOT, "sovereign-moe_amd"))
    from sovlp.workloads import known_optimum_lp
    out = []
    for (m, n, d, seed, m_eq, K, iters) in ((6, 10, 2, 3, 3, 1, 5), (12, 24, 3, 5, 6, 3, 8),
                                             (20, 40, 3, 7, 20, 5, 12), (15, 30, 3, 9, 5, 4, 6)):
        ko = known_optimum_lp(m, n, d, seed, m_eq=m_eq, ub=10.0)
        p = ko.problem
        rng = np.random.default_rng(100 + seed)
        L = np.repeat(p.col_lb[:, None], K, 1)
        U = np.repeat(p.col_ub[:, None], K, 1)
        for q in range(1, K):
            js = rng.choice(n, 3, replace=False)
            U[js, q] = np.minimum(U[js, q], rng.uniform(0.0, 5.0, 3))
            ls = rng.choice(n, 2, replace=False)
            L[ls, q] = np.minimum(U[ls, q], rng.uniform(0.0, 2.0, 2))
        if K >= 4:                       # one empty box
            L[0, K - 1], U[0, K - 1] = 3.0, 2.0
        cols = np.repeat(np.arange(n), np.diff(p.A.cp))
        out.append(dict(name=f"bnr_{m}x{n}_K{K}_s{seed}", m=m, n=n, K=K, iters=iters,
                        rows=p.A.ci.tolist(), cols=cols.tolist(), vals=p.A.cx.tolist(),
                        c=p.c.tolist(), rl=p.row_lb.tolist(), ru=p.row_ub.tolist(),
                        L=L.tolist(), U=U.tolist()))
    return out


def main():
    os.environ["SOVOPT_FETCH"] = "0"
    sys.path.insert(0, REF_SRC)
    from sovopt.core.problem import Problem
    from sovopt.core.sparse import SparseMatrix
    from sovopt.mip.bnr import BNRConfig, BNREngine

    class _Scaling:            # the engine reads only the objective factor
        obj = 1.0

    results = []
    for case in _cases():
        m, n, K = case["m"], case["n"], case["K"]
        A = SparseMatrix.from_triplets(np.array(case["rows"]), np.array(case["cols"]),
                                       np.array(case["vals"], dtype=float), m, n)
        prob = Problem(A=A, c=np.array(case["c"], float), row_lb=np.array(case["rl"], float),
                       row_ub=np.array(case["ru"], float), col_lb=np.zeros(n),
                       col_ub=np.full(n, 10.0), name=case["name"])
        L = np.array(case["L"], float)
        U = np.array(case["U"], float)
        eng = BNREngine(prob, _Scaling(), BNRConfig(batch=K, iters=case["iters"], device="cpu"))
        visits = []
        warm = None
        for _ in range(2):
            r = eng.solve_batch(L, U, warm=warm)
            X, Y = eng.X[:, :K].copy(), eng.Y[:, :K].copy()
            visits.append(dict(X=X.tolist(), Y=Y.tolist(), eta=eng.eta[:K].tolist(),
                               omega=eng.omega[:K].tolist(),
                               bounds=[b if np.isfinite(b) else (str(b)) for b in
                                       r.bounds.tolist()],
                               empty=[bool(e) for e in r.infeasible]))
            warm = (X, Y)
        rec = dict(case, visits=visits)
        results.append(rec)
        print(f"{case['name']:>22s}  bounds {visits[0]['bounds']} -> {visits[1]['bounds']}",
              flush=True)
    with open(os.path.join(HERE, "reference_bnr.json"), "w") as fh:
        json.dump(dict(source="sovopt BNREngine.solve_batch (CPU backend), two visits",
                       cases=results), fh)


if __name__ == "__main__":
    main()
