"""Write golden input/output vectors from the reference's CPU PDLP.

Run in the development container only (the reference tree is not present on
the GPU machines):

    SOVOPT_FETCH=0 python tests/golden/make_reference_vectors.py

It imports the reference package from ``/root/reference/src`` and records, for
a handful of tiny LPs built here from seeds, what only running the reference
can pin: the ``scale_problem(..., "pdlp")`` factors and the ``solve_pdlp``
trajectory (iterations, restarts, rejected steps, final step size and primal
weight, history) with the returned ``x``, ``y`` and objective. The output,
``reference_pdlp.json``, holds data only: each case's problem in triplet form
and the reference's outputs.
"""

from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    """Problems as plain arrays: (name, m, n, rows, cols, vals, c, rl, ru, lo, hi,
    offset, sense, params)."""
    INF = 1e30
    out = []
    # toy max LP: max 3x+2y, x+y<=4, x+3y<=6, x<=3
    out.append(dict(name="toy_max", m=3, n=2, rows=[0, 1, 2, 0, 1], cols=[0, 0, 0, 1, 1],
                    vals=[1, 1, 1, 1, 3], c=[3, 2], rl=[-INF] * 3, ru=[4, 6, 3],
                    lo=[0, 0], hi=[INF, INF], offset=0.0, sense=-1, params={}))
    # testprob.mps: min x1+2x2+3x3; x1+x2<=4; x1+x3>=1; -x2+x3=7; x1<=4; -1<=x2<=1
    out.append(dict(name="testprob", m=3, n=3, rows=[0, 1, 0, 2, 1, 2], cols=[0, 0, 1, 1, 2, 2],
                    vals=[1, 1, 1, -1, 1, 1], c=[1, 2, 3], rl=[-INF, 1, 7], ru=[4, INF, 7],
                    lo=[0, -1, 0], hi=[4, 1, INF], offset=0.0, sense=1, params={}))
    # known-optimum LPs (same construction as the synthetic workloads), tiny
    sys.path.insert(0, os.path.join(HERE, "..", "..", "sovereign-moe_amd"))
    from sovlp.workloads import known_optimum_lp
    for (m, n, d, seed, tol) in ((12, 30, 3, 5, 1e-8), (30, 80, 4, 6, 1e-6), (40, 100, 5, 7, 1e-8)):
        ko = known_optimum_lp(m, n, d, seed, m_eq=m // 2, ub=10.0)
        p = ko.problem
        cols = np.repeat(np.arange(p.n), np.diff(p.A.cp))
        out.append(dict(name=f"known_{m}x{n}_s{seed}", m=m, n=n, rows=p.A.ci.tolist(),
                        cols=cols.tolist(), vals=p.A.cx.tolist(), c=p.c.tolist(),
                        rl=p.row_lb.tolist(), ru=p.row_ub.tolist(), lo=p.col_lb.tolist(),
                        hi=p.col_ub.tolist(), offset=0.0, sense=1,
                        params=dict(eps_abs=tol, eps_rel=tol), f_star=ko.f_star))
    # badly scaled rows (row magnitudes 1e-4..1e4), objective scale != 1
    rng = np.random.default_rng(1)
    m, n = 30, 40
    cols = np.repeat(np.arange(n), 4)
    rows = rng.integers(0, m, cols.size)
    vals = rng.standard_normal(cols.size) * (10.0 ** rng.integers(-4, 5, size=m).astype(float))[rows]
    x0 = rng.random(n)
    dense = np.zeros((m, n))
    np.add.at(dense, (rows, cols), vals)
    b = dense @ x0 + rng.random(m)
    out.append(dict(name="badly_scaled", m=m, n=n, rows=rows.tolist(), cols=cols.tolist(),
                    vals=vals.tolist(), c=(rng.standard_normal(n) * 37.0).tolist(),
                    rl=[-INF] * m, ru=b.tolist(), lo=[0.0] * n, hi=[10.0] * n, offset=1.5,
                    sense=1, params=dict(max_iter=60000)))
    return out


def main():
    if os.environ.get("SOVOPT_FETCH") != "0":
        os.environ["SOVOPT_FETCH"] = "0"
    sys.path.insert(0, REF_SRC)
    from sovopt.core.problem import ObjSense, Problem
    from sovopt.core.sparse import SparseMatrix
    from sovopt.lp.pdlp import PDLPParams, solve_pdlp
    from sovopt.numerics.scaling import scale_problem

    results = []
    for case in _cases():
        A = SparseMatrix.from_triplets(np.array(case["rows"]), np.array(case["cols"]),
                                       np.array(case["vals"], dtype=float), case["m"], case["n"])
        prob = Problem(A=A, c=np.array(case["c"], float), row_lb=np.array(case["rl"], float),
                       row_ub=np.array(case["ru"], float), col_lb=np.array(case["lo"], float),
                       col_ub=np.array(case["hi"], float), obj_offset=case["offset"],
                       sense=ObjSense(case["sense"]), name=case["name"])
        work = prob
        if prob.sense == ObjSense.MAXIMISE:
            work = prob.copy()
            work.c = -work.c
            work.obj_offset = -work.obj_offset
            work.sense = ObjSense.MINIMISE
        _, sc = scale_problem(work, method="pdlp")
        sol = solve_pdlp(prob, PDLPParams(device="cpu", **case["params"]))
        info = sol.info
        rec = dict(case)
        rec["expected"] = dict(
            scale_row=sc.row.tolist(), scale_col=sc.col.tolist(), scale_obj=sc.obj,
            status=int(sol.status), objective=None if not math.isfinite(sol.objective) else sol.objective,
            iterations=int(sol.iterations), restarts=int(info.restarts),
            rejected_steps=int(info.rejected_steps), step_size=float(info.step_size),
            primal_weight=float(info.primal_weight), x=sol.x.tolist(), y=sol.y.tolist(),
            reduced_costs=sol.reduced_costs.tolist(),
            history=[list(map(float, h)) for h in info.history],
            primal_res=float(info.primal_res), dual_res=float(info.dual_res), gap=float(info.gap))
        results.append(rec)
        print(f"{case['name']:>22s}  status={sol.status.name} obj={sol.objective:.12g} "
              f"it={sol.iterations} restarts={info.restarts} rej={info.rejected_steps}", flush=True)
    with open(os.path.join(HERE, "reference_pdlp.json"), "w") as fh:
        json.dump(dict(source="sovopt solve_pdlp(device='cpu') and scale_problem(method='pdlp')",
                       numba=False, cases=results), fh)


if __name__ == "__main__":
    main()
