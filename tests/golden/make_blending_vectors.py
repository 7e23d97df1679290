"""Write golden objectives of small multi-period blending LPs with SciPy's HiGHS.

    python tests/golden/make_blending_vectors.py           (small instances)
    python tests/golden/make_blending_vectors.py --spec    (served-workload sizes)

The blending LPs (``sovlp.workloads.blending_lp``) and their 1 % resolves
(``sovlp.workloads.perturbed``) have no closed-form optimum, so an independent
LP solver pins them: HiGHS (dual simplex / IPM with crossover, SciPy 1.15) solves
each instance here and the objective is stored. The output,
``blending_highs.json``, holds the generator arguments and the objectives only;
the tests rebuild each LP from its seed. ``--spec`` writes
``blending_spec_highs.json``: instances inside the served workload's ranges
(K >= 20 components, P >= 5 products, T = 100-120 hours) and a 1 % resolve.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "sovereign-moe_amd"))

from sovlp.core.tolerances import INF  # noqa: E402
from sovlp.workloads import blending_lp, perturbed  # noqa: E402


def highs_objective(p) -> float:
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix
    A = csr_matrix((p.A.rx, p.A.ri, p.A.rp), shape=(p.m, p.n))
    eq = p.row_lb == p.row_ub
    le = ~eq
    assert np.all(p.row_lb[le] <= -INF), "only equality and <= rows are generated"
    bounds = list(zip(p.col_lb, np.where(p.col_ub >= INF, None, p.col_ub)))
    res = linprog(p.c, A_ub=A[le], b_ub=p.row_ub[le], A_eq=A[eq], b_eq=p.row_ub[eq],
                  bounds=bounds, method="highs")
    assert res.status == 0, res.message
    return float(res.fun) + p.obj_offset


def cases():
    out = []
    for seed in range(8):
        rng = np.random.default_rng(1000 + seed)
        K, P, T = int(rng.integers(4, 13)), int(rng.integers(2, 5)), int(rng.integers(4, 25))
        out.append(dict(seed=seed, K=K, P=P, T=T, perturb_seed=None))
        out.append(dict(seed=seed, K=K, P=P, T=T, perturb_seed=77 + seed))
    return out


def spec_cases():
    """Instances inside the served workload's ranges (K 20-60, P 5-20, T 24-720)
    that PDLP finishes in seconds on one GPU, and a 1 % resolve of the first."""
    return [dict(seed=11, K=20, P=5, T=100, perturb_seed=None),
            dict(seed=11, K=20, P=5, T=100, perturb_seed=311),
            dict(seed=13, K=22, P=6, T=120, perturb_seed=None)]


def build(case):
    p = blending_lp(case["seed"], K=case["K"], P=case["P"], T=case["T"])
    if case["perturb_seed"] is not None:
        p = perturbed(p, case["perturb_seed"], 0.01)
    return p


def main():
    spec = "--spec" in sys.argv[1:]
    res = []
    for c in (spec_cases() if spec else cases()):
        p = build(c)
        c = dict(c, m=p.m, n=p.n, nnz=int(p.A.nnz), objective=highs_objective(p))
        res.append(c)
        print(c, flush=True)
    out = "blending_spec_highs.json" if spec else "blending_highs.json"
    with open(os.path.join(HERE, out), "w") as f:
        json.dump({"solver": "scipy.optimize.linprog(method='highs')", "cases": res}, f, indent=1)


if __name__ == "__main__":
    main()
