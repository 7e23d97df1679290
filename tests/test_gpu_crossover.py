"""Crossover from device PDLP solves to optimal bases (the path's next step).

``solve(..., method="pdlp", device="gpu", crossover=True)`` as the reference's
``solve`` runs it (``src/sovopt/cli.py:167-182``): the PDLP point from the GPU,
then the identified basis and the simplex on the host.  Pinned on the LPs'
closed-form optima, on the reference's own figure for the hourly blending
plan (``tests/test_simplex.py:230-245``), on the optimality conditions of the
returned basis, and on the reference's claim for crossover: fewer pivots than
a cold simplex (``tests/test_crossover.py:150-172``).
"""

import numpy as np
import pytest

from sovlp.core.problem import Status

from test_crossover import MODELS, assert_optimal_basis

pytestmark = pytest.mark.gpu


def test_crossover_from_the_500_x_2000_lp():
    from sovlp.lp.simplex import solve_simplex
    from sovlp.solve import solve
    from sovlp.workloads import small_lp
    print("[gpu] crossover: 500 x 2000, PDLP at 1e-8 then crossover", flush=True)
    ko = small_lp(0)
    p = ko.problem
    s = solve(p, method="pdlp", device="gpu", tol=1e-8, crossover=True)
    cold = solve_simplex(p)
    print(f"[gpu] crossover {s.status.name} {s.iterations} pivots vs cold {cold.iterations}",
          flush=True)
    assert s.status == Status.OPTIMAL and s.info["crossover"]
    assert s.info["repairs"] == 0
    assert abs(s.objective - ko.f_star) <= 1e-9 * max(1.0, abs(ko.f_star))
    assert_optimal_basis(p, s)
    assert s.iterations < cold.iterations


def test_crossover_from_batched_fp32_solutions():
    """Points from the batched FP32 + refinement engine cross over too."""
    from sovlp.lp.batch import solve_batch
    from sovlp.lp.crossover import crossover
    from sovlp.lp.pdlp import PDLPParams
    from sovlp.workloads import uneven_batch
    print("[gpu] crossover: 8 batched FP32 solutions", flush=True)
    kos = uneven_batch(8, base_seed=7000)
    res = solve_batch([k.problem for k in kos], PDLPParams(eps_abs=1e-6, eps_rel=1e-6,
                                                           feas_cap=1e-6), precision=32)
    for ko, s in zip(kos, res.solutions):
        if s.status != Status.OPTIMAL:
            continue
        xo = crossover(ko.problem, s.x, s.y)
        assert xo.status == Status.OPTIMAL
        assert abs(xo.objective - ko.f_star) <= 1e-9 * max(1.0, abs(ko.f_star))
        assert_optimal_basis(ko.problem, xo)


def test_crossover_finishes_the_hourly_blending_plan():
    """PDLP does not converge on blend_plan_6h (654 x 1104, a maximisation):
    at a 20k-iteration cap it stops 6e-4 short of the optimum.  The crossover
    from that point reaches the reference simplex's optimum, 6421.520443, with
    an optimal basis, in fewer pivots than a cold simplex."""
    from sovlp.io import read_model
    from sovlp.lp.crossover import crossover
    from sovlp.lp.pdlp import PDLPParams, solve_pdlp
    from sovlp.lp.simplex import SimplexParams, solve_simplex
    print("[gpu] crossover: blend_plan_6h from a 20k-iteration PDLP point", flush=True)
    p = read_model(MODELS + "blend_plan_6h.mps.gz")
    s = solve_pdlp(p, PDLPParams(device="gpu", eps_abs=1e-8, eps_rel=1e-8, max_iter=20_000))
    assert s.status == Status.ITERATION_LIMIT and s.x is not None
    sp = SimplexParams(feas_tol=1e-9, opt_tol=1e-9)
    xo = crossover(p, s.x, s.y, sp)
    cold = solve_simplex(p, sp)
    print(f"[gpu] blend: PDLP point {p.objective(s.x):.6f}; crossover {xo.status.name} "
          f"{xo.objective:.9f} in {xo.iterations} iterations, cold {cold.iterations}", flush=True)
    assert xo.status == Status.OPTIMAL
    assert abs(xo.objective - 6421.520443) <= 5e-7
    assert_optimal_basis(p, xo, tol=1e-9)
    assert xo.iterations < cold.iterations
