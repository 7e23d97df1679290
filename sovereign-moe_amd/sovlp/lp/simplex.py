"""The revised simplex the crossover finishes with (host code, ``libsovlp.so``).

Reference: ``solve_simplex(prob, params, warm_basis)``
(``src/sovopt/lp/simplex.py:1181``) and ``SimplexParams``
(``simplex.py:94-187``).  Same call, same ``Solution`` fields (x, row duals
``y`` with ``c - Aᵀy`` the reduced costs, ``basis_status``, iterations), same
status codes.  The native simplex (``crossover.cpp``) is a bounded primal
revised simplex over ``[A | -I]``: phase 1 on the sum of infeasibilities,
Dantzig pricing, a Harris two-pass ratio test with bound flips, a random cost
perturbation after ``perturb_after`` degenerate pivots (removed before the end),
a dense LU of the basis refactorised every ``refactor_freq`` pivots with
product-form updates between.  Dense factors limit it to 8192 rows.
``sensitivity=True`` adds the ranging of the optimal basis
(``sovlp.lp.sensitivity``).  The reference's dual simplex, node kernel,
Forrest-Tomlin updates and scaling are not provided.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..core.problem import ObjSense, Problem, Solution, Status
from .basis import host_lp

__all__ = ["SimplexParams", "solve_simplex", "result_solution"]


@dataclass
class SimplexParams:
    max_iter: int = 1_000_000
    time_limit: float = 300.0
    feas_tol: float = 1e-7
    opt_tol: float = 1e-7
    pivot_tol: float = 1e-7
    refactor_freq: int = 150
    perturb_after: int = 500
    perturb_scale: float = 1e-7
    sensitivity: bool = False
    verbose: bool = False

    def to_c(self) -> _lib.sovlp_simplex_params:
        return _lib.sovlp_simplex_params(int(self.max_iter), float(self.time_limit),
                                         float(self.feas_tol), float(self.opt_tol),
                                         float(self.pivot_tol), int(self.refactor_freq),
                                         int(self.perturb_after), float(self.perturb_scale))


def solve_simplex(prob: Problem, params: SimplexParams | None = None,
                  warm_basis=None) -> Solution:
    """Solve an LP with the revised simplex from ``warm_basis`` (``n + m``
    statuses with exactly ``m`` BASIC) or from the all-logical basis."""
    params = params or SimplexParams()
    t0 = time.perf_counter()
    if prob.Q is not None:
        raise NotImplementedError("solve_simplex optimises a linear objective; this model "
                                  "has a quadratic term")
    flip = prob.sense == ObjSense.MAXIMISE
    c = -prob.c if flip else prob.c
    h = host_lp(prob, c)
    m, n = prob.m, prob.n
    warm = None
    if warm_basis is not None:
        warm = np.ascontiguousarray(warm_basis, dtype=np.uint8)
        if warm.shape != (n + m,):
            raise ValueError(f"warm_basis has {warm.size} entries, the LP n + m = {n + m}")
    x = np.zeros(n)
    y = np.zeros(m)
    d = np.zeros(n)
    st = np.zeros(n + m, dtype=np.uint8)
    res = _lib.sovlp_simplex_result()
    cp = params.to_c()
    _lib.call("sovlp_simplex_solve", h.ref(), C.byref(cp),
              None if warm is None else warm.ctypes.data, x.ctypes.data, y.ctypes.data,
              d.ctypes.data, st.ctypes.data, C.byref(res))
    return result_solution(prob, res, x, y, d, st, t0, warm is not None,
                           sensitivity=params.sensitivity)


def result_solution(prob: Problem, res, x, y, d, st, t0: float, warm: bool,
                    extra: dict | None = None, sensitivity: bool = False) -> Solution:
    """The ``Solution`` of a native simplex result (duals in the caller's
    sense: the native code minimises, so a maximisation flips y and d), with
    the ranging of an optimal basis when ``sensitivity``."""
    status = Status(int(res.status))
    if prob.sense == ObjSense.MAXIMISE:
        y, d = -y, -d
    sol = Solution(status=status, x=x, objective=float(prob.c @ x + prob.obj_offset), y=y,
                   reduced_costs=d, basis_status=st, iterations=int(res.iterations),
                   time=time.perf_counter() - t0, method="simplex")
    sol.dual_bound = sol.objective if status == Status.OPTIMAL else float("nan")
    sol.work_units = float(res.iterations)
    sol.info = {"pivots": int(res.pivots), "phase1_iterations": int(res.phase1_iterations),
                "repairs": int(res.repairs), "primal_infeasibility": res.primal_infeasibility,
                "dual_infeasibility": res.dual_infeasibility, "native_seconds": res.seconds,
                "warm": bool(warm), **(extra or {})}
    if sensitivity and status == Status.OPTIMAL:
        from .sensitivity import compute_sensitivity
        sol.sensitivity = compute_sensitivity(prob, st)
    return sol.drop_objective_if_unsolved()
