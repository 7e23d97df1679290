"""The solve entry point callers use (reference ``src/sovopt/cli.py:82-86, 157-182``).

``solve(prob, method="pdlp", device="gpu", time_limit, gap, tol, verbose)``
maps ``tol`` to ``eps_abs = eps_rel = tol`` and runs ``solve_pdlp``; ``"auto"``
means PDLP here.  ``crossover=True`` then crosses the PDLP point over to an
optimal basis (``sovlp.lp.crossover``, simplex tolerances ``max(tol, 1e-9)``
as the reference sets them), and ``method="simplex"`` runs the host simplex
alone.  ``solve_file`` reads a model file first (``sovlp.io.read_model``), so
``solve_file("model.mps", method="pdlp", device="gpu")`` is the library form
of ``sovopt solve model.mps --method pdlp --device gpu``.
"""

from __future__ import annotations

from .core.problem import Problem, Solution

__all__ = ["solve", "solve_file"]


def solve(prob: Problem, method: str = "auto", device: str = "auto",
          time_limit: float = 300.0, gap: float = 1e-4, tol: float = 1e-8,
          verbose: bool = False, crossover: bool = False, sensitivity: bool = False,
          **options) -> Solution:
    """Solve an LP with PDLP on the GPU (``crossover``: then cross over to an
    optimal basis), or with the host simplex (``method="simplex"``).  ``gap``
    is accepted for the reference's signature (it applies to branch-and-bound,
    not to an LP); ``sensitivity`` adds the ranging of the simplex's or the
    crossover's optimal basis.  The reference's other methods and options
    (``ipm``, ``presolve``) are not provided and raise."""
    from .lp.pdlp import PDLPParams, solve_pdlp
    from .lp.simplex import SimplexParams, solve_simplex
    if method not in ("auto", "pdlp", "simplex"):
        raise NotImplementedError(f"method {method!r} is not provided ('pdlp' or 'simplex')")
    unsupported = [k for k, v in options.items() if v]
    if unsupported:
        raise NotImplementedError(f"options not provided: {', '.join(sorted(unsupported))}")
    sp = SimplexParams(time_limit=time_limit, feas_tol=max(tol, 1e-9),
                       opt_tol=max(tol, 1e-9), sensitivity=sensitivity, verbose=verbose)
    if method == "simplex":
        return solve_simplex(prob, sp)
    point = solve_pdlp(prob, PDLPParams(device=device, eps_abs=tol, eps_rel=tol,
                                        time_limit=time_limit, verbose=verbose))
    if not crossover or point.x is None:
        return point
    from .lp.crossover import crossover as _crossover
    return _crossover(prob, point.x, point.y, sp)


def solve_file(path, **kwargs) -> Solution:
    """``solve(read_model(path), **kwargs)``."""
    from .io import read_model
    return solve(read_model(path), **kwargs)
