"""Boundary types: ``Problem``, ``Solution``, ``Status``, ``ObjSense``.

Field names, enum values and semantics follow the reference
(``src/sovopt/core/problem.py:45-343``) so a caller of SOVOPT's PDLP path can
hand the same objects to this engine:

* canonical row-bounds LP ``min cᵀx + c0  s.t.  rl <= Ax <= ru,  l <= x <= u``
  with ``INF = 1e30`` marking absent bounds;
* ``Status`` integer codes ``OPTIMAL=0 ... NUMERICAL=8 ... NOT_SOLVED=10``
  (``problem.py:58-69``) and ``has_solution`` (``problem.py:75-78``);
* ``Solution.drop_objective_if_unsolved`` erases the objective of a status that
  carries no point (``problem.py:311-328``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

from .sparse import SparseMatrix, VAL
from .tolerances import INF

__all__ = ["Problem", "VarKind", "ObjSense", "Status", "Solution"]


class VarKind(IntEnum):
    CONTINUOUS = 0
    INTEGER = 1
    BINARY = 2
    SEMI_CONTINUOUS = 3
    SEMI_INTEGER = 4


class ObjSense(IntEnum):
    MINIMISE = 1
    MAXIMISE = -1


class Status(IntEnum):
    OPTIMAL = 0
    INFEASIBLE = 1
    UNBOUNDED = 2
    INFEASIBLE_OR_UNBOUNDED = 3
    ITERATION_LIMIT = 4
    TIME_LIMIT = 5
    NODE_LIMIT = 6
    GAP_LIMIT = 7
    NUMERICAL = 8
    INTERRUPTED = 9
    NOT_SOLVED = 10

    @property
    def is_solved(self) -> bool:
        return self is Status.OPTIMAL

    @property
    def has_solution(self) -> bool:
        return self in (Status.OPTIMAL, Status.ITERATION_LIMIT, Status.TIME_LIMIT,
                        Status.NODE_LIMIT, Status.GAP_LIMIT)


@dataclass
class Problem:
    """An LP in canonical row-bounds form (quadratic and integer data are carried
    for interface compatibility; the PDLP engine refuses both)."""

    A: SparseMatrix
    c: np.ndarray
    row_lb: np.ndarray
    row_ub: np.ndarray
    col_lb: np.ndarray
    col_ub: np.ndarray
    kind: np.ndarray = None
    Q: SparseMatrix | None = None
    obj_offset: float = 0.0
    sense: ObjSense = ObjSense.MINIMISE
    name: str = "model"
    col_names: list | None = None
    row_names: list | None = None
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        for attr in ("c", "row_lb", "row_ub", "col_lb", "col_ub"):
            setattr(self, attr, np.ascontiguousarray(getattr(self, attr), dtype=VAL))
        if self.kind is None:
            self.kind = np.zeros(self.n, dtype=np.uint8)
        else:
            self.kind = np.ascontiguousarray(self.kind, dtype=np.uint8)
        m, n = self.A.shape
        for arr, size, what in ((self.c, n, "c"), (self.col_lb, n, "col_lb"),
                                (self.col_ub, n, "col_ub"), (self.kind, n, "kind"),
                                (self.row_lb, m, "row_lb"), (self.row_ub, m, "row_ub")):
            if arr.shape[0] != size:
                raise ValueError(f"{what} has length {arr.shape[0]}, expected {size}")

    @property
    def m(self) -> int:
        return self.A.m

    @property
    def n(self) -> int:
        return self.A.n

    @property
    def nnz(self) -> int:
        return self.A.nnz

    @property
    def is_mip(self) -> bool:
        return bool((self.kind != VarKind.CONTINUOUS).any())

    @property
    def is_qp(self) -> bool:
        return self.Q is not None

    @property
    def integer_mask(self) -> np.ndarray:
        """Columns that must take integer values (core/problem.py:151-153)."""
        return (self.kind == VarKind.INTEGER) | (self.kind == VarKind.BINARY)

    def row_types(self):
        lo, hi = self.row_lb, self.row_ub
        flo, fhi = lo > -INF, hi < INF
        eq = int((flo & fhi & (hi - lo <= 1e-12)).sum())
        rng = int((flo & fhi & (hi - lo > 1e-12)).sum())
        return eq, rng, int((~flo & fhi).sum()), int((flo & ~fhi).sum()), int((~flo & ~fhi).sum())

    def objective(self, x) -> float:
        return float(self.c @ np.asarray(x, dtype=VAL)) + self.obj_offset

    def violation(self, x):
        """``(row violation, bound violation, integrality violation)`` maxima."""
        x = np.asarray(x, dtype=VAL)
        act = self.A.matvec(x)
        row_v = max(float(np.maximum(self.row_lb - act, 0.0).max(initial=0.0)),
                    float(np.maximum(act - self.row_ub, 0.0).max(initial=0.0)))
        col_v = max(float(np.maximum(self.col_lb - x, 0.0).max(initial=0.0)),
                    float(np.maximum(x - self.col_ub, 0.0).max(initial=0.0)))
        mask = (self.kind == VarKind.INTEGER) | (self.kind == VarKind.BINARY)
        int_v = float(np.abs(x[mask] - np.round(x[mask])).max(initial=0.0)) if mask.any() else 0.0
        return row_v, col_v, int_v

    def summary(self) -> str:
        eq, rng, le, ge, free = self.row_types()
        lo, hi = self.A.coeff_range()
        ratio = hi / lo if lo > 0 else float("inf")
        return (f"{self.name}  [LP]\n"
                f"  {self.m:,} rows x {self.n:,} cols, {self.nnz:,} nonzeros "
                f"(density {self.A.density:.2e})\n"
                f"  rows: {eq:,} eq, {le:,} <=, {ge:,} >=, {rng:,} ranged, {free:,} free\n"
                f"  |a_ij| in [{lo:.3e}, {hi:.3e}]  ratio {ratio:.2e}")

    def copy(self) -> "Problem":
        return Problem(A=self.A.copy(), c=self.c.copy(), row_lb=self.row_lb.copy(),
                       row_ub=self.row_ub.copy(), col_lb=self.col_lb.copy(),
                       col_ub=self.col_ub.copy(), kind=self.kind.copy(),
                       Q=self.Q.copy() if self.Q is not None else None,
                       obj_offset=self.obj_offset, sense=self.sense, name=self.name,
                       col_names=list(self.col_names) if self.col_names else None,
                       row_names=list(self.row_names) if self.row_names else None,
                       meta=dict(self.meta))


@dataclass
class Solution:
    status: Status
    x: np.ndarray | None = None
    objective: float = float("nan")
    y: np.ndarray | None = None
    reduced_costs: np.ndarray | None = None
    basis_status: np.ndarray | None = None
    dual_bound: float = float("nan")
    nodes: int = 0
    gap_rel: float = float("nan")
    iterations: int = 0
    time: float = 0.0
    work_units: float = 0.0
    method: str = ""
    log: list = field(default_factory=list)
    sensitivity: object = None
    farkas: np.ndarray | None = None
    info: object = None

    def drop_objective_if_unsolved(self) -> "Solution":
        if not self.status.has_solution:
            self.objective = float("nan")
            self.dual_bound = float("nan")
        return self

    @property
    def gap(self) -> float:
        if not np.isfinite(self.objective) or not np.isfinite(self.dual_bound):
            return float("inf")
        return abs(self.objective - self.dual_bound) / max(1.0, abs(self.objective))

    def __repr__(self):
        obj = f"{self.objective:.10g}" if np.isfinite(self.objective) else "-"
        return f"Solution({self.status.name}, obj={obj}, iters={self.iterations}, {self.time:.3f}s)"
