"""Ranging of an optimal basis: cost and right-hand-side intervals, shadow prices.

Reference: ``Sensitivity`` and ``compute_sensitivity``
(``src/sovopt/lp/sensitivity.py:79-332``), the report a planner reads off a
basis: each cost's interval over which the basis stays optimal, each binding
row's shadow price and the interval of its bound over which the basis stays
feasible.  Only a basis can answer these, so they come with the simplex and
the crossover (``SimplexParams(sensitivity=True)``), never from PDLP alone.
The computation is native (``sovlp_simplex_ranging``: one BTRAN per basic
structural, one FTRAN per binding row); this module converts it to the
caller's sense.  At a degenerate optimum the intervals are those of the basis
the simplex stopped at, and zero-width ranges are the usual sign of it
(``degenerate_cols``/``degenerate_rows`` count them).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..core.problem import ObjSense, Problem
from ..core.tolerances import INF
from .basis import host_lp

__all__ = ["Sensitivity", "compute_sensitivity"]


@dataclass
class Sensitivity:
    """Ranging report in the model's own sense and units."""

    # per variable
    value: np.ndarray = None
    reduced_cost: np.ndarray = None
    cost: np.ndarray = None
    cost_lo: np.ndarray = None
    cost_hi: np.ndarray = None

    # per row
    activity: np.ndarray = None
    dual: np.ndarray = None
    rhs: np.ndarray = None
    rhs_lo: np.ndarray = None
    rhs_hi: np.ndarray = None
    binding: np.ndarray = None

    col_names: list | None = None
    row_names: list | None = None
    degenerate_cols: int = 0
    degenerate_rows: int = 0

    def binding_rows(self) -> np.ndarray:
        """Rows with a nonzero shadow price, largest magnitude first."""
        idx = np.flatnonzero(np.abs(self.dual) > 1e-9)
        return idx[np.argsort(-np.abs(self.dual[idx]), kind="stable")]

    def report(self, max_rows: int = 20, max_cols: int = 20) -> str:
        """Plain-text tables: shadow prices of the binding rows with their
        right-hand-side ranges, then the cost ranges of the nonzero columns."""
        cols = self.col_names or [f"C{j}" for j in range(self.value.size)]
        rows = self.row_names or [f"R{i}" for i in range(self.dual.size)]

        def num(v: float) -> str:
            return "-inf" if v <= -INF else ("+inf" if v >= INF else f"{v:.6g}")

        def rng(a: float, b: float) -> str:
            return f"[{num(a)}, {num(b)}]"

        lines = ["Shadow prices of binding rows (largest first):",
                 f"  {'row':<20} {'dual':>14} {'activity':>14}  {'rhs range':<32}"]
        picked = self.binding_rows()[:max_rows]
        for i in picked:
            lines.append(f"  {rows[i][:20]:<20} {self.dual[i]:>14.6g} {self.activity[i]:>14.6g}  "
                         f"{rng(self.rhs_lo[i], self.rhs_hi[i]):<32}")
        if picked.size == 0:
            lines.append("  none")
        nz = np.flatnonzero(np.abs(self.value) > 1e-9)
        lines += ["", "Cost ranges of nonzero columns:",
                  f"  {'column':<20} {'value':>14} {'cost':>12} {'reduced':>12}  {'cost range':<32}"]
        for j in nz[:max_cols]:
            lines.append(f"  {cols[j][:20]:<20} {self.value[j]:>14.6g} {self.cost[j]:>12.6g} "
                         f"{self.reduced_cost[j]:>12.6g}  {rng(self.cost_lo[j], self.cost_hi[j]):<32}")
        if nz.size > max_cols:
            lines.append(f"  ({nz.size - max_cols} more)")
        if self.degenerate_cols or self.degenerate_rows:
            lines += ["", f"Degenerate optimum: {self.degenerate_cols} cost ranges and "
                          f"{self.degenerate_rows} right-hand-side ranges have zero width; "
                          f"those prices hold for this basis but are not unique."]
        return "\n".join(lines)


def compute_sensitivity(prob: Problem, basis_status: np.ndarray) -> Sensitivity:
    """The ranging of the optimal basis ``basis_status`` (``n + m`` statuses,
    exactly ``m`` BASIC) of ``prob``, in the model's sense: a maximisation's
    cost intervals, duals and reduced costs come back with its signs."""
    m, n = prob.m, prob.n
    st = np.ascontiguousarray(basis_status, dtype=np.uint8)
    if st.shape != (n + m,):
        raise ValueError(f"basis_status has {st.size} entries, the LP n + m = {n + m}")
    flip = prob.sense == ObjSense.MAXIMISE
    h = host_lp(prob, -prob.c if flip else prob.c)
    x, clo, chi = np.zeros(n), np.zeros(n), np.zeros(n)
    rhs, rlo, rhi, act, dual = (np.zeros(m) for _ in range(5))
    binding = np.zeros(m, dtype=np.uint8)
    _lib.call("sovlp_simplex_ranging", h.ref(), st.ctypes.data, x.ctypes.data, clo.ctypes.data,
              chi.ctypes.data, rhs.ctypes.data, rlo.ctypes.data, rhi.ctypes.data,
              act.ctypes.data, dual.ctypes.data, binding.ctypes.data)
    c_min = -prob.c if flip else prob.c
    red = c_min - prob.A.rmatvec(dual)
    if flip:
        clo, chi = -chi, -clo            # negation reverses the interval
        dual, red = -dual, -red
    dcols = int(np.sum(np.isfinite(clo) & np.isfinite(chi) & (chi - clo <= 1e-9)))
    bind = binding.astype(bool)
    drows = int(np.sum(bind & np.isfinite(rlo) & np.isfinite(rhi) & (rhi - rlo <= 1e-9)))
    return Sensitivity(value=x, reduced_cost=red, cost=prob.c.copy(), cost_lo=clo, cost_hi=chi,
                       activity=act, dual=dual, rhs=rhs, rhs_lo=rlo, rhs_hi=rhi, binding=bind,
                       col_names=prob.col_names, row_names=prob.row_names,
                       degenerate_cols=dcols, degenerate_rows=drows)
