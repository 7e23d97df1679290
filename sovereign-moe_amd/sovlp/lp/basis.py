"""Basis status codes and the nonsingularity check of a basis.

The reference's conventions (``src/sovopt/lp/basis.py:75-79, 203-215``): the
variables are ``[x ; s]`` with ``A x - s = 0``, ``n`` structurals then one
logical per row bounded by that row's bounds, and a status per variable:
BASIC, AT_LOWER, AT_UPPER, FREE (nonbasic at zero), FIXED.  The factorisation
itself is native (``crossover.cpp`` in ``libsovlp.so``); ``Basis`` keeps only
what a caller checks a basis with: its statuses, its basic variables, and how
many columns ``factorize`` had to replace by logicals (``n_repairs``; 0 for a
nonsingular basis).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from ..core.problem import Problem

__all__ = ["BASIC", "AT_LOWER", "AT_UPPER", "FREE", "FIXED", "Basis", "host_lp"]

BASIC = np.uint8(0)
AT_LOWER = np.uint8(1)
AT_UPPER = np.uint8(2)
FREE = np.uint8(3)
FIXED = np.uint8(4)


class _HostLP:
    """The ``sovlp_host_lp`` view of a Problem (holds the arrays it points to)."""

    def __init__(self, prob: Problem, c: np.ndarray | None = None):
        A = prob.A
        self.arrays = [np.ascontiguousarray(A.cp, dtype=np.int64),
                       np.ascontiguousarray(A.ci, dtype=np.int32),
                       np.ascontiguousarray(A.cx, dtype=np.float64),
                       np.ascontiguousarray(prob.c if c is None else c, dtype=np.float64),
                       np.ascontiguousarray(prob.col_lb, dtype=np.float64),
                       np.ascontiguousarray(prob.col_ub, dtype=np.float64),
                       np.ascontiguousarray(prob.row_lb, dtype=np.float64),
                       np.ascontiguousarray(prob.row_ub, dtype=np.float64)]
        ptr = [a.ctypes.data for a in self.arrays]
        self.desc = _lib.sovlp_host_lp(prob.m, prob.n, int(A.nnz), *ptr)

    def ref(self):
        return C.byref(self.desc)


def host_lp(prob: Problem, c: np.ndarray | None = None) -> _HostLP:
    """``prob`` (objective ``c`` if given) as the native code's host LP."""
    return _HostLP(prob, c)


class Basis:
    """A basis of ``prob`` by status vector (``n + m`` entries)."""

    def __init__(self, prob: Problem):
        self.prob = prob
        self.m, self.n = prob.m, prob.n
        self.N = self.m + self.n
        self.status = np.full(self.N, AT_LOWER, dtype=np.uint8)
        self.status[self.n:] = BASIC
        self.basic = np.arange(self.n, self.N, dtype=np.int64)
        self.n_repairs = 0

    def factorize(self) -> None:
        """Factorises the basis the statuses name; ``n_repairs`` counts the
        columns that had no pivot and would be replaced by logicals."""
        if int((self.status == BASIC).sum()) != self.m:
            raise ValueError("a basis has exactly m BASIC entries")
        h = host_lp(self.prob)
        st = np.ascontiguousarray(self.status, dtype=np.uint8)
        rep = C.c_int32()
        _lib.call("sovlp_basis_repairs", h.ref(), st.ctypes.data, C.byref(rep))
        self.n_repairs = int(rep.value)
