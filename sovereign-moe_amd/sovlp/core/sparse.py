"""Host-side sparse matrix held in both CSC and CSR.

Interface follows ``SparseMatrix`` in the reference (``src/sovopt/core/sparse.py``):

* ``from_triplets`` sorts entries by (column, row), sums exact duplicates and
  drops structural zeros (``sparse.py:198-236``);
* CSR is derived from CSC by a stable counting transpose so both orders hold
  the same values (``sparse.py:239-255``);
* pointers are ``int64``, indices ``int32``, values ``float64``
  (``sparse.py:41-42, 280-287``).

This module only builds and inspects matrices on the host. Every product the
solver takes runs on the GPU through ``libsovlp``; the NumPy ``matvec`` here is
for model inspection (``Problem.violation``) and result reporting, as in the
reference CLI.
"""

from __future__ import annotations

import numpy as np

__all__ = ["SparseMatrix", "coo_to_csc", "csc_to_csr", "locality_order", "relabel_columns",
           "IDX", "VAL"]

IDX = np.int32
VAL = np.float64


def coo_to_csc(rows, cols, vals, m: int, n: int, sum_duplicates: bool = True):
    """Triplets -> CSC ``(indptr int64[n+1], rows int32[nnz], vals f64[nnz])``.

    Entries are ordered by column then row. Repeated ``(row, col)`` pairs are
    added together (an MPS file that lists a coefficient twice means their
    sum) and exact zeros are removed afterwards.
    """
    rows = np.asarray(rows, dtype=np.int64).ravel()
    cols = np.asarray(cols, dtype=np.int64).ravel()
    vals = np.asarray(vals, dtype=VAL).ravel()
    if not (rows.size == cols.size == vals.size):
        raise ValueError("triplet arrays must have equal length")
    if rows.size and (rows.min() < 0 or rows.max() >= m or cols.min() < 0 or cols.max() >= n):
        raise ValueError("triplet index out of range")

    key = cols * np.int64(max(m, 1)) + rows
    order = np.argsort(key, kind="stable")
    key = key[order]
    vals = vals[order]
    if sum_duplicates and key.size:
        head = np.empty(key.size, dtype=bool)
        head[0] = True
        np.not_equal(key[1:], key[:-1], out=head[1:])
        if not head.all():
            grp = np.cumsum(head) - 1
            vals = np.bincount(grp, weights=vals)
            key = key[head]
    keep = vals != 0.0
    if not keep.all():
        key, vals = key[keep], vals[keep]
    out_cols = key // max(m, 1)
    out_rows = (key - out_cols * max(m, 1)).astype(IDX)
    indptr = np.zeros(n + 1, dtype=np.int64)
    if out_cols.size:
        np.cumsum(np.bincount(out_cols, minlength=n), out=indptr[1:])
    return indptr, np.ascontiguousarray(out_rows), np.ascontiguousarray(vals, dtype=VAL)


def csc_to_csr(indptr, indices, values, m: int):
    """Transpose storage order. Within each row, columns come out ascending."""
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=IDX)
    nnz = indices.size
    n = indptr.size - 1
    col_of = np.repeat(np.arange(n, dtype=IDX), np.diff(indptr))
    order = np.argsort(indices, kind="stable")   # stable keeps column order per row
    t_ptr = np.zeros(m + 1, dtype=np.int64)
    if nnz:
        np.cumsum(np.bincount(indices, minlength=m), out=t_ptr[1:])
    return (t_ptr, np.ascontiguousarray(col_of[order], dtype=IDX),
            np.ascontiguousarray(np.asarray(values, dtype=VAL)[order]))


class SparseMatrix:
    """``m x n`` matrix stored as CSC (``cp, ci, cx``) and CSR (``rp, ri, rx``)."""

    __slots__ = ("m", "n", "cp", "ci", "cx", "rp", "ri", "rx")

    def __init__(self, m: int, n: int, cp, ci, cx, rp=None, ri=None, rx=None):
        self.m = int(m)
        self.n = int(n)
        self.cp = np.ascontiguousarray(cp, dtype=np.int64)
        self.ci = np.ascontiguousarray(ci, dtype=IDX)
        self.cx = np.ascontiguousarray(cx, dtype=VAL)
        if rp is None:
            rp, ri, rx = csc_to_csr(self.cp, self.ci, self.cx, self.m)
        self.rp = np.ascontiguousarray(rp, dtype=np.int64)
        self.ri = np.ascontiguousarray(ri, dtype=IDX)
        self.rx = np.ascontiguousarray(rx, dtype=VAL)

    @classmethod
    def from_triplets(cls, rows, cols, vals, m: int, n: int) -> "SparseMatrix":
        cp, ci, cx = coo_to_csc(rows, cols, vals, m, n)
        return cls(m, n, cp, ci, cx)

    @classmethod
    def from_dense(cls, A) -> "SparseMatrix":
        A = np.asarray(A, dtype=VAL)
        r, c = np.nonzero(A)
        return cls.from_triplets(r, c, A[r, c], A.shape[0], A.shape[1])

    @property
    def shape(self):
        return (self.m, self.n)

    @property
    def nnz(self) -> int:
        return int(self.ci.size)

    @property
    def density(self) -> float:
        d = self.m * self.n
        return self.nnz / d if d else 0.0

    def __repr__(self):
        return f"SparseMatrix({self.m}x{self.n}, nnz={self.nnz})"

    # Host products for inspection and reporting only (see module docstring).
    def matvec(self, x):
        x = np.asarray(x, dtype=VAL)
        rows = np.repeat(np.arange(self.m), np.diff(self.rp))
        return np.bincount(rows, weights=self.rx * x[self.ri], minlength=self.m)

    def rmatvec(self, y):
        y = np.asarray(y, dtype=VAL)
        cols = np.repeat(np.arange(self.n), np.diff(self.cp))
        return np.bincount(cols, weights=self.cx * y[self.ci], minlength=self.n)

    def row_counts(self):
        return np.diff(self.rp).astype(IDX)

    def col_counts(self):
        return np.diff(self.cp).astype(IDX)

    def coeff_range(self):
        if self.nnz == 0:
            return (0.0, 0.0)
        a = np.abs(self.cx)
        return (float(a.min()), float(a.max()))

    def copy(self) -> "SparseMatrix":
        return SparseMatrix(self.m, self.n, self.cp.copy(), self.ci.copy(), self.cx.copy(),
                            self.rp.copy(), self.ri.copy(), self.rx.copy())

    def to_dense(self):
        A = np.zeros((self.m, self.n), dtype=VAL)
        cols = np.repeat(np.arange(self.n), np.diff(self.cp))
        A[self.ci, cols] = self.cx
        return A


def locality_order(A: SparseMatrix, hot_degree: int = 8) -> np.ndarray:
    """A new label for each column, so that the gathers of ``A x`` and the rows of
    ``A^T y`` hit cache: columns with at least ``hot_degree`` nonzeros first, by
    degree descending, then the rest by the first row that holds them (ties by
    old label). On a matrix whose few long columns are scattered over a long x,
    the gathers of the long columns then come from a short prefix of x."""
    deg = np.diff(A.cp)
    first = np.full(A.n, np.iinfo(np.int64).max, dtype=np.int64)
    has = deg > 0
    first[has] = A.ci[A.cp[:-1][has]]
    hot = deg >= hot_degree
    order = np.lexsort((np.arange(A.n), first, np.where(hot, -deg, 0), ~hot))
    new = np.empty(A.n, dtype=np.int64)
    new[order] = np.arange(A.n)
    return new


def relabel_columns(A: SparseMatrix, new) -> SparseMatrix:
    """A with column ``j`` renamed ``new[j]`` (a permutation of ``range(n)``).

    Each row keeps its entries in their order and each column its entries, so
    with ``x_new[new[j]] = x[j]`` every row sum of ``A x`` and every column sum of
    ``A^T y``, taken in storage order, is bit for bit the same as before. (The
    device's tile pass over ``A^T`` plans its tiles over the moved rows and may
    reduce a row over a different number of lanes: those sums agree to
    rounding.)"""
    new = np.asarray(new, dtype=np.int64)
    if new.shape != (A.n,) or not np.array_equal(np.sort(new), np.arange(A.n)):
        raise ValueError("new must be a permutation of the column labels")
    order = np.argsort(new)                     # old column at each new position
    lens = np.diff(A.cp)[order]
    cp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    take = np.arange(A.nnz, dtype=np.int64) + np.repeat(A.cp[:-1][order] - cp[:-1], lens)
    return SparseMatrix(A.m, A.n, cp, A.ci[take], A.cx[take], A.rp.copy(),
                        new[A.ri].astype(IDX), A.rx.copy())
