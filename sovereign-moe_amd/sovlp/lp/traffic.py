"""Algorithmic HBM bytes of the device PDLP phases.

Each phase of the device state machine (one launch's work for one LP) reads
and writes a known set of arrays.  ``phase_bytes`` counts every compulsory
array once: matrix values and indices once per pass, the gathered vector once
(its length, not once per nonzero), pointer and tile arrays once, and each
row/column vector the phase touches once per read and once per write.  Cache
reuse of a gathered vector is therefore not charged, and neither are re-reads
that a perfect cache would avoid.  This is the yardstick that a kernel's
achieved bandwidth (bytes / time) is measured against; hardware counters give
the actual traffic.

An SpMV ``y = M v`` over R rows, C columns and Z nonzeros costs
``Z (s_val + 4) + (R + 1) 4 + C s_v + R s_y`` (4-byte indices and pointers).

The engines count how often each phase ran (``sovlp_*_phase_counts``); the
B phases that also formed the next primal step are counted a second time in
the ``*_NEXT`` slots, which carry only the extra bytes.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib

__all__ = ["phase_names", "phase_bytes", "bytes_of_counts"]

_F64 = 8
_IDX = 4


def phase_names() -> list[str]:
    """Slot names in the library's order."""
    n = C.c_int()
    _lib.call("sovlp_phase_slots", C.byref(n))
    return [_lib.lib().sovlp_phase_name(i).decode() for i in range(n.value)]


def phase_bytes(name: str, m: int, n: int, nnz: int, ntiles_a: int, ntiles_t: int,
                precision: int) -> int:
    """Algorithmic bytes one execution of phase ``name`` moves for one LP."""
    s = 8 if precision == 64 else 4
    # structure of one pass over A (rows m) or A^T (rows n): indices, pointers, tiles
    pa = nnz * _IDX + (m + 1) * _IDX + (ntiles_a + 1) * _IDX
    pt = nnz * _IDX + (n + 1) * _IDX + (ntiles_t + 1) * _IDX
    b = {
        "INIT": _F64 * (m + n),
        # |a| r c products over A and A^T (FP64 original values); factors gathered
        "RUIZ_MAX": pa + pt + 2 * nnz * _F64 + 3 * _F64 * (m + n),
        "RUIZ_APPLY": 3 * _F64 * (m + n),
        "PC_SUM": pa + pt + 2 * nnz * _F64 + 3 * _F64 * (m + n),
        "PC_APPLY": 3 * _F64 * m + 4 * _F64 * n,
        # scaled copy of A and A^T, bounds, c, x0, y0
        "SCALE": pa + pt + 2 * nnz * (_F64 + s) + 2 * _F64 * (m + n)
                 + n * (5 * _F64 + 7 * s) + m * (3 * _F64 + 5 * s),
        "AX0": pa + nnz * s + n * s + m * s + n * (_F64 + s),
        "POW_A": pa + nnz * s + n * s + m * s,
        "POW_T": pt + nnz * s + m * s + n * s,
        "KKT0": 4 * s * m + 5 * s * n,
        "P_TRIAL": 7 * s * n,
        "A_TRIAL": pa + nnz * s + n * s + 7 * s * m,
        "B_ADAPT": pt + nnz * s + m * s + 6 * s * n + 7 * s * m,
        "A_FUSED": pa + nnz * s + n * s + 9 * s * m,
        "B_FUSED": pt + nnz * s + m * s + 5 * s * n,
        "RESTART": 8 * s * (m + n),
        "COPY_AVG": 2 * s * (m + n),
        "FINAL": n * (s + 5 * _F64) + m * (s + 2 * _F64),
        "FINAL_D": pt + nnz * _F64 + m * _F64 + 2 * _F64 * n,
        # refinement rounds (FP32 engines): X += x col / Delta
        "REF_X": n * (s + 5 * _F64),
        # FP64 KKT of X on the original: A X (kept) and A^T y
        "REF_KKT": (pa + nnz * _F64 + n * _F64 + m * (s + 5 * _F64))
                   + (pt + nnz * _F64 + m * (s + _F64) + n * 4 * _F64),
        # shifted, scaled bounds of the next round and its starting point
        "REF_SHIFT": m * (6 * _F64 + 5 * s) + n * (6 * _F64 + 5 * s),
        "DONE": 0,
        # extra bytes of a B phase that also formed the next primal step:
        # read c, lo, hi; write dx and x (or x')
        "B_ADAPT_NEXT": 5 * s * n,
        "B_FUSED_NEXT": 5 * s * n,
    }
    if name == "CHECK":
        # two-vector passes over A and A^T (current and average points)
        v = (pa + nnz * s + 2 * s * n + 7 * s * m) + (pt + nnz * s + 2 * s * m + 8 * s * n)
        if precision == 64:
            # unscaled KKT from the same products: row/col factors and original bounds
            v += 3 * _F64 * m + 4 * _F64 * n
        else:
            # unscaled KKT recomputed in FP64 from the original matrix, one
            # vector per pass: current point (sum kept in tmpm / tmpn), then average
            v += pa + nnz * _F64 + n * (s + _F64) + m * _F64
            v += pa + nnz * _F64 + n * (s + _F64) + m * (2 * s + 4 * _F64)
            v += pt + nnz * _F64 + m * (s + _F64) + n * _F64
            v += pt + nnz * _F64 + m * (s + _F64) + n * (2 * s + 5 * _F64)
        return int(v)
    return int(b.get(name, 0))


def bytes_of_counts(counts: np.ndarray, sizes, precision: int) -> float:
    """Total algorithmic bytes of a batch: ``counts`` is [LPs, slots] as the
    engine reports it, ``sizes`` a list of (m, n, nnz, ntiles_a, ntiles_t)."""
    names = phase_names()
    counts = np.asarray(counts, dtype=np.float64).reshape(len(sizes), len(names))
    used = [k for k in range(len(names)) if counts[:, k].any()]
    total = 0.0
    for k in used:
        per_lp = np.array([phase_bytes(names[k], *sz, precision) for sz in sizes], np.float64)
        total += float(counts[:, k] @ per_lp)
    return total
