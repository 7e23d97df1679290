"""CPU restatement of the batched node relaxation (test infrastructure only).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of
``sovlp_bnr_solve_batch``, never by the ``sovlp`` package.

Restates ``BNREngine.solve_batch`` (reference ``src/sovopt/mip/bnr.py:176-301``)
and ``safe_dual_bound_batch`` (``src/sovopt/mip/safebound.py:194-235``) in
NumPy, with the same per-element operations.  The sparse products walk each
row's nonzeros in CSR order for all nodes at once, as the reference's
``_csr_spmm`` (``core/sparse.py:72-83``).  The per-node sums over rows (the
step-size test's movement and interaction, the bound's two sums) are taken in
the device's fixed order -- blocks of ``256 / KP`` rows (KP: K rounded up to a
power of two), then the blocks over 256 lanes in a fixed order -- where the reference's NumPy sums
run down the rows.  The interaction sum can cancel heavily, so any other
order moves the step sizes visibly; with the same order the whole batch is
the device's bit for bit.  ``order="rows"`` sums as the reference does; run
that way the oracle reproduces the reference's own batches
(``tests/golden/reference_bnr.json``, written by running ``BNREngine``).
"""

from __future__ import annotations

import numpy as np

INF = 1.0e30
QUOTIENT_MAX = 2.0 ** 1000


def spmm(rp, ri, rx, V, rows):
    """``out[i, :] = sum_p rx[p] V[ri[p], :]`` over row i's nonzeros, in order."""
    K = V.shape[1]
    out = np.zeros((rows, K))
    for i in range(rows):
        acc = np.zeros(K)
        for p in range(rp[i], rp[i + 1]):
            acc = acc + rx[p] * V[ri[p]]
        out[i] = acc
    return out


def block_rows(K: int) -> int:
    """Rows per device workgroup for a batch of K nodes."""
    kp = 1
    while kp < K:
        kp <<= 1
    return 256 // kp


def blocked_sum(V, rb: int):
    """Per-column sum of V (rows, K) in the device's order: each block of rb
    rows in order from 0; then lane t (of 256) adds blocks t, t + 256, ... in
    order from 0, and the 256 lane sums are added in order from 0."""
    K = V.shape[1]
    nblocks = max(1, -(-V.shape[0] // rb))
    parts = []
    for g in range(nblocks):
        part = np.zeros(K)
        for r in range(g * rb, min((g + 1) * rb, V.shape[0])):
            part = part + V[r]
        parts.append(part)
    total = np.zeros(K)
    for t in range(256):
        lane = np.zeros(K)
        for g in range(t, nblocks, 256):
            lane = lane + parts[g]
        total = total + lane
    return total


def row_sum(V):
    """Per-column sum of V (rows, K) as the reference takes it: NumPy's
    ``V.sum(axis=0)`` of a C-ordered array (down the rows for K > 1; NumPy's
    pairwise order when K == 1 makes the rows contiguous)."""
    return np.ascontiguousarray(V).sum(axis=0)


def _sum(V, rb: int, order: str):
    return blocked_sum(V, rb) if order == "device" else row_sum(V)


def safe_dual_bound_batch(prob, L, U, Y, order: str = "device"):
    """(K,) bounds from the (m, K) duals; -inf where a term is vacuous.
    ``order``: "device" sums in the device's order, "rows" in the reference's."""
    A = prob.A
    aty = spmm(A.cp, A.ci, A.cx, Y, A.n)
    d = prob.c[:, None] - aty
    rl, ru = prob.row_lb[:, None], prob.row_ub[:, None]
    ypos = Y > 0.0
    row_bad = np.where(ypos, rl <= -INF, ru >= INF) & (Y != 0.0)
    row_pick = np.where(ypos, rl, ru)
    row_c = np.where(row_bad, 0.0, Y * np.where(row_bad, 0.0, row_pick))
    dpos = d > 0.0
    col_bad = np.where(dpos, L <= -INF, U >= INF) & (d != 0.0)
    col_pick = np.where(dpos, L, U)
    col_c = np.where(col_bad, 0.0, d * np.where(col_bad, 0.0, col_pick))
    vac = row_bad.any(axis=0) | col_bad.any(axis=0)
    rb = block_rows(Y.shape[1])
    s = _sum(row_c, rb, order) if row_c.shape[0] else np.zeros(Y.shape[1])
    t = _sum(col_c, rb, order)
    return np.where(vac, -np.inf, s + t)


def solve_batch(prob, L, U, iters=60, warm=None, eta=None, omega=None, growth=0.3, damp=0.6,
                dual_sign=1.0, order: str = "device"):
    """One batch visit: returns (bounds scaled, X, Y, eta, empty).  The bound
    reads ``dual_sign * Y`` (1: as the reference; -1: textbook signs).
    ``order`` of the per-node sums: "device" (the GPU's fixed order) or "rows"
    (down the rows, the reference's NumPy order)."""
    A = prob.A
    m, n, K = A.m, A.n, L.shape[1]
    L = np.asarray(L, dtype=np.float64)
    U = np.asarray(U, dtype=np.float64)
    empty = (L > U + 1e-12).any(axis=0)
    X = np.zeros((n, K)) if warm is None else np.array(warm[0], dtype=np.float64)
    Y = np.zeros((m, K)) if warm is None else np.array(warm[1], dtype=np.float64)
    X = np.minimum(np.maximum(X, L), U)
    eta = np.ones(K) if eta is None else np.array(eta, dtype=np.float64)
    omega = np.ones(K) if omega is None else np.array(omega, dtype=np.float64)
    AX = spmm(A.rp, A.ri, A.rx, X, m)
    AtY = spmm(A.cp, A.ci, A.cx, Y, n)
    c, rl, ru = prob.c[:, None], prob.row_lb[:, None], prob.row_ub[:, None]
    attempts = 0
    for _ in range(iters):
        attempts += 1
        tau = (eta / omega)[None, :]
        sigma = (eta * omega)[None, :]
        Xn = np.minimum(np.maximum(X - tau * (c + AtY), L), U)
        dX = Xn - X
        AdX = spmm(A.rp, A.ri, A.rx, dX, m)
        W = Y + sigma * (AX + 2.0 * AdX)
        Yn = W - sigma * np.minimum(np.maximum(W / sigma, rl), ru)
        dY = Yn - Y
        AtdY = spmm(A.cp, A.ci, A.cx, dY, n)
        rb = block_rows(K)
        inter = _sum(dY * AdX, rb, order) if m else np.zeros(K)
        dyy = _sum(dY * dY, rb, order) if m else np.zeros(K)
        dxx = _sum(dX * dX, rb, order)
        interaction = np.abs(inter)
        movement = 0.5 * (omega * dxx + dyy / omega)
        floor = np.maximum(movement, 1.0) / QUOTIENT_MAX
        with np.errstate(divide="ignore", invalid="ignore"):
            limit = np.where(interaction > 0.0, movement / np.maximum(interaction, floor), np.inf)
        accept = (eta <= limit) | (movement == 0.0)
        a = accept[None, :]
        X = np.where(a, Xn, X)
        Y = np.where(a, Yn, Y)
        AX = np.where(a, AX + AdX, AX)
        AtY = np.where(a, AtY + AtdY, AtY)
        t = attempts + 1
        eta = np.minimum((1.0 - t ** -growth) * limit, (1.0 + t ** -damp) * eta)
        eta = np.maximum(np.where(np.isfinite(eta), eta, 1.0), 1e-14)
    bounds = safe_dual_bound_batch(prob, L, U, dual_sign * Y, order)
    return bounds, X, Y, eta, empty
