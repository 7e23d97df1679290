"""Seeded synthetic LPs with a known optimum, and the blending request stream.

Construction of the "known optimum" LP (used for the single CPU case, the large
skewed LP and the uneven batch):

* ``m`` rows, the first ``m_eq`` equalities and the rest ``<=`` rows; ``n``
  columns, each with ``deg[j]`` entries at uniformly random rows (drawn with
  replacement, duplicates summed as the matrix builder does), values ``N(0,1)``;
  bounds ``0 <= x <= ub``.
* ``x*``: 30 % of the columns at 0, 10 % at ``ub``, the rest ``U(0, ub)``.
* ``y*`` (textbook sign): ``N(0,1)`` on equality rows; half of the ``<=`` rows
  active with ``y* = -|N(0,1)|``, the other half ``y* = 0``.
* ``b = A x*`` on equality and active rows, ``A x* + U(0,1)`` on inactive ones.
* ``c = Aᵀ y* + z`` with ``z = U(0.1,1)`` at the lower bound, ``-U(0.1,1)`` at the
  upper bound and 0 in the interior.

``(x*, y*)`` then satisfies the KKT conditions exactly, so ``f* = cᵀx*`` equals
the dual objective ``bᵀy* + Σ_{x*_j = ub_j} ub_j z_j``; both are returned,
computed with compensated (``math.fsum``) sums.

The stored ``A`` is exactly the one ``b`` and ``c`` are computed from, so the
optimum is that of the problem handed to the solver.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .core.problem import Problem
from .core.sparse import SparseMatrix
from .core.tolerances import INF

__all__ = ["KnownOptimum", "known_optimum_lp", "small_lp", "skewed_lp_degrees",
           "skewed_lp", "uneven_batch", "uneven_batch_member", "blending_lp",
           "blending_stream"]


@dataclass
class KnownOptimum:
    problem: Problem
    x_star: np.ndarray
    y_star: np.ndarray
    f_star: float
    dual_star: float


def known_optimum_lp(m: int, n: int, deg, seed: int, *, m_eq: int | None = None,
                     ub=10.0, name: str | None = None, rng=None) -> KnownOptimum:
    """Build the known-optimum LP. ``deg`` is an int (same for every column) or an
    int array of length ``n``; ``ub`` is a scalar or length-``n`` array."""
    rng = np.random.default_rng(seed) if rng is None else rng
    if m_eq is None:
        m_eq = m // 2
    deg = np.full(n, int(deg), dtype=np.int64) if np.isscalar(deg) else np.asarray(deg, np.int64)
    cols = np.repeat(np.arange(n, dtype=np.int64), deg)
    rows = rng.integers(0, m, cols.size)
    vals = rng.standard_normal(cols.size)
    A = SparseMatrix.from_triplets(rows, cols, vals, m, n)
    ubv = np.full(n, float(ub)) if np.isscalar(ub) else np.asarray(ub, dtype=np.float64)

    perm = rng.permutation(n)
    n_lo, n_hi = int(round(0.3 * n)), int(round(0.1 * n))
    at_lo, at_hi, inner = perm[:n_lo], perm[n_lo:n_lo + n_hi], perm[n_lo + n_hi:]
    x = np.empty(n)
    x[at_lo] = 0.0
    x[at_hi] = ubv[at_hi]
    x[inner] = rng.uniform(0.0, 1.0, inner.size) * ubv[inner]

    n_le = m - m_eq
    y = np.zeros(m)
    y[:m_eq] = rng.standard_normal(m_eq)
    le_perm = m_eq + rng.permutation(n_le)
    active = le_perm[: n_le // 2]
    inactive = le_perm[n_le // 2:]
    y[active] = -np.abs(rng.standard_normal(active.size))

    ax = A.matvec(x)
    b = ax.copy()
    b[inactive] += rng.uniform(0.0, 1.0, inactive.size)

    z = np.zeros(n)
    z[at_lo] = rng.uniform(0.1, 1.0, at_lo.size)
    z[at_hi] = -rng.uniform(0.1, 1.0, at_hi.size)
    c = A.rmatvec(y) + z

    row_lb = np.full(m, -INF)
    row_lb[:m_eq] = b[:m_eq]
    row_ub = b.copy()
    prob = Problem(A=A, c=c, row_lb=row_lb, row_ub=row_ub, col_lb=np.zeros(n),
                   col_ub=ubv.copy(), name=name or f"known_{m}x{n}_s{seed}")
    f_star = math.fsum((c * x).tolist())
    dual = math.fsum((y * b).tolist()) + math.fsum((ubv[at_hi] * z[at_hi]).tolist())
    return KnownOptimum(prob, x, y, f_star, dual)


def small_lp(seed: int = 0) -> KnownOptimum:
    """The single CPU-runnable case: m=500 (250 eq, 250 <=), n=2000, 5 entries per
    column, bounds [0, 10]."""
    return known_optimum_lp(500, 2000, 5, seed, m_eq=250, ub=10.0, name=f"small_500x2000_s{seed}")


def skewed_lp_degrees(n: int = 1_000_000, total: int = 8_000_000, s: float = 1.3,
                      dmin: int = 2, dmax: int = 5000) -> np.ndarray:
    """Column degrees by the Zipf rank-frequency law ``d_r = clip(rint(C r^-s))``.

    ``C`` is found by bisection so that the degrees sum to ``total`` (to within
    one column's worth). Sampling degrees i.i.d. from a clipped Zipf pmf would
    give a mean degree of 200+ and hundreds of millions of nonzeros, contrary
    to "about 8M nonzeros", so the rank-frequency reading is used.
    """
    r = np.arange(1, n + 1, dtype=np.float64) ** (-s)

    def tot(C):
        return int(np.clip(np.rint(C * r), dmin, dmax).sum())

    lo, hi = 1.0, 1e12
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if tot(mid) < total:
            lo = mid
        else:
            hi = mid
    return np.clip(np.rint(hi * r), dmin, dmax).astype(np.int64)


def skewed_lp(seed: int = 1, m: int = 200_000, n: int = 1_000_000,
              total_nnz: int = 8_000_000) -> KnownOptimum:
    """The large skewed LP: ``m`` rows (half equality), ``n`` columns whose degrees
    follow the Zipf rank law (s=1.3, 2..5000), shuffled by the seed."""
    rng = np.random.default_rng(seed)
    deg = skewed_lp_degrees(n, total_nnz)
    deg = deg[rng.permutation(n)]
    return known_optimum_lp(m, n, deg, seed, m_eq=m // 2, ub=10.0,
                            name=f"skewed_{m}x{n}_s{seed}", rng=rng)


def uneven_batch_member(seed: int) -> KnownOptimum:
    """One LP of the uneven batch: ``m ~ exp(U(ln 20, ln 800))`` rows, 60 %
    equalities, ``n = floor(U(2,4)·m)`` columns with ``U{3..8}`` entries each,
    one upper bound ``U(5, 50)`` shared by all columns of the LP."""
    rng = np.random.default_rng(seed)
    m = int(round(math.exp(rng.uniform(math.log(20.0), math.log(800.0)))))
    n = int(math.floor(rng.uniform(2.0, 4.0) * m))
    deg = rng.integers(3, 9, n)
    ub = float(rng.uniform(5.0, 50.0))
    return known_optimum_lp(m, n, deg, seed, m_eq=int(round(0.6 * m)), ub=ub,
                            name=f"batch_{seed}", rng=rng)


def uneven_batch(count: int = 4096, base_seed: int = 1000) -> list[KnownOptimum]:
    return [uneven_batch_member(base_seed + i) for i in range(count)]


# --------------------------------------------------------------------------- #
# multi-period blending (the served workload)                                  #
# --------------------------------------------------------------------------- #


def blending_lp(seed: int, K: int | None = None, P: int | None = None,
                T: int | None = None, *, max_T: int = 720) -> Problem:
    """A multi-period blending LP, feasible by construction.

    Sizes: ``K ~ U{20..60}`` components, ``P ~ U{5..20}`` products, ``T ~
    U{24..max_T}`` hours (each may be fixed by the caller). Variables: flows
    ``f[k,p,t] in [0, cap_k]`` (component k into product p in hour t) and end-of-hour
    inventories ``s[k,t] in [0, INF)``. Rows of hour t:

    * inventory balance (equality) per component:
      ``s[k,t] - s[k,t-1] + Σ_p f[k,p,t] = a[k,t]`` (``a`` = the hour's arrivals;
      no ``s[k,t-1]`` term at t = 0);
    * supply (``<=``) per component: ``Σ_p f[k,p,t] <= avail_k``;
    * demand (equality) per product: ``Σ_k f[k,p,t] = d_p``;
    * three quality rows per product: ``Σ_k (q[k,j] - qmax[p,j]) f[k,p,t] <= 0``.

    Draws: ``cap ~ U(50,500)``, ``avail ~ U(100,1000)``, ``d ~ U(50,300)``,
    ``q ~ U(0,1)``, component price ``U(40,80)``, product price ``U(70,110)``.
    Feasibility: a random blend ``w[p,:]`` (Dirichlet, weighted by the caps) is
    fixed first; its flows ``d_p w[p,k]`` and per-component use ``u_k`` define
    ``qmax[p,j] = 1.05 (w[p,:] · q[:,j])``, raise ``cap_k`` and ``avail_k`` to
    at least 1.05 times what the blend needs (a no-op for most components), and
    set arrivals ``a[k,t] = u_k U(1.0, 1.3)``, so the blend with its inventories
    satisfies every row. The objective is component cost minus product revenue.
    """
    rng = np.random.default_rng(seed)
    K = int(rng.integers(20, 61)) if K is None else int(K)
    P = int(rng.integers(5, 21)) if P is None else int(P)
    T = int(rng.integers(24, max_T + 1)) if T is None else int(T)
    cap = rng.uniform(50.0, 500.0, K)
    avail = rng.uniform(100.0, 1000.0, K)
    demand = rng.uniform(50.0, 300.0, P)
    q = rng.uniform(0.0, 1.0, (K, 3))
    price_k = rng.uniform(40.0, 80.0, K)
    price_p = rng.uniform(70.0, 110.0, P)
    w = rng.dirichlet(np.ones(K), P) * cap[None, :]
    w /= w.sum(axis=1, keepdims=True)                   # (P, K) blend fractions
    flow = demand[:, None] * w                          # the blend's f[k,p,t], any t
    use = flow.sum(axis=0)                              # (K,) per-hour use of each component
    qmax = 1.05 * (w @ q)                               # (P, 3)
    cap = np.maximum(cap, 1.05 * flow.max(axis=0))
    avail = np.maximum(avail, 1.05 * use)
    arrive = use[:, None] * rng.uniform(1.0, 1.3, (K, T))

    nf, ns = K * P * T, K * T
    n = nf + ns
    R = 2 * K + 4 * P                                   # rows per hour
    m = R * T
    t3, k3, p3 = np.meshgrid(np.arange(T), np.arange(K), np.arange(P), indexing="ij")
    t3, k3, p3 = t3.ravel(), k3.ravel(), p3.ravel()
    fcol = (t3 * K + k3) * P + p3
    base = t3 * R
    tk_t, tk_k = np.meshgrid(np.arange(T), np.arange(K), indexing="ij")
    tk_t, tk_k = tk_t.ravel(), tk_k.ravel()
    scol = nf + tk_t * K + tk_k
    prev = tk_t > 0
    qual_r = (base[:, None] + 2 * K + P + 3 * p3[:, None] + np.arange(3)[None, :]).ravel()
    qual_v = (q[k3] - qmax[p3]).ravel()
    rows = np.concatenate([base + k3, tk_t * R + tk_k, (tk_t * R + tk_k)[prev],
                           base + K + k3, base + 2 * K + p3, qual_r])
    cols = np.concatenate([fcol, scol, (scol - K)[prev], fcol, fcol, np.repeat(fcol, 3)])
    vals = np.concatenate([np.ones(nf), np.ones(ns), -np.ones(int(prev.sum())), np.ones(nf),
                           np.ones(nf), qual_v])
    A = SparseMatrix.from_triplets(rows, cols, vals, m, n)
    hour_lb = np.concatenate([np.zeros(K), np.full(K, -INF), demand, np.full(3 * P, -INF)])
    hour_ub = np.concatenate([np.zeros(K), avail, demand, np.zeros(3 * P)])
    rl = np.tile(hour_lb, T)
    ru = np.tile(hour_ub, T)
    bal = (np.arange(T)[:, None] * R + np.arange(K)[None, :]).ravel()
    rl[bal] = arrive.T.ravel()
    ru[bal] = arrive.T.ravel()
    c = np.zeros(n)
    c[:nf] = price_k[k3] - price_p[p3]
    col_lb = np.zeros(n)
    col_ub = np.full(n, INF)
    col_ub[:nf] = cap[k3]
    return Problem(A=A, c=c, row_lb=rl, row_ub=ru, col_lb=col_lb, col_ub=col_ub,
                   name=f"blend_K{K}_P{P}_T{T}_s{seed}")


def blending_stream(rate: float = 200.0, duration: float = 60.0, seed: int = 0,
                    resolve_frac: float = 0.3, perturb: float = 0.01, max_T: int = 720):
    """Poisson arrivals of blending requests. Yields ``(t_arrival, key, kind,
    seed, perturb_seed)`` where ``kind`` is ``"new"`` or ``"resolve"``; a resolve
    refers to an earlier request's model with ``b`` and ``c`` perturbed by
    ``perturb`` (relative, uniform in ±perturb)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    i = 0
    issued = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        if issued and rng.random() < resolve_frac:
            base = issued[int(rng.integers(0, len(issued)))]
            yield (t, i, "resolve", base, int(rng.integers(0, 2**31 - 1)))
        else:
            s = int(rng.integers(0, 2**31 - 1))
            issued.append(s)
            yield (t, i, "new", s, 0)
        i += 1


def perturbed(prob: Problem, pseed: int, rel: float = 0.01) -> Problem:
    """``prob`` with finite row bounds and costs scaled by ``1 + U(-rel, rel)``."""
    rng = np.random.default_rng(pseed)
    q = prob.copy()
    fb = q.row_lb > -INF
    fu = q.row_ub < INF
    eq = fb & fu & (q.row_lb == q.row_ub)
    f = 1.0 + rng.uniform(-rel, rel, q.m)
    q.row_ub = np.where(fu, q.row_ub * f, q.row_ub)
    q.row_lb = np.where(eq, q.row_ub, np.where(fb, q.row_lb * f, q.row_lb))
    q.c = q.c * (1.0 + rng.uniform(-rel, rel, q.n))
    return q
