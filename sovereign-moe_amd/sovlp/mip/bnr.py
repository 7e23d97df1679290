"""Batched node relaxation: bounds for a whole branch-and-bound frontier at once.

The nodes of a tree share ``A``, ``c`` and the row bounds; branching changes
only column bounds.  ``BNREngine.solve_batch(col_lb, col_ub)`` advances ``K``
nodes together by a fixed number of PDHG iterations with a per-node adaptive
step and returns a valid lower bound per node from the batched
Neumaier-Shcherbina bound, which holds for any dual iterate, converged or not
(reference ``src/sovopt/mip/bnr.py:95-301``, ``mip/safebound.py:194-235``).
Everything runs in one library call (``sovlp_bnr_solve_batch``): the shared
matrix stays on the device, the ``(n, K)`` blocks are node-fastest, so one
row's K values are one coalesced run, and each product reads ``A`` once for
all K nodes.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..core.problem import Problem

__all__ = ["BNREngine", "BNRConfig", "BatchResult"]


@dataclass
class BNRConfig:
    batch: int = 128          # nodes per call (at most 256)
    iters: int = 60           # PDHG iterations per node visit
    warm_start: bool = True
    eps: float = 1e-9
    device: str = "auto"
    step_growth: float = 0.3
    step_damp: float = 0.6
    # The bound reads the dual iterate as the reference does (False): that
    # iterate is the negative of the textbook dual, so the bound is valid but
    # weak, and vacuous when a one-sided row's multiplier is nonzero.  True
    # reads it with textbook signs: the bound then tightens to each node's
    # optimum as the iterates converge.
    textbook_duals: bool = False


@dataclass
class BatchResult:
    bounds: np.ndarray                    # (K,) valid lower bounds, original objective scale
    x: np.ndarray                         # (n, K) primal iterates
    iterations: int = 0
    infeasible: np.ndarray | None = None  # (K,) nodes whose bounds are empty


class BNREngine:
    """The shared matrix on the device; ``solve_batch`` advances node batches.

    ``prob`` is the working problem the nodes branch on (scaled or not);
    ``scaling.obj`` (default 1) is its objective scale, which the bounds are
    divided by, as the reference does.
    """

    def __init__(self, prob: Problem, scaling=None, config: BNRConfig | None = None):
        import torch
        from ..core.backend import get_backend
        self.cfg = config or BNRConfig()
        if not 1 <= self.cfg.batch <= 256:
            raise ValueError("batch must be in [1, 256]")
        get_backend("gpu" if self.cfg.device in ("auto", "gpu") else self.cfg.device)
        self.prob = prob
        self.sc = scaling
        self.m, self.n = prob.m, prob.n
        A = prob.A

        def up(a, dt):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to("cuda")

        self.t = dict(rp=up(A.rp, np.int64), ri=up(A.ri, np.int32), rx=up(A.rx, np.float64),
                      cp=up(A.cp, np.int64), ci=up(A.ci, np.int32), cx=up(A.cx, np.float64),
                      c=up(prob.c, np.float64), rl=up(prob.row_lb, np.float64),
                      ru=up(prob.row_ub, np.float64))
        self.obj_scale = getattr(scaling, "obj", 1.0) or 1.0
        self.obj_offset = prob.obj_offset
        self._alloc(self.cfg.batch)
        self.total_iterations = 0
        self.total_batches = 0
        self.total_nodes = 0

    def _alloc(self, K: int):
        import torch
        self.K = K
        z = lambda *s: torch.zeros(*s, dtype=torch.float64, device="cuda")   # noqa: E731
        self.X = z(self.n, K)
        self.Y = z(max(self.m, 1), K)
        self.eta = np.ones(K)
        self.omega = np.ones(K)

    def _ensure(self, K: int):
        if K > self.K:
            self._alloc(min(256, max(K, self.K * 2)))

    def _buffers(self, K: int) -> dict:
        """Device buffers and the ``sovlp_bnr_batch`` descriptor of a K-node call."""
        import torch
        calls = self.__dict__.setdefault("_calls", {})
        if K in calls:
            return calls[K]
        m, n = self.m, self.n
        f = lambda *s: torch.zeros(*s, dtype=torch.float64, device="cuda")   # noqa: E731
        c = dict(L=f(n, K), U=f(n, K), X=f(n, K), Y=f(max(m, 1), K), eta=f(K), omega=f(K),
                 bounds=f(K), empty=torch.zeros(K, dtype=torch.int32, device="cuda"))
        nb = C.c_int64()
        _lib.call("sovlp_bnr_workspace_bytes", m, n, K, C.byref(nb))
        c["ws"] = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")
        calls[K] = c
        return c

    def _descriptor(self, K: int, c: dict):
        t = self.t
        return _lib.sovlp_bnr_batch(
            self.m, self.n, K, int(self.cfg.iters), t["rp"].data_ptr(), t["cp"].data_ptr(),
            t["ri"].data_ptr(), t["ci"].data_ptr(), t["rx"].data_ptr(), t["cx"].data_ptr(),
            t["c"].data_ptr(), t["rl"].data_ptr(), t["ru"].data_ptr(), c["L"].data_ptr(),
            c["U"].data_ptr(), c["X"].data_ptr(), c["Y"].data_ptr(), c["eta"].data_ptr(),
            c["omega"].data_ptr(), float(self.cfg.step_growth), float(self.cfg.step_damp),
            c["bounds"].data_ptr(), c["empty"].data_ptr(),
            -1.0 if self.cfg.textbook_duals else 1.0)

    def solve_batch(self, col_lb, col_ub, warm=None) -> BatchResult:
        """Advance ``K`` nodes (``col_lb``, ``col_ub``: ``(n, K)``) and bound them.
        ``warm``: an optional ``(X, Y)`` from an earlier visit."""
        import torch
        col_lb = np.asarray(col_lb, dtype=np.float64)
        col_ub = np.asarray(col_ub, dtype=np.float64)
        if col_lb.ndim != 2 or col_lb.shape[0] != self.n or col_lb.shape != col_ub.shape:
            raise ValueError(f"column bounds must be ({self.n}, K)")
        K = col_lb.shape[1]
        if not 1 <= K <= 256:
            raise ValueError("K must be in [1, 256]")
        self._ensure(K)
        m = self.m
        # the call's device buffers persist per batch size, so repeated calls
        # pass the library the same arguments (it then replays them as a graph)
        c = self._buffers(K)
        L, U, X, Y = c["L"], c["U"], c["X"], c["Y"]
        eta, omega, bounds, empty, ws = c["eta"], c["omega"], c["bounds"], c["empty"], c["ws"]
        L.copy_(torch.from_numpy(np.ascontiguousarray(col_lb)))
        U.copy_(torch.from_numpy(np.ascontiguousarray(col_ub)))
        if warm is not None:
            X.copy_(torch.from_numpy(np.ascontiguousarray(warm[0], dtype=np.float64)))
            if m > 0:
                Y.copy_(torch.from_numpy(np.ascontiguousarray(warm[1], dtype=np.float64)))
        else:
            X.zero_()
            Y.zero_()
        eta.copy_(torch.from_numpy(self.eta[:K].copy()))
        omega.copy_(torch.from_numpy(self.omega[:K].copy()))
        # the stream the uploads above were queued on
        desc = self._descriptor(K, c)
        _lib.call("sovlp_bnr_solve_batch", C.byref(desc), ws.data_ptr(), ws.numel(),
                  torch.cuda.current_stream().cuda_stream)
        self.X[:, :K] = X
        self.Y[:, :K] = Y[: self.Y.shape[0]]
        self.eta[:K] = eta.cpu().numpy()
        self.omega[:K] = omega.cpu().numpy()
        empty_h = empty.cpu().numpy().astype(bool)
        bounds_h = (bounds.cpu().numpy() + self.obj_offset) / self.obj_scale
        bounds_h = np.where(empty_h, np.inf, bounds_h)
        self.total_iterations += self.cfg.iters
        self.total_batches += 1
        self.total_nodes += K
        return BatchResult(bounds=bounds_h, x=X.cpu().numpy(), iterations=self.cfg.iters,
                           infeasible=empty_h)

    def iterates(self, K: int) -> tuple[np.ndarray, np.ndarray]:
        """``(X, Y)`` of the last batch's first ``K`` nodes: a warm start."""
        return self.X[:, :K].cpu().numpy(), self.Y[: self.m, :K].cpu().numpy()
