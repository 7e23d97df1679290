"""PDLP (restarted-average PDHG) on the GPU: the drop-in ``solve_pdlp``.

Same API, parameters, status codes and outputs as the reference
(``src/sovopt/lp/pdlp.py:79-132, 238-494``):

* ``solve_pdlp(prob, params) -> Solution`` with ``method = "pdlp[gpu]"``;
* textbook dual signs on output, ``d = c - Aᵀy`` (``pdlp.py:24-26, 470-475, 484``);
* ``dual_bound = objective`` and ``work_units = iterations`` (``pdlp.py:490-492``);
* ``drop_objective_if_unsolved`` on exit (``pdlp.py:494``).

The whole solve — scaling, step-size search, restarts, primal weight, the
termination test on the unscaled problem and the report — runs on the device in
``libsovlp.so``; the host only uploads the model, launches, and reads the
result. Two extensions, both defaulted to the reference behaviour:
``precision`` (64: FP64 iterates as in the reference; 32: FP32 matrix and
iterates with FP64 reductions) and ``team`` ("auto", "block", "grid").
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from .. import _lib
from ..core.backend import get_backend
from ..core.problem import ObjSense, Problem, Solution, Status
from ..core.sparse import locality_order, relabel_columns

__all__ = ["LARGE_NNZ", "PDLPParams", "PDLPInfo", "solve_pdlp", "DeviceLP", "upload_problem"]


# An LP of this many nonzeros or more is "large": its columns default to the
# locality order and its refinement rounds to a looser tolerance.  On the
# 200k x 1M skewed LP (8.3M nonzeros, seeds 1-3, device seconds) FP32 +
# refinement took 1.08 / 1.62 / 1.36 s with both, against 3.58 / 2.26 / 3.64 s
# in the given order at 3e-4; FP64 took 1.94 / 1.58 / 2.02 s (locality) and
# 1.87 / 2.29 / 2.35 s (given).
LARGE_NNZ = 1 << 20


@dataclass
class PDLPParams:
    eps_abs: float = 1e-8
    eps_rel: float = 1e-8
    feas_cap: float = 1e-6
    max_iter: int = 200_000
    time_limit: float = 600.0
    check_every: int = 64
    stepsize_every: int = 8
    beta_sufficient: float = 0.2
    beta_necessary: float = 0.8
    beta_artificial: float = 0.36
    primal_weight_smoothing: float = 0.5
    scaling: str = "pdlp"
    device: str = "auto"
    verbose: bool = False
    log_every: int = 2000
    # extensions of this engine
    precision: int = 64
    team: str = "auto"
    chunk_iters: int = 8192
    # iterative refinement of the FP32 engine (None: 6 rounds at precision
    # 32, none at 64; each round to relative KKT max(eps_rel, 3e-4), which
    # FP32 iterates reach, or max(eps_rel, 1e-3) for an LP of LARGE_NNZ or
    # more nonzeros, whose rounds then stop sooner and zoom further; the
    # rounds together reach eps and feas_cap).  The stall lengths trade LPs
    # finished for time: on 4096 uneven LPs the defaults finish 99.5 % in
    # 0.82 s, refine_stall_open=1024 98.2 % in 0.33 s.
    refine_rounds: int | None = None
    refine_eps: float | None = None
    refine_tol: float = 1e-9
    # iterations without a halving of a round's KKT ratio that end it: rounds
    # that polish an X already meeting the criterion (0: 1024), and the rounds
    # before that (0: 4096)
    refine_stall: int = 0
    refine_stall_open: int = 0
    # "locality": solve with the columns relabelled by
    # ``sovlp.core.sparse.locality_order`` (long columns first), which puts the
    # gathers of A x on the long columns into a short prefix of x; x and the
    # reduced costs come back in the caller's order. A x is bit for bit that of
    # the given order; A^T y (whose tiles move with the columns) and the sums
    # over columns agree to rounding.  "auto": "locality" for an LP of
    # LARGE_NNZ or more nonzeros (x no longer fits one XCD's L2 and the
    # gathers dominate A x), "given" below.
    column_order: str = "auto"
    # The one-workgroup engines (batch, single LP on one workgroup, serving
    # slots) hold an LP's scaled matrix, tile plans and step vectors in the
    # workgroup's LDS for the PDHG loop when they fit its staging buffer (about
    # 5000 nonzeros on 512-thread workgroups in FP32, 1300 on 128-thread ones):
    # the loop then reads the matrix from LDS instead of memory.  The results
    # are bit for bit the same; False keeps every matrix in memory.
    lds_matrix: bool = True

    def resolved_refine(self, precision: int | None = None,
                        nnz: int | None = None) -> tuple[int, float]:
        """(rounds, per-round tolerance) of the refinement; ``nnz``: the LP's
        nonzeros (None: a small LP)."""
        prec = int(self.precision if precision is None else precision)
        rounds = self.refine_rounds
        if rounds is None:
            rounds = 6 if prec == 32 else 0
        floor = 1e-3 if nnz is not None and nnz >= LARGE_NNZ else 3e-4
        eps = self.refine_eps if self.refine_eps is not None else max(self.eps_rel, floor)
        return int(rounds), float(eps)

    def resolved_column_order(self, nnz: int) -> str:
        if self.column_order == "auto":
            return "locality" if nnz >= LARGE_NNZ else "given"
        return self.column_order

    def to_c(self, precision: int | None = None, nnz: int | None = None) -> _lib.sovlp_params:
        rounds, eps = self.resolved_refine(precision, nnz)
        return _lib.sovlp_params(self.eps_abs, self.eps_rel, self.feas_cap, int(self.max_iter),
                                 float(self.time_limit), int(self.check_every),
                                 int(self.stepsize_every), self.beta_sufficient,
                                 self.beta_necessary, self.beta_artificial,
                                 self.primal_weight_smoothing, rounds,
                                 int(self.refine_stall), eps,
                                 float(self.refine_tol), int(self.refine_stall_open),
                                 0 if self.lds_matrix else 1)


@dataclass
class PDLPInfo:
    iterations: int = 0
    restarts: int = 0
    rejected_steps: int = 0
    primal_res: float = float("nan")
    dual_res: float = float("nan")
    gap: float = float("nan")
    primal_obj: float = float("nan")
    dual_obj: float = float("nan")
    step_size: float = float("nan")
    primal_weight: float = float("nan")
    device: str = "gpu"
    history: list = field(default_factory=list)
    scale_obj: float = float("nan")
    device_seconds: float = float("nan")
    team: str = ""
    workgroups: int = 1
    precision: int = 64
    # per refinement round: (iterations so far, primal residual, dual
    # residual, gap, objective) of the accumulated X on the original problem
    refinement: list = field(default_factory=list)


_TEAMS = {"auto": 0, "block": 1, "grid": 2}


def tile_geometry(precision: int) -> tuple[int, int, int]:
    """(workgroup size, tile row cap, tile nonzero cap) of the single-LP engine."""
    b, r, t = C.c_int(), C.c_int(), C.c_int()
    _lib.call("sovlp_tile_geometry", int(precision), C.byref(b), C.byref(r), C.byref(t))
    return b.value, r.value, t.value


def plan_tiles(ptr: np.ndarray, tile_nnz: int, tile_rows: int, block: int = 0) -> np.ndarray:
    """Host tile plan of a CSR pointer array (``sovlp_plan_tiles``); ``block``:
    lanes of the workgroup that runs it (0: no cap on a lane's serial work)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    rows = ptr.size - 1
    out = np.zeros(rows + 1, dtype=np.int32)
    nt = C.c_int32()
    _lib.call("sovlp_plan_tiles", rows, ptr.ctypes.data, int(tile_nnz), int(tile_rows),
              int(block), out.ctypes.data, C.byref(nt))
    return out[: nt.value + 1].copy()


class DeviceLP:
    """A problem's arrays on the device in the layout the engines take, with the
    ``sovlp_lp`` descriptor pointing at them. Keeps the tensors alive."""

    def __init__(self, prob: Problem, precision: int = 64,
                 geometry: tuple[int, int, int] | None = None, v0=None):
        """``geometry`` = (workgroup size, tile rows, tile nonzeros) of the engine
        that will run the LP (default: the single-LP engine's); ``v0`` a device
        tensor whose first ``n`` entries are ``default_rng(0).standard_normal(n)``
        (default: a fresh one)."""
        import torch
        flip = prob.sense == ObjSense.MAXIMISE
        A = prob.A
        if A.nnz >= 2**31 - 1:
            raise ValueError("more than 2^31-1 nonzeros")
        block, rows, tnnz = geometry if geometry is not None else tile_geometry(precision)
        self.tiles_a = plan_tiles(A.rp, tnnz, rows, block)
        self.tiles_t = plan_tiles(A.cp, tnnz, rows, block)
        dev = "cuda"

        def up(a, dt):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)

        c = -prob.c if flip else prob.c
        self.flip = flip
        self.t = dict(
            rp=up(A.rp, np.int32), ri=up(A.ri, np.int32), rx=up(A.rx, np.float64),
            cp=up(A.cp, np.int32), ci=up(A.ci, np.int32), cx=up(A.cx, np.float64),
            c=up(c, np.float64), rl=up(prob.row_lb, np.float64), ru=up(prob.row_ub, np.float64),
            lo=up(prob.col_lb, np.float64), hi=up(prob.col_ub, np.float64),
            tiles_a=up(self.tiles_a, np.int32), tiles_t=up(self.tiles_t, np.int32))
        if v0 is None:
            v0 = up(np.random.default_rng(0).standard_normal(A.n), np.float64)
        elif v0.numel() < A.n:
            raise ValueError("v0 shorter than the LP's column count")
        self.t["v0"] = v0
        t = self.t
        self.desc = _lib.sovlp_lp(
            A.m, A.n, A.nnz, t["rp"].data_ptr(), t["ri"].data_ptr(), t["rx"].data_ptr(),
            t["cp"].data_ptr(), t["ci"].data_ptr(), t["cx"].data_ptr(), t["c"].data_ptr(),
            t["rl"].data_ptr(), t["ru"].data_ptr(), t["lo"].data_ptr(), t["hi"].data_ptr(),
            t["v0"].data_ptr(), t["tiles_a"].data_ptr(), self.tiles_a.size - 1,
            t["tiles_t"].data_ptr(), self.tiles_t.size - 1,
            float(-prob.obj_offset if flip else prob.obj_offset), 1 if flip else 0)


def refinement_rows(v: np.ndarray, flip: bool) -> list:
    """Rows of a device refinement record, objective in the caller's sense."""
    return [(int(r[0]), float(r[1]), float(r[2]), float(r[3]), float(-r[4] if flip else r[4]))
            for r in v]


def upload_problem(prob: Problem, precision: int = 64) -> DeviceLP:
    return DeviceLP(prob, precision)


class Engine:
    """One device solve (``sovlp_pdlp_*``) over a ``DeviceLP``."""

    def __init__(self, dlp: DeviceLP, params: PDLPParams, history_cap: int | None = None):
        import torch
        self.dlp = dlp
        self.params = params
        prec = int(params.precision)
        if prec not in (64, 32):
            raise ValueError("precision must be 64 or 32")
        cap = history_cap if history_cap is not None else \
            params.max_iter // max(params.check_every, 1) + 2
        self.hist_cap = int(cap)
        nbytes = C.c_int64()
        _lib.call("sovlp_pdlp_workspace_bytes", C.byref(dlp.desc), prec, self.hist_cap,
                  C.byref(nbytes))
        self.ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device="cuda")
        self.cparams = params.to_c(prec, int(dlp.desc.nnz))
        h = C.c_void_p()
        if params.team not in _TEAMS:
            raise ValueError(f"team must be one of {sorted(_TEAMS)}")
        self.stream = torch.cuda.current_stream().cuda_stream
        _lib.call("sovlp_pdlp_create", C.byref(dlp.desc), C.byref(self.cparams), prec,
                  self.ws.data_ptr(), nbytes.value, self.hist_cap, _TEAMS[params.team],
                  self.stream, C.byref(h))
        self.h = h

    def advance(self, iters: int, seconds_left: float) -> None:
        """Replays the phase graph until ``iters`` more iterations ran or the solve
        ended (synchronous)."""
        _lib.call("sovlp_pdlp_advance", self.h, int(iters), float(seconds_left), self.stream)

    def width(self) -> int:
        w = C.c_int()
        _lib.call("sovlp_pdlp_width", self.h, C.byref(w))
        return w.value

    def info(self) -> _lib.sovlp_info:
        inf = _lib.sovlp_info()
        _lib.call("sovlp_pdlp_info", self.h, C.byref(inf), self.stream)
        return inf

    def _view(self, ptr: int, count: int):
        import torch
        off = ptr - self.ws.data_ptr()
        return self.ws[off: off + 8 * count].view(torch.float64)

    def outputs(self):
        x, y, d, h = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        _lib.call("sovlp_pdlp_outputs", self.h, C.byref(x), C.byref(y), C.byref(d), C.byref(h))
        m, n = self.dlp.desc.m, self.dlp.desc.n
        return self._view(x.value, n), self._view(y.value, m), self._view(d.value, n), \
            self._view(h.value, 5 * self.hist_cap)

    def scaling(self) -> tuple[np.ndarray, np.ndarray]:
        """Row and column factors of the device's Ruiz + Pock-Chambolle scaling
        (FP64 powers of two, as ``scale_problem(prob, "pdlp")`` returns them,
        reference ``numerics/scaling.py:374-478``), once the solve has passed
        its scaling phases."""
        r, c = C.c_void_p(), C.c_void_p()
        _lib.call("sovlp_pdlp_scaling", self.h, C.byref(r), C.byref(c))
        m, n = self.dlp.desc.m, self.dlp.desc.n
        return (self._view(r.value, m).cpu().numpy().copy(),
                self._view(c.value, n).cpu().numpy().copy())

    def refinement(self, rows: int) -> list:
        """The refinement record (``rows`` rows, as ``sovlp_pdlp_refinement``)."""
        if rows <= 0:
            return []
        r = C.c_void_p()
        _lib.call("sovlp_pdlp_refinement", self.h, C.byref(r))
        v = self._view(r.value, 5 * rows).cpu().numpy().reshape(rows, 5)
        return refinement_rows(v, self.dlp.flip)

    def close(self):
        if getattr(self, "h", None) is not None:
            _lib.lib().sovlp_pdlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_pdlp(prob: Problem, params: PDLPParams | None = None) -> Solution:
    """Solve an LP with restarted average PDHG on the GPU."""
    params = params or PDLPParams()
    t0 = time.perf_counter()
    if params.scaling != "pdlp":
        raise ValueError("only scaling='pdlp' (Ruiz + Pock-Chambolle) is provided")
    if params.column_order not in ("auto", "given", "locality"):
        raise ValueError("column_order is 'auto', 'given' or 'locality'")
    bk = get_backend(params.device)          # raises without a usable GPU
    new = None
    v0 = None
    if params.resolved_column_order(int(prob.A.nnz)) == "locality" and prob.n > 0:
        import torch
        new = locality_order(prob.A)
        order = np.argsort(new)
        prob = Problem(A=relabel_columns(prob.A, new), c=prob.c[order], row_lb=prob.row_lb,
                       row_ub=prob.row_ub, col_lb=prob.col_lb[order], col_ub=prob.col_ub[order],
                       kind=prob.kind[order], obj_offset=prob.obj_offset, sense=prob.sense,
                       name=prob.name)
        # the power iteration's start vector, relabelled with the columns
        v0 = torch.from_numpy(np.random.default_rng(0).standard_normal(prob.n)[order]).cuda()
    dlp = DeviceLP(prob, params.precision, v0=v0)
    eng = Engine(dlp, params)
    width = eng.width()
    try:
        last = None
        stalls = 0
        while True:
            left = params.time_limit - (time.perf_counter() - t0)
            eng.advance(params.chunk_iters, max(left, 0.0))
            inf = eng.info()
            if inf.stage == 3:
                break
            # every replay runs phases, so the state must move; guard anyway
            key = (inf.iterations, inf.stage, inf.restarts, inf.rejected_steps)
            stalls = stalls + 1 if key == last else 0
            last = key
            if stalls > 1000 or time.perf_counter() - t0 > params.time_limit + 600.0:
                raise _lib.SovlpError(f"device solve made no progress (state {key})")
            if params.verbose:
                print(f"  it {inf.iterations:>7d}  restarts {inf.restarts}  "
                      f"eta {inf.step_size:.2e}  w {inf.primal_weight:.2e}")
        x, y, d, hist = eng.outputs()
        x_h = x.cpu().numpy().copy()
        y_h = y.cpu().numpy().copy()
        d_h = d.cpu().numpy().copy()
        if new is not None:
            x_h, d_h = x_h[new], d_h[new]
        nh = min(int(inf.history_len), eng.hist_cap)
        hist_h = hist[: 5 * nh].cpu().numpy().reshape(nh, 5) if nh else np.zeros((0, 5))
        ref_rows = eng.refinement(int(inf.refine_rows))
    finally:
        eng.close()
    status = Status(int(inf.status))
    hz = C.c_double()
    _lib.call("sovlp_clock_rate", 0, C.byref(hz))
    info = PDLPInfo(iterations=int(inf.iterations), restarts=int(inf.restarts),
                    rejected_steps=int(inf.rejected_steps), primal_res=inf.primal_res,
                    dual_res=inf.dual_res, gap=inf.gap, primal_obj=inf.primal_obj,
                    dual_obj=inf.dual_obj, step_size=inf.step_size,
                    primal_weight=inf.primal_weight, device=bk.device,
                    history=[(int(r[0]), float(r[1]), float(r[2]), float(r[3]), float(r[4]))
                             for r in hist_h],
                    scale_obj=inf.scale_obj,
                    device_seconds=inf.device_ticks / hz.value if hz.value else float("nan"),
                    team="grid" if width > 1 else "block", workgroups=width,
                    precision=int(params.precision), refinement=ref_rows)
    sol = Solution(status=status, x=x_h, objective=float(inf.objective), y=y_h,
                   reduced_costs=d_h, iterations=int(inf.iterations),
                   time=time.perf_counter() - t0, method=f"pdlp[{bk.device}]")
    sol.log = info.history
    sol.dual_bound = sol.objective
    sol.work_units = float(inf.iterations)
    sol.info = info
    return sol.drop_objective_if_unsolved()
