"""Served PDLP: continuous batching of LP requests on one GPU.

No reference counterpart. SOVOPT solves one LP per ``solve_pdlp`` call
(``src/sovopt/lp/pdlp.py:238``), always from the cold start ``x0 = clip(0)``,
``y0 = 0`` (``pdlp.py:274-275``), and its only memory guard is a host-RAM probe
for the IPM (``src/sovopt/core/memory.py:33-67``). Here:

* ``DeviceServer`` wraps the ``sovlp_server_*`` slot table of ``libsovlp.so``:
  every admitted LP occupies one slot, and all occupied slots advance together
  in each launch of one recorded graph (one workgroup per slot). A request that
  arrives while others run joins the next launch, so its work overlaps theirs
  on the device. An LP's arithmetic depends only on its own data and the
  workgroup size, so a result is the same alone and in any mix.
* ``HbmAdmission`` admits a request only when its device bytes (inputs plus the
  engine workspace) fit in what is left of a budget taken from the device's free
  HBM; requests wait in arrival order otherwise, and one that could never fit is
  refused.
* ``ServeLoop`` runs the request stream: admission into free slots, steps of the
  device, harvesting of finished slots, and warm starts — a resolve of an earlier
  model (same matrix shape, perturbed ``b`` and ``c``) starts from that model's
  last solution ``(x, y)`` and primal weight.
* Latency of a request = the loop's clock when its result is read minus its
  arrival time; ``summarize`` reports p50/p99 and LPs/s.

The loop's host logic (admission, slots, warm-start routing, accounting) talks to
the device through five methods, so the CPU tests drive it with a stand-in device.
"""

from __future__ import annotations

import ctypes as C
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .. import _lib
from ..core.backend import get_backend
from ..core.problem import Problem, Solution, Status
from .batch import batch_geometry
from .pdlp import DeviceLP, PDLPInfo, PDLPParams, refinement_rows

__all__ = ["Request", "Served", "HbmAdmission", "hbm_budget", "DeviceServer", "ServeLoop",
           "summarize", "WarmStart", "resolve_warm", "poisson_blending_requests"]


@dataclass
class WarmStart:
    """A solution to start from: x, y in report units and signs, the primal weight."""
    x: np.ndarray
    y: np.ndarray
    omega: float = 0.0


@dataclass
class Request:
    key: int
    problem: Problem
    arrival: float = 0.0          # seconds after the loop starts
    model: int | None = None      # requests with the same model share warm starts
    resolve: bool = False         # warm-start from the model's last solution


@dataclass
class Served:
    key: int
    solution: Solution | None
    arrival: float
    admitted: float = float("nan")
    finished: float = float("nan")
    slot: int = -1
    nbytes: int = 0
    warm: bool = False
    refused: str = ""
    escalated: bool = False       # re-solved on the fallback device
    first_status: int = -1        # the first device's status when escalated

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class HbmAdmission:
    """Device-memory admission: a budget in bytes, taken and given back per request."""

    def __init__(self, budget: int):
        if budget <= 0:
            raise ValueError("HBM budget must be positive")
        self.budget = int(budget)
        self.used = 0
        self.peak = 0

    def fits(self, nbytes: int) -> bool:
        return self.used + nbytes <= self.budget

    def possible(self, nbytes: int) -> bool:
        return nbytes <= self.budget

    def take(self, nbytes: int) -> None:
        if not self.fits(nbytes):
            raise RuntimeError("admission over the HBM budget")
        self.used += int(nbytes)
        self.peak = max(self.peak, self.used)

    def give(self, nbytes: int) -> None:
        self.used -= int(nbytes)
        if self.used < 0:
            raise RuntimeError("HBM accounting went negative")


def hbm_budget(fraction: float = 0.8, probe=None) -> int:
    """``fraction`` of the device's free HBM now (``probe`` returns (free, total)
    bytes; default ``torch.cuda.mem_get_info``)."""
    if probe is None:
        import torch
        probe = torch.cuda.mem_get_info
    free, _total = probe()
    return int(free * fraction)


def _input_bytes(p: Problem, warm: bool) -> int:
    m, n, z = p.m, p.n, p.A.nnz
    b = (m + 1 + n + 1) * 4 + z * (4 + 8) * 2      # CSR + CSC: int32 pointers / indices, FP64
    b += (3 * n + 2 * m) * 8                       # c, lo, hi, rl, ru
    b += (m + n + 4) * 4                           # tile plans (at most one tile per row)
    if warm:
        b += (m + n) * 8
    return b


class DeviceServer:
    """The ``sovlp_server_*`` slot table with per-request uploads (device memory
    is torch's; each request's tensors live until its slot is released)."""

    def __init__(self, slots: int, params: PDLPParams | None = None, precision: int = 32,
                 block: int = 256, phases_per_launch: int = 32, history_cap: int = 1,
                 team: int = 1):
        """``team`` > 1: every slot runs each phase over ``team`` workgroups of
        256 threads (``sovlp_server_create_team``; ``phases_per_launch`` phases
        per replay as launch pairs), for LPs too large for one workgroup."""
        import torch
        self.params = params or PDLPParams()
        get_backend(self.params.device)
        self.slots, self.team = int(slots), int(team)
        if self.team > 1:
            block = 256
        self.precision, self.block = int(precision), int(block)
        self.history_cap = int(history_cap)
        self.geometry = (self.block,) + tuple(batch_geometry(self.precision, self.block))
        nb = C.c_int64()
        _lib.call("sovlp_server_team_workspace_bytes", self.slots, self.team, self.precision,
                  C.byref(nb))
        self.table = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")
        self.cparams = self.params.to_c(self.precision)
        self.stream = torch.cuda.current_stream().cuda_stream
        h = C.c_void_p()
        if self.team > 1:
            _lib.call("sovlp_server_create_team", self.slots, self.team, C.byref(self.cparams),
                      self.precision, int(phases_per_launch), self.table.data_ptr(), nb.value,
                      self.stream, C.byref(h))
        else:
            _lib.call("sovlp_server_create", self.slots, C.byref(self.cparams), self.precision,
                      self.block, int(phases_per_launch), self.table.data_ptr(), nb.value,
                      self.stream, C.byref(h))
        self.h = h
        hz = C.c_double()
        _lib.call("sovlp_server_clock_hz", self.h, C.byref(hz))
        self.hz = hz.value
        self._v0 = []                 # every v0 ever handed out (kept until close)
        self._live = {}               # slot -> (DeviceLP, workspace, warm tensors)
        self._fin = np.zeros(self.slots, dtype=np.int32)

    # -- v0: default_rng(0).standard_normal(n), one prefix-shared tensor
    def _v0_for(self, n: int):
        import torch
        if not self._v0 or self._v0[-1].numel() < n:
            size = max(n, 2 * (self._v0[-1].numel() if self._v0 else 1024))
            v = np.random.default_rng(0).standard_normal(size)
            self._v0.append(torch.from_numpy(v).to("cuda"))
        return self._v0[-1]

    def _desc_bytes(self, p: Problem) -> int:
        d = _lib.sovlp_lp()
        d.m, d.n, d.nnz = p.m, p.n, p.A.nnz
        nb = C.c_int64()
        _lib.call("sovlp_server_lp_bytes", C.byref(d), self.precision, self.history_cap,
                  C.byref(nb))
        return int(nb.value)

    def request_bytes(self, p: Problem, warm: bool = False) -> int:
        return self._desc_bytes(p) + _input_bytes(p, warm)

    def admit(self, slot: int, p: Problem, warm: WarmStart | None = None) -> None:
        import torch
        dlp = DeviceLP(p, self.precision, geometry=self.geometry, v0=self._v0_for(p.n))
        nb = self._desc_bytes(p)
        ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
        wx = wy = None
        omega = 0.0
        if warm is not None:
            if warm.x.shape != (p.n,) or warm.y.shape != (p.m,):
                raise ValueError("warm start does not match the LP's shape")
            wx = torch.from_numpy(np.ascontiguousarray(warm.x, dtype=np.float64)).to("cuda")
            wy = torch.from_numpy(np.ascontiguousarray(warm.y, dtype=np.float64)).to("cuda")
            omega = float(warm.omega) if np.isfinite(warm.omega) else 0.0
        _lib.call("sovlp_server_admit", self.h, int(slot), C.byref(dlp.desc),
                  wx.data_ptr() if wx is not None else None,
                  wy.data_ptr() if wy is not None else None, omega, ws.data_ptr(), nb,
                  self.history_cap, self.stream)
        self._live[slot] = (dlp, ws, wx, wy)

    def step(self, launches: int = 1) -> list[int]:
        """Advances every occupied slot; returns the slots whose LP has finished."""
        running = C.c_int32()
        _lib.call("sovlp_server_step", self.h, int(launches), self.stream,
                  self._fin.ctypes.data, C.byref(running))
        return [int(s) for s in np.flatnonzero(self._fin)]

    def result(self, slot: int) -> Solution:
        inf = _lib.sovlp_info()
        _lib.call("sovlp_server_info", self.h, int(slot), C.byref(inf))
        dlp, ws, _wx, _wy = self._live[slot]
        m, n = dlp.desc.m, dlp.desc.n
        x, y, d = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _lib.call("sovlp_server_outputs", self.h, int(slot), C.byref(x), C.byref(y), C.byref(d))
        import torch
        base = ws.data_ptr()

        def view(ptr, cnt):
            return ws[ptr - base: ptr - base + 8 * cnt].view(torch.float64)

        out = torch.cat([view(x.value, n), view(y.value, m), view(d.value, n)]).cpu().numpy()
        ref = []
        if inf.refine_rows > 0:
            r = C.c_void_p()
            _lib.call("sovlp_server_refinement", self.h, int(slot), C.byref(r))
            ref = refinement_rows(view(r.value, 5 * int(inf.refine_rows)).cpu().numpy()
                                  .reshape(-1, 5), dlp.flip)
        hist = []
        nh = min(int(inf.history_len), self.history_cap)
        if nh > 0:
            hp = C.c_void_p()
            _lib.call("sovlp_server_history", self.h, int(slot), C.byref(hp))
            rows = view(hp.value, 5 * nh).cpu().numpy().reshape(-1, 5)
            hist = [(int(r[0]), float(r[1]), float(r[2]), float(r[3]), float(r[4])) for r in rows]
        secs = inf.device_ticks / self.hz if self.hz else float("nan")
        st = Status(int(inf.status)) if inf.status >= 0 else Status.NOT_SOLVED
        info = PDLPInfo(iterations=int(inf.iterations), restarts=int(inf.restarts),
                        rejected_steps=int(inf.rejected_steps), primal_res=inf.primal_res,
                        dual_res=inf.dual_res, gap=inf.gap, primal_obj=inf.primal_obj,
                        dual_obj=inf.dual_obj, step_size=inf.step_size,
                        primal_weight=inf.primal_weight, scale_obj=inf.scale_obj,
                        device_seconds=secs, team="served", precision=self.precision,
                        refinement=ref, history=hist)
        sol = Solution(status=st, x=out[:n].copy(), objective=float(inf.objective),
                       y=out[n:n + m].copy(), reduced_costs=out[n + m:].copy(),
                       iterations=int(inf.iterations), time=secs, method="pdlp[gpu,served]")
        sol.dual_bound = sol.objective
        sol.work_units = float(inf.iterations)
        sol.info = info
        return sol.drop_objective_if_unsolved()

    def release(self, slot: int) -> None:
        _lib.call("sovlp_server_release", self.h, int(slot))
        self._live.pop(slot, None)

    def close(self):
        if getattr(self, "h", None) is not None:
            _lib.lib().sovlp_server_destroy(self.h)
            self.h = None
        self._live = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Lane:
    """One device of the loop with its queue, free slots and running requests."""

    def __init__(self, device):
        self.device = device
        self.free = list(range(device.slots))[::-1]
        self.queue: deque = deque()        # (Request, WarmStart | None, Served | None)
        self.active: dict = {}             # slot -> (Served, Request)


_ESCALATE = (int(Status.NUMERICAL), int(Status.ITERATION_LIMIT))


@dataclass
class ServeLoop:
    """Runs a request stream through a device (``DeviceServer`` or a stand-in with
    ``slots``, ``request_bytes``, ``admit``, ``step``, ``result``, ``release``).

    ``fallback`` (optional) is a second device, normally an FP64 server: a request
    the first device ends NUMERICAL or at the iteration limit is queued there,
    from a cold start (``escalate_warm``: from the first device's ``(x, y)`` and
    primal weight), and its latency runs until the fallback's result."""
    device: object
    admission: HbmAdmission
    launches_per_step: int = 1
    clock: object = time.perf_counter
    sleep: object = time.sleep
    warm_starts: bool = True
    fallback: object = None
    # an escalated request restarts cold on the fallback: the point the first
    # device ended NUMERICAL at is a poor start (FP64 PDLP from the FP32 point of
    # a 22 x 6 x 120 blending LP hit 200k iterations, against 121k cold)
    escalate_warm: bool = False
    warm: dict = field(default_factory=dict)
    progress: object = None       # called with a status dict every progress_every seconds
    progress_every: float = 2.0

    def _admit(self, lane: _Lane, t0: float, out: list) -> None:
        adm = self.admission
        # admission in arrival order: the head waits for memory, later requests
        # do not overtake it
        while lane.queue and lane.free:
            req, w, prev = lane.queue[0]
            if w is not None and (w.x.shape != (req.problem.n,) or w.y.shape != (req.problem.m,)):
                w = None
            need = lane.device.request_bytes(req.problem, w is not None)
            if not adm.possible(need):
                lane.queue.popleft()
                s = prev or Served(req.key, None, req.arrival)
                s.nbytes, s.refused = need, f"needs {need} B of HBM, budget {adm.budget} B"
                out.append(s)
                continue
            if not adm.fits(need):
                break
            lane.queue.popleft()
            slot = lane.free.pop()
            lane.device.admit(slot, req.problem, w)
            adm.take(need)
            if prev is None:
                prev = Served(req.key, None, req.arrival, admitted=self.clock() - t0,
                              warm=w is not None)
            prev.slot, prev.nbytes = slot, need
            lane.active[slot] = (prev, req)

    def run(self, requests: list[Request]) -> list[Served]:
        adm = self.admission
        lanes = [_Lane(self.device)] + ([_Lane(self.fallback)] if self.fallback else [])
        order = sorted(requests, key=lambda r: (r.arrival, r.key))
        out: list[Served] = []
        i = 0
        t0 = self.clock()
        next_report = self.progress_every
        while i < len(order) or any(ln.queue or ln.active for ln in lanes):
            now = self.clock() - t0
            if self.progress is not None and now >= next_report:
                next_report = now + self.progress_every
                self.progress(dict(t=now, arrived=i, finished=len(out),
                                   queued=[len(ln.queue) for ln in lanes],
                                   running=[len(ln.active) for ln in lanes],
                                   hbm_used=adm.used))
            while i < len(order) and order[i].arrival <= now:
                r = order[i]
                w = self.warm.get(r.model) if (self.warm_starts and r.resolve) else None
                lanes[0].queue.append((r, w, None))
                i += 1
            for ln in lanes:
                self._admit(ln, t0, out)
            stepped = False
            for li, ln in enumerate(lanes):
                if not ln.active:
                    continue
                stepped = True
                done = ln.device.step(self.launches_per_step)
                now = self.clock() - t0
                for slot in done:
                    s, req = ln.active.pop(slot)
                    sol = ln.device.result(slot)
                    ln.device.release(slot)
                    adm.give(s.nbytes)
                    ln.free.append(slot)
                    st = int(sol.status) if sol is not None else -1
                    if li == 0 and len(lanes) > 1 and st in _ESCALATE and sol.x is not None:
                        s.escalated = True
                        s.first_status = st
                        w = WarmStart(sol.x, sol.y, getattr(sol.info, "primal_weight", 0.0)) \
                            if self.escalate_warm else None
                        lanes[1].queue.append((req, w, s))
                        continue
                    s.solution = sol
                    s.finished = now
                    if (req.model is not None and sol is not None and sol.x is not None
                            and st == int(Status.OPTIMAL)):
                        self.warm[req.model] = WarmStart(sol.x, sol.y,
                                                         getattr(sol.info, "primal_weight", 0.0))
                    out.append(s)
            if not stepped and i < len(order):
                self.sleep(max(0.0, min(order[i].arrival - (self.clock() - t0), 0.05)))
            elif not stepped and any(ln.queue for ln in lanes):
                raise RuntimeError("requests are queued but none can be admitted")
        return out


def summarize(served: list[Served], span: float | None = None) -> dict:
    """p50/p99 latency, LPs/s (finished LPs over the span from the first arrival
    to the last result), outcome counts."""
    ok = [s for s in served if s.solution is not None]
    lat = np.array([s.latency for s in ok]) if ok else np.zeros(0)
    if span is None and ok:
        span = max(s.finished for s in ok) - min(s.arrival for s in served)
    opt = [s for s in ok if int(s.solution.status) == int(Status.OPTIMAL)]
    warm = [s for s in ok if s.warm]
    cold = [s for s in ok if not s.warm]
    return {
        "requests": len(served),
        "finished": len(ok),
        "refused": sum(1 for s in served if s.refused),
        "optimal": len(opt),
        "warm_started": len(warm),
        "escalated": sum(1 for s in ok if s.escalated),
        "lps_per_s": len(ok) / span if span and span > 0 else float("nan"),
        "latency_p50_s": float(np.percentile(lat, 50)) if lat.size else float("nan"),
        "latency_p99_s": float(np.percentile(lat, 99)) if lat.size else float("nan"),
        "mean_iterations_cold": float(np.mean([s.solution.iterations for s in cold]))
        if cold else float("nan"),
        "mean_iterations_warm": float(np.mean([s.solution.iterations for s in warm]))
        if warm else float("nan"),
        "span_s": span,
    }


def resolve_warm(device, problems: list[Problem], warm: list[WarmStart],
                 budget: int = 1 << 40) -> list[Solution]:
    """Solve ``problems`` on ``device`` (normally an FP64 ``DeviceServer``), each
    warm-started from ``warm[i]``: the escalation of LPs that a batch solved in
    FP32 ended NUMERICAL or at the iteration limit."""
    loop = ServeLoop(device, HbmAdmission(budget))
    reqs = []
    for i, (p, w) in enumerate(zip(problems, warm)):
        loop.warm[i] = w
        reqs.append(Request(i, p, 0.0, model=i, resolve=True))
    out = {s.key: s for s in loop.run(reqs)}
    return [out[i].solution for i in range(len(problems))]


def poisson_blending_requests(rate: float, duration: float, seed: int, resolve: float, tmin: int,
                              tmax: int, kmax: int = 60, pmax: int = 20) -> list[Request]:
    """A seeded Poisson stream of blending requests, every model built up front:
    arrivals at ``rate`` per second for ``duration`` seconds; a fraction
    ``resolve`` re-solves an earlier model with ``b`` and ``c`` perturbed by 1 %
    (warm-startable); the rest are new models with ``K ~ U{20..kmax}``,
    ``P ~ U{5..pmax}``, ``T ~ U{tmin..tmax}``."""
    from ..workloads import blending_lp, perturbed
    rng = np.random.default_rng(seed)
    t, key = 0.0, 0
    reqs, models = [], {}
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        if models and rng.random() < resolve:
            mk = int(rng.choice(list(models)))
            p = perturbed(models[mk], int(rng.integers(1 << 30)), 0.01)
            reqs.append(Request(key, p, t, model=mk, resolve=True))
        else:
            K = int(rng.integers(20, kmax + 1))
            P = int(rng.integers(5, pmax + 1))
            T = int(rng.integers(tmin, tmax + 1))
            p = blending_lp(int(rng.integers(1 << 30)), K=K, P=P, T=T)
            models[key] = p
            reqs.append(Request(key, p, t, model=key))
        key += 1
    return reqs
