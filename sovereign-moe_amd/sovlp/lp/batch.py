"""Batched PDLP: many independent, uneven LPs solved together on one GPU.

No reference counterpart: SOVOPT solves one LP per ``solve_pdlp`` call
(``src/sovopt/lp/pdlp.py:238``); its only batched PDHG (``src/sovopt/mip/bnr.py:176-301``)
shares one ``A`` across branch-and-bound nodes. Here every LP keeps its own
matrix, scaling, step size, restarts and termination test — exactly the
single-LP algorithm — and the device runs one workgroup per LP per launch,
every launch advancing each unfinished LP by one phase. The launches are
replayed from a hipGraph until every LP has terminated.

An LP's arithmetic depends only on its own data and on the workgroup size,
which the autotuner (``sovlp.lp.autotune``) picks from the LP's own shape, so
a result is the same whichever other LPs share the batch.

All arrays of the batch live in a handful of concatenated device tensors;
``sovlp_lp`` descriptors point into them.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..core.backend import get_backend
from ..core.problem import ObjSense, Problem, Solution, Status
from .pdlp import PDLPInfo, PDLPParams, plan_tiles, refinement_rows

__all__ = ["BatchResult", "BatchEngine", "solve_batch", "DeviceBatch", "batch_geometry",
           "lds_mode", "lds_resident"]


def phases_per_launch() -> int:
    """Phases a run-kernel launch advances each LP by, at most."""
    v = C.c_int()
    _lib.call("sovlp_phases_per_launch", C.byref(v))
    return v.value


def batch_geometry(precision: int, block: int) -> tuple[int, int]:
    r, t = C.c_int(), C.c_int()
    _lib.call("sovlp_batch_geometry", int(precision), int(block), C.byref(r), C.byref(t))
    return r.value, t.value


def lds_mode(precision: int, block: int, m: int, n: int, nnz: int, tiles_a: int,
             tiles_t: int) -> int:
    """What the run kernel on ``block`` threads holds in LDS for the PDHG loop
    of an LP of this shape (the device's own rule): 1 its matrix and step
    vectors, 2 only its step vectors, 0 nothing."""
    r = C.c_int()
    _lib.call("sovlp_lds_resident", int(precision), int(block), int(m), int(n), int(nnz),
              int(tiles_a), int(tiles_t), C.byref(r))
    return int(r.value)


def lds_resident(precision: int, block: int, m: int, n: int, nnz: int, tiles_a: int,
                 tiles_t: int) -> bool:
    """Whether the run kernel holds an LP of this shape, matrix and all, in LDS."""
    return lds_mode(precision, block, m, n, nnz, tiles_a, tiles_t) == 1


class DeviceBatch:
    """Uploads many LPs into concatenated device tensors and builds their
    descriptors (minimisation form; costs negated for maximisations)."""

    def __init__(self, problems: list[Problem], precision: int = 32, block: int = 256,
                 configs: list | None = None):
        """``configs``: per LP (block, tile rows, tile nonzeros), e.g. an
        autotuner's choice; default: ``block`` with its full geometry."""
        import torch
        self.count = len(problems)
        if configs is None:
            rows, tnnz = batch_geometry(precision, block)
            configs = [(int(block), rows, tnnz)] * self.count
        if len(configs) != self.count:
            raise ValueError("one kernel configuration per LP")
        for b, r, t in set(configs):
            rmax, tmax = batch_geometry(precision, b)
            if not (1 <= r <= rmax and 1 <= t <= tmax):
                raise ValueError(f"tile geometry ({r} rows, {t} nonzeros) exceeds block {b}'s")
        self.configs = [tuple(int(v) for v in c) for c in configs]
        self.blocks = np.array([c[0] for c in self.configs], dtype=np.int32)
        nmax = max(p.n for p in problems)
        v0 = np.random.default_rng(0).standard_normal(nmax)   # each LP uses a prefix
        cat = {k: [] for k in ("rp", "ri", "rx", "cp", "ci", "cx", "c", "rl", "ru", "lo", "hi",
                               "ta", "tt")}
        meta = []
        for p, (blk, rows, tnnz) in zip(problems, self.configs):
            A = p.A
            flip = p.sense == ObjSense.MAXIMISE
            ta = plan_tiles(A.rp, tnnz, rows, blk)
            tt = plan_tiles(A.cp, tnnz, rows, blk)
            cat["rp"].append(A.rp.astype(np.int32))
            cat["ri"].append(A.ri)
            cat["rx"].append(A.rx)
            cat["cp"].append(A.cp.astype(np.int32))
            cat["ci"].append(A.ci)
            cat["cx"].append(A.cx)
            cat["c"].append(-p.c if flip else p.c)
            cat["rl"].append(p.row_lb)
            cat["ru"].append(p.row_ub)
            cat["lo"].append(p.col_lb)
            cat["hi"].append(p.col_ub)
            cat["ta"].append(ta)
            cat["tt"].append(tt)
            meta.append((A.m, A.n, A.nnz, ta.size - 1, tt.size - 1,
                         -p.obj_offset if flip else p.obj_offset, int(flip)))
        dt = {"rp": np.int32, "ri": np.int32, "cp": np.int32, "ci": np.int32, "ta": np.int32,
              "tt": np.int32}
        self.t = {}
        offs = {}
        for k, parts in cat.items():
            arr = np.concatenate(parts).astype(dt.get(k, np.float64), copy=False)
            self.t[k] = torch.from_numpy(np.ascontiguousarray(arr)).to("cuda")
            offs[k] = np.concatenate([[0], np.cumsum([a.size for a in parts])])
        self.t["v0"] = torch.from_numpy(v0).to("cuda")
        descs = (_lib.sovlp_lp * self.count)()
        for i, (m, n, nnz, nta, ntt, off, flip) in enumerate(meta):
            def ptr(k, elem):
                return self.t[k].data_ptr() + int(offs[k][i]) * elem
            descs[i] = _lib.sovlp_lp(m, n, nnz, ptr("rp", 4), ptr("ri", 4), ptr("rx", 8),
                                     ptr("cp", 4), ptr("ci", 4), ptr("cx", 8), ptr("c", 8),
                                     ptr("rl", 8), ptr("ru", 8), ptr("lo", 8), ptr("hi", 8),
                                     self.t["v0"].data_ptr(), ptr("ta", 4), nta, ptr("tt", 4),
                                     ntt, float(off), flip)
        self.descs = descs
        self.meta = meta


@dataclass
class BatchResult:
    solutions: list
    seconds: float            # wall time of the device solve (graph replays), inputs resident
    graphs: int               # graph replays
    lps_per_second: float
    latency_p50: float        # per-LP device time from its first phase to its last
    latency_p99: float
    precision: int
    block: int


class BatchEngine:
    """A batch laid out on the device with its phase graph recorded; ``run``
    solves every LP from scratch (the same batch can be solved many times)."""

    def __init__(self, problems: list[Problem], params: PDLPParams | None = None,
                 precision: int = 32, block: int = 256, history_cap: int = 1,
                 tuner=None, configs: list | None = None):
        """``tuner`` (an ``sovlp.lp.autotune.Autotuner``) picks each LP's kernel
        configuration from its own shape; ``configs`` gives them explicitly;
        neither: every LP uses ``block`` threads and its full tile geometry."""
        import torch
        self.params = params or PDLPParams()
        get_backend(self.params.device)
        self.precision, self.block = int(precision), int(block)
        if configs is None and tuner is not None:
            configs = [tuner.choose(p) for p in problems]
        self.db = DeviceBatch(problems, precision, block, configs)
        self.count = self.db.count
        nbytes = C.c_int64()
        _lib.call("sovlp_batch_workspace_bytes", self.count, self.db.descs, self.precision,
                  int(history_cap), C.byref(nbytes))
        self.ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device="cuda")
        self.cparams = self.params.to_c(self.precision)
        self.stream = torch.cuda.current_stream().cuda_stream
        h = C.c_void_p()
        _lib.call("sovlp_batch_create_mixed", self.count, self.db.descs, C.byref(self.cparams),
                  self.precision, self.db.blocks.ctypes.data, self.ws.data_ptr(), nbytes.value,
                  int(history_cap), self.stream, C.byref(h))
        self.h = h
        # a bound on the replays of one solve (an iteration takes at most 61
        # phases; the device enforces the iteration and time limits itself)
        self.max_graphs = (int(self.params.max_iter) * 61 + 400) // phases_per_launch() + 16

    def run(self) -> int:
        """Solves the whole batch (synchronous); returns the graph replays."""
        g = C.c_int64()
        _lib.call("sovlp_batch_run", self.h, float(self.params.time_limit), self.max_graphs,
                  self.stream, C.byref(g))
        return int(g.value)

    def infos(self):
        infos = (_lib.sovlp_info * self.count)()
        _lib.call("sovlp_batch_infos", self.h, infos, self.stream)
        return infos

    def outputs_of(self, index: int, infos=None) -> tuple[np.ndarray, np.ndarray, float]:
        """(x, y, primal weight) of LP ``index`` as reported (original units,
        textbook signs): the warm start of a re-solve. ``infos``: a recent
        ``infos()`` result, to skip reading the states again."""
        import torch
        x, y, d = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _lib.call("sovlp_batch_outputs", self.h, int(index), C.byref(x), C.byref(y), C.byref(d))
        m, n = self.db.meta[index][0], self.db.meta[index][1]
        base = self.ws.data_ptr()
        xs = self.ws[x.value - base: x.value - base + 8 * n].view(torch.float64)
        ys = self.ws[y.value - base: y.value - base + 8 * m].view(torch.float64)
        infos = self.infos() if infos is None else infos
        return xs.cpu().numpy().copy(), ys.cpu().numpy().copy(), float(infos[index].primal_weight)

    def phase_counts(self) -> np.ndarray:
        slots = C.c_int()
        _lib.call("sovlp_phase_slots", C.byref(slots))
        out = np.zeros((self.count, slots.value), dtype=np.int64)
        _lib.call("sovlp_batch_phase_counts", self.h, out.ctypes.data, self.stream)
        return out

    def phase_ticks(self) -> np.ndarray:
        """(count, slots) device clock ticks each LP spent in each phase."""
        slots = C.c_int()
        _lib.call("sovlp_phase_slots", C.byref(slots))
        out = np.zeros((self.count, slots.value), dtype=np.int64)
        _lib.call("sovlp_batch_phase_ticks", self.h, out.ctypes.data, self.stream)
        return out

    def lds_modes(self) -> np.ndarray:
        """Per LP: what the run kernel holds in LDS for its PDHG loop, 1 its
        matrix and step vectors, 2 only its step vectors, 0 nothing (always 0
        with ``PDLPParams(lds_matrix=False)``)."""
        if not self.params.lds_matrix:
            return np.zeros(self.count, dtype=np.int32)
        return np.array([lds_mode(self.precision, int(b), m, n, nnz, ta, tt)
                         for (m, n, nnz, ta, tt, _o, _f), b in zip(self.db.meta, self.db.blocks)],
                        dtype=np.int32)

    def resident(self) -> np.ndarray:
        """Per LP: whether its matrix is held in LDS for the PDHG loop."""
        return self.lds_modes() == 1

    def sizes(self) -> list[tuple[int, int, int, int, int]]:
        """(m, n, nnz, tiles of A, tiles of A^T) of every LP."""
        return [(m, n, nnz, ta, tt) for (m, n, nnz, ta, tt, _o, _f) in self.db.meta]

    def solutions(self, keep_outputs: bool = True) -> tuple[list, np.ndarray]:
        """Per-LP ``Solution`` objects and device latencies (seconds)."""
        import torch
        infos = self.infos()
        hz = C.c_double()
        _lib.call("sovlp_clock_rate", 0, C.byref(hz))
        if keep_outputs:
            xs, ys, ds = [], [], []
            ws, base = self.ws, self.ws.data_ptr()
            for i in range(self.count):
                x, y, d = C.c_void_p(), C.c_void_p(), C.c_void_p()
                _lib.call("sovlp_batch_outputs", self.h, i, C.byref(x), C.byref(y), C.byref(d))
                m, n = self.db.meta[i][0], self.db.meta[i][1]
                xs.append(ws[x.value - base: x.value - base + 8 * n])
                ys.append(ws[y.value - base: y.value - base + 8 * m])
                ds.append(ws[d.value - base: d.value - base + 8 * n])
            xall = torch.cat(xs).view(torch.float64).cpu().numpy()
            yall = torch.cat(ys).view(torch.float64).cpu().numpy()
            dall = torch.cat(ds).view(torch.float64).cpu().numpy()
        ox = oy = 0
        sols, lat = [], []
        for i in range(self.count):
            inf = infos[i]
            m, n = self.db.meta[i][0], self.db.meta[i][1]
            secs = inf.device_ticks / hz.value if hz.value else float("nan")
            lat.append(secs)
            st = Status(int(inf.status)) if inf.status >= 0 else Status.NOT_SOLVED
            ref = []
            if inf.refine_rows > 0:
                r = C.c_void_p()
                _lib.call("sovlp_batch_refinement", self.h, i, C.byref(r))
                off = r.value - self.ws.data_ptr()
                rows = self.ws[off: off + 40 * int(inf.refine_rows)].view(torch.float64)
                ref = refinement_rows(rows.cpu().numpy().reshape(-1, 5), bool(self.db.meta[i][6]))
            info = PDLPInfo(iterations=int(inf.iterations), restarts=int(inf.restarts),
                            rejected_steps=int(inf.rejected_steps), primal_res=inf.primal_res,
                            dual_res=inf.dual_res, gap=inf.gap, primal_obj=inf.primal_obj,
                            dual_obj=inf.dual_obj, step_size=inf.step_size,
                            primal_weight=inf.primal_weight, scale_obj=inf.scale_obj,
                            device_seconds=secs, team="batch", precision=self.precision,
                            refinement=ref)
            sol = Solution(status=st, objective=float(inf.objective),
                           iterations=int(inf.iterations), time=secs, method="pdlp[gpu,batch]")
            if keep_outputs:
                sol.x = xall[ox: ox + n].copy()
                sol.reduced_costs = dall[ox: ox + n].copy()
                sol.y = yall[oy: oy + m].copy()
                ox += n
                oy += m
            sol.dual_bound = sol.objective
            sol.work_units = float(inf.iterations)
            sol.info = info
            sols.append(sol.drop_objective_if_unsolved())
        return sols, np.array(lat)

    def close(self):
        if getattr(self, "h", None) is not None:
            _lib.lib().sovlp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batch(problems: list[Problem], params: PDLPParams | None = None, precision: int = 32,
                block: int = 256, history_cap: int = 1, repeat: int = 1,
                keep_outputs: bool = True, tuner=None) -> BatchResult:
    """Solve every LP of ``problems`` on the GPU in one batch.

    ``repeat`` > 1 solves the same batch again from scratch (timing); ``tuner``
picks each LP's kernel configuration (``sovlp.lp.autotune``). Returns
    per-LP ``Solution`` objects (same fields as ``solve_pdlp``'s, method
    ``"pdlp[gpu,batch]"``) and the batch throughput.
    """
    import torch
    eng = BatchEngine(problems, params, precision, block, history_cap, tuner=tuner)
    try:
        best, graphs = None, 0
        for _ in range(max(1, repeat)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            graphs = eng.run()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        sols, lat = eng.solutions(keep_outputs)
    finally:
        eng.close()
    return BatchResult(sols, best, graphs, eng.count / best,
                       float(np.percentile(lat, 50)), float(np.percentile(lat, 99)),
                       int(precision), int(block))
