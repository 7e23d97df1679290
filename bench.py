#!/usr/bin/env python3
"""Throughput of batched PDLP on one MI355X: LPs/s solved to relative KKT 1e-6.

Workload: 4096 independent synthetic LPs with known optima (seeds 1000..5095),
m ~ log-uniform[20, 800] with 60 % equality rows, n = U[2,4] m, 3-8 nonzeros
per column, values N(0,1), bounds [0, U(5,50)] per LP.  A step solves the
whole batch once, from scratch (scaling, power iteration, PDHG to termination
under the reference's _converged test with eps_abs = eps_rel = feas_cap =
1e-6), with the inputs resident in HBM.  Default arithmetic: FP32 iterates
and matrix, FP64 reductions, and an outer iterative refinement (FP64 X, Y)
that takes each objective to within ~1e-8 of the FP64 optimum, with rounds
cut short once their KKT ratio stops halving for 1024 iterations.  `value`
counts only the LPs that end OPTIMAL (the criterion met), per second of the
timed steps.

Beside the headline number the JSON line carries:
  plain_fp64    the same LPs one at a time through the single-LP engine in
                FP64 (inputs resident, a sample), and the ratio
  engine_default_stalls  the batch again with the engine's default stall
                lengths (4096 before a refined point meets the criterion):
                more LPs finish, in more time
  fp64_batch    the batch in FP64 on the same batched engine (one timed
                step): what batching alone buys, and the ratio to it
  roofline      algorithmic bytes of the step kernel (phase counts x the byte
                model of each phase) / its average launch time (HIP events on
                the launch stream), against 8 TB/s; `traffic` the HBM bytes per
                launch from the committed FETCH_SIZE / WRITE_SIZE passes when
                they were taken on this workload and launch structure
  cpu_baseline  the FP64 C restatement of the reference solver: the batch
                sample one LP per host thread (the headline's unit), the same
                on one thread, and the 500 x 2000 LP's wall time per solve
                (median of 3 after a warm-up) on the affinity's cores and on one
  spmv          the tile-pass SpMV of A and A^T of the 200k x 1M skewed LP,
                FP64 and FP32, in the given and the locality column order
                (HIP events, algorithmic GB/s against 8 TB/s)
  large_lp      the 200k x 1M skewed LP solved once in FP64 (rel KKT 1e-6)
                and once in FP32 + iterative refinement, with the engine's
                defaults for an LP of this size (locality column order, rounds
                at 1e-3): status, iterations, device time, objective error
                against the closed form, and the refinement record
  bnr           batched node relaxation, 128 nodes of a 500 x 2000 LP
                (optional, --bnr)
  certify       the certificate of a point of the 200k x 1M LP (optional,
                --certify)
  serve         a Poisson stream of multi-period blending LPs (10 req/s for
                10 s, T 24-36 hours) through the serving loop: FP64, 128 slots
                of 8 workgroups, admission by free HBM; p50/p99 latency,
                LPs/s, outcomes (--no-serve to skip)

`--only SECTION` runs one section by itself (spmv, large_lp, bnr, certify,
serve, or copy: a plain 256 MiB device copy whose counter readings calibrate
the FETCH_SIZE / WRITE_SIZE passes) and prints its JSON: the short runs the
committed rocprofv3 profiles were taken on.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--lps 4096]
                       [--precision 64|32] [--block 256|64] [--only SECTION]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sovereign-moe_amd"))
sys.path.insert(0, ROOT)

METRIC = ("LPs/s at rel KKT≤1e-6: batched FP32+refine vs plain FP64 GPU PDLP; p50/p99")
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lps", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--precision", type=int, default=32, choices=(64, 32),
                    help="32: FP32 iterates + FP64 reductions + iterative refinement (default)")
    ap.add_argument("--block", type=int, default=256, choices=(256, 64))
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--refine-stall", type=int, default=0,
                    help="FP32: iterations without a halving of a round's KKT ratio that end a "
                         "round polishing a point that met the criterion (0: the engine's 1024)")
    ap.add_argument("--refine-stall-open", type=int, default=1024,
                    help="FP32: the same for rounds before any refined point met the criterion "
                         "(0: the engine's 4096, which finishes more LPs in twice the time)")
    ap.add_argument("--robust-sample", type=int, default=1,
                    help="FP32: steps also run with the engine's default stall lengths, "
                         "reported beside the headline (0: skip)")
    ap.add_argument("--fp64-batch-steps", type=int, default=1,
                    help="FP32: timed steps of the same batch in FP64, reported beside the "
                         "headline (0: skip)")
    ap.add_argument("--plain-sample", type=int, default=48,
                    help="LPs solved one at a time for the FP64 single-LP rate (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="wall budget of the CPU baseline sample (0: skip)")
    ap.add_argument("--no-spmv", action="store_true",
                    help="skip the SpMV section (A and A^T of the 200k x 1M LP)")
    ap.add_argument("--no-large-lp", action="store_true",
                    help="skip the 200k x 1M LP solves (FP64, FP32 + refinement)")
    ap.add_argument("--no-lds-matrix", action="store_true",
                    help="keep every LP's matrix in memory (by default the run kernel holds "
                         "an LP's matrix in LDS for the PDHG loop when it fits)")
    ap.add_argument("--panels", action="store_true",
                    help="SpMV section: also A x as eight XCD column panels")
    ap.add_argument("--large-lp-given", action="store_true",
                    help="large LP: also solve in the given column order at the "
                         "small-LP refinement tolerance")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving section (a Poisson stream of blending LPs)")
    ap.add_argument("--serve-rate", type=float, default=10.0)
    ap.add_argument("--serve-duration", type=float, default=10.0)
    ap.add_argument("--serve-tmin", type=int, default=24)
    ap.add_argument("--serve-tmax", type=int, default=36)
    ap.add_argument("--serve-slots", type=int, default=128)
    ap.add_argument("--serve-team", type=int, default=8)
    ap.add_argument("--serve-precision", type=int, default=64, choices=(32, 64))
    ap.add_argument("--only", default="",
                    choices=("", "spmv", "large_lp", "bnr", "certify", "serve", "copy"),
                    help="run only this section and print its JSON")
    ap.add_argument("--bnr", action="store_true",
                    help="also time a batched node relaxation: 128 nodes of the 500 x 2000 LP, "
                         "60 PDHG iterations, and the oracle on a sample")
    ap.add_argument("--certify", action="store_true",
                    help="also time the certificate (bound, reduced costs, compensated row "
                         "activities) of a point of the 200k x 1M LP, and the C oracle on it")
    ap.add_argument("--tuning", default="auto",
                    help="autotuner table (JSON) choosing each LP's kernel configuration; "
                         "'auto': the packaged MI355X table for the precision if present; "
                         "'none': every LP on --block")
    ap.add_argument("--escalate", action="store_true",
                    help="FP32: re-solve the LPs the refinement cannot finish in FP64, "
                         "warm-started from their FP32 point, inside the timed step")
    return ap.parse_args()


def refinement_config(par, precision: int):
    """Refinement settings the engine ran with (None at FP64)."""
    rounds, eps = par.resolved_refine(precision)
    if rounds == 0:
        return None
    return {"rounds": rounds, "round_eps": eps, "stall": par.refine_stall or 1024,
            "stall_open": par.refine_stall_open or 4096, "target": par.refine_tol}


def fp64_batch_rate(problems, tol: float, steps: int) -> dict:
    """The same batch in FP64 on the same batched engine (256-thread
    workgroups, no refinement): the batching without the mixed precision,
    so the headline's gain splits into what batching and what FP32 +
    refinement buy.  One untimed step, then ``steps`` timed ones."""
    import torch
    from sovlp.lp.batch import BatchEngine
    from sovlp.lp.pdlp import PDLPParams
    par = PDLPParams(eps_abs=tol, eps_rel=tol, feas_cap=tol, device="gpu")
    eng = BatchEngine(problems, par, precision=64, block=256)
    try:
        eng.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.run()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        sols, lat = eng.solutions(keep_outputs=False)
    finally:
        eng.close()
    import numpy as np
    opt = sum(1 for s in sols if int(s.status) == 0)
    return {"value": opt * steps / wall, "unit": "LPs/s", "optimal": opt,
            "ms_per_step": wall / steps * 1e3, "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)), "max_iterations":
            int(max(s.iterations for s in sols))}


def robust_rate(problems, par, tuner, a) -> dict:
    """The same batch with the engine's default stall lengths (rounds before a
    refined point meets the criterion may run 4x longer): more LPs finish, in
    more time.  One untimed step, then ``a.robust_sample`` timed ones."""
    import dataclasses
    import torch
    from sovlp.lp.batch import BatchEngine
    p2 = dataclasses.replace(par, refine_stall_open=0)
    eng = BatchEngine(problems, p2, precision=a.precision, block=a.block, tuner=tuner)
    try:
        eng.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.robust_sample):
            eng.run()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        sols, _lat = eng.solutions(keep_outputs=False)
    finally:
        eng.close()
    opt = sum(1 for s in sols if int(s.status) == 0)
    return {"value": opt * a.robust_sample / wall, "unit": "LPs/s", "optimal": opt,
            "ms_per_step": wall / a.robust_sample * 1e3,
            "refinement": refinement_config(p2, a.precision)}


def bnr_rate(K: int = 128, iters: int = 60, reps: int = 10) -> dict:
    """Batched node relaxation: K nodes of the 500 x 2000 LP (random sub-boxes),
    `iters` PDHG iterations per call; device time per call (wall, synchronised),
    node-iterations per second, algorithmic bytes per iteration, and the oracle
    on 5 iterations of the same batch (one host thread, NumPy)."""
    import numpy as np
    import torch
    from sovlp.mip.bnr import BNRConfig, BNREngine
    from sovlp.workloads import small_lp
    p = small_lp(0).problem
    rng = np.random.default_rng(5)
    L = np.repeat(p.col_lb[:, None], K, 1).copy()
    U = np.repeat(p.col_ub[:, None], K, 1).copy()
    for q in range(1, K):
        js = rng.choice(p.n, 20, replace=False)
        U[js, q] = np.minimum(U[js, q], rng.uniform(0.0, 5.0, 20))
    eng = BNREngine(p, config=BNRConfig(batch=K, iters=iters, textbook_duals=True))
    for _ in range(2):          # the second call captures the launches as a graph
        eng.solve_batch(L, U)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        r = eng.solve_batch(L, U)
    torch.cuda.synchronize()
    s = (time.perf_counter() - t) / reps
    m, n, nnz = p.m, p.n, p.A.nnz
    # per iteration: two shared-A products (A once, K gathered values per
    # nonzero, K outputs per row), the primal (6 values per (j, k)), dual (5 per
    # (i, k)) and update (4 per entry) passes
    prod = lambda rows: 12 * nnz + 8 * (rows + 1) + 8 * nnz * K + 8 * rows * K  # noqa: E731
    per_iter = prod(m) + prod(n) + 8 * K * (6 * n + 5 * m + 4 * (n + m))
    out = {"lp": f"{m} x {n}, {nnz} nonzeros", "nodes": K, "iters": iters,
           "ms_per_call": s * 1e3, "node_iterations_per_s": K * iters / s,
           "algorithmic_bytes_per_iteration": per_iter,
           "achieved_gbs": per_iter * iters / s / 1e9,
           "frac_of_hbm_peak": per_iter * iters / s / 1e9 / HBM_PEAK_GBS,
           "bounds_finite": int(np.isfinite(r.bounds).sum())}
    try:
        from oracle import bnr_oracle
        t = time.perf_counter()
        bnr_oracle.solve_batch(p, L, U, iters=5, dual_sign=-1.0)
        cpu = (time.perf_counter() - t) / 5
        out["cpu_baseline"] = {"node_iterations_per_s": K / cpu, "cores": 1, "kind": "port",
                               "sample": f"5 iterations of the same {K}-node batch"}
    except ImportError as e:
        out["cpu_baseline"] = {"error": str(e)}
    return out


def certify_rate(reps: int = 20, k=None) -> dict:
    """One certificate of a point (x inside the box, y random) of the 200k x 1M
    skewed LP: device time per call (kernels, the 48-byte result copy and the
    synchronisation), its algorithmic bytes, and the C oracle on the same call
    (one host thread)."""
    import numpy as np
    import torch
    from sovlp.lp.certify import DeviceCertifier
    from sovlp.workloads import skewed_lp
    t0 = time.perf_counter()
    k = skewed_lp(seed=1) if k is None else k
    p = k.problem
    rng = np.random.default_rng(9)
    x = rng.uniform(0.0, 1.0, p.n) * np.minimum(p.col_ub, 10.0)
    y = rng.standard_normal(p.m)
    log(f"certificate: 200k x 1M LP built in {time.perf_counter() - t0:.1f} s")
    dc = DeviceCertifier(p)
    xs = torch.from_numpy(x).cuda()
    ys = torch.from_numpy(y).cuda()
    for _ in range(3):
        dc.certify(xs, ys)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        cert = dc.certify(xs, ys)
    us = (time.perf_counter() - t) / reps * 1e6
    m, n, nnz = p.m, p.n, p.A.nnz
    # rows: CSR (12 B per nonzero + 8 B pointers), x gathers, rl ru y, yy act
    # written; columns: CSC, yy gathers, c x lo hi, d written; terms: d lo hi
    alg = 20 * nnz + 8 * (m + 1) + 40 * m + 20 * nnz + 8 * (n + 1) + 40 * n + 24 * n
    out = {"lp": "200000 x 1000000, %d nonzeros" % nnz, "us_per_certificate": us,
           "algorithmic_bytes": alg, "achieved_gbs": alg / (us * 1e-6) / 1e9,
           "frac_of_hbm_peak": alg / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
           "bound": cert.bound, "perturbation": cert.perturbation}
    try:
        from oracle import oracle
        t = time.perf_counter()
        ref = oracle.certify(p, x, y)
        cpu = time.perf_counter() - t
        out["cpu_baseline"] = {"seconds_per_certificate": cpu, "cores": 1, "kind": "port",
                               "bound_rel_diff": abs(ref.bound - cert.bound) /
                               max(1.0, abs(ref.bound))}
    except (ImportError, OSError) as e:
        out["cpu_baseline"] = {"error": str(e)}
    return out


def traffic_key(lps: int, seed: int, precision: int, block: int, tuned: bool,
                refine: dict | None, lds_matrix: bool = True) -> str:
    """The workload a traffic record was measured on, as the record names it."""
    want = f"{lps} uneven LPs, seeds {seed}..{seed + lps - 1}, " \
           f"{'FP32 + refinement' if precision == 32 else 'FP64'}, block {block}" \
           f"{', autotuned' if tuned else ''}"
    if refine:
        want += f", {refine['rounds']} rounds at {refine['round_eps']:g}, stall " \
                f"{refine['stall_open']}/{refine['stall']}"
    if lds_matrix:
        want += ", matrices or step vectors in LDS where they fit"
    return want


SPMV_KEYS = ("op", "kernel", "precision", "rows", "cols", "nnz", "grid", "tiles")


def spmv_traffic() -> dict:
    """Counted traffic per launch of the SpMV section's products from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes over `bench.py --only spmv`,
    keyed by the product's shape and launch (assembled by ``python -m
    sovlp.lp.traffic``)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_spmv_tiled_200k_x_1M.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return {}
    return {tuple(e[k] for k in SPMV_KEYS): float(e["traffic_bytes_per_launch"])
            for e in d.get("products", [])}


def measured_traffic(kernel: str, lps: int, seed: int, precision: int, block: int, tuned: bool,
                     refine: dict | None, replays: int, lds_matrix: bool = True):
    """HBM bytes per launch of ``kernel`` on this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a plain copy's reading;
    see profiles/, assembled by ``python -m sovlp.lp.traffic``), or None when
    they were taken on another configuration or with another number of graph
    replays per step."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_pdlp_run_kernel_fp32_4096_lps.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    want = traffic_key(lps, seed, precision, block, tuned, refine, lds_matrix)
    if d.get("kernel") != kernel or d.get("workload") != want \
            or d.get("replays_per_step") != replays:
        return None
    return float(d["traffic_bytes_per_launch"])


def slowest_lps(eng, sols, lat, problems, k: int = 8) -> list:
    """The LPs that set the step, latest finish first: index, size, kernel
    block, iterations, refinement rounds, status, start, latency and finish on
    the device clock (ms from the step's first LP start)."""
    import ctypes
    import numpy as np
    from sovlp import _lib
    hz = ctypes.c_double()
    _lib.call("sovlp_clock_rate", 0, ctypes.byref(hz))
    infos = eng.infos()
    t0 = min(int(infos[i].start_tick) for i in range(eng.count))
    start = np.array([(int(infos[i].start_tick) - t0) / hz.value for i in range(eng.count)])
    end = start + np.asarray(lat, dtype=np.float64)
    out = []
    for i in np.argsort(end)[::-1][:k]:
        i = int(i)
        p = problems[i]
        out.append({"lp": i, "m": p.m, "n": p.n, "nnz": int(p.A.nnz),
                    "block": int(eng.db.blocks[i]), "iterations": int(sols[i].iterations),
                    "rounds": len(sols[i].info.refinement), "status": sols[i].status.name,
                    "start_ms": start[i] * 1e3, "latency_ms": float(lat[i]) * 1e3,
                    "end_ms": end[i] * 1e3})
    return out


def phase_profile(counts, ticks, lat) -> dict:
    """Device time per phase of the last step, from the run kernel's own clock:
    for all LPs and for the 1 % slowest (the step's tail), each phase's calls
    and mean microseconds per call."""
    import ctypes
    import numpy as np
    from sovlp import _lib
    L = _lib.lib()
    hz = ctypes.c_double()
    _lib.call("sovlp_clock_rate", 0, ctypes.byref(hz))
    names = [L.sovlp_phase_name(k).decode() for k in range(counts.shape[1])]
    tail = np.argsort(lat)[-max(1, len(lat) // 100):]
    out = {}
    for label, rows in (("all", slice(None)), ("slowest_1pct", tail)):
        c = counts[rows].sum(axis=0)
        t = ticks[rows].sum(axis=0) / hz.value * 1e6
        out[label] = {names[k]: {"calls": int(c[k]), "us_per_call": float(t[k] / c[k]),
                                 "share": float(t[k] / t.sum())}
                      for k in range(len(names)) if c[k] > 0 and names[k]
                      and not names[k].endswith("_NEXT")}
    return out


def plain_fp64_rate(problems, tol, sample):
    """LPs/s of the single-LP FP64 engine, one LP at a time, inputs resident."""
    import torch
    from sovlp.lp.pdlp import DeviceLP, Engine, PDLPParams
    par = PDLPParams(eps_abs=tol, eps_rel=tol, feas_cap=tol, device="gpu", precision=64)
    total, iters, solved = 0.0, 0, 0
    for p in problems[:sample]:
        dlp = DeviceLP(p, 64)
        eng = Engine(dlp, par, history_cap=1)
        try:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            while True:
                eng.advance(par.chunk_iters, par.time_limit)
                inf = eng.info()
                if inf.stage == 3:
                    break
            total += time.perf_counter() - t0
            iters += int(inf.iterations)
            solved += int(inf.status == 0)
        finally:
            eng.close()
    return dict(lps=sample, seconds=total, value=solved / total, unit="LPs/s",
                optimal=solved, mean_iterations=iters / max(sample, 1))


def cpu_baseline(problems, tol, budget_s):
    """The FP64 C restatement of the reference solver (the oracle).

    * ``value``: LPs/s over the batch, one LP per host thread, as many LPs as
      finish within about ``budget_s``, on the affinity's cores (at most 16,
      the host share of one GPU here; ``cores`` says how many ran);
    * ``one_thread``: the same on one thread, a quarter of the budget;
    * ``configs0``: the 500 x 2000 LP at rel KKT 1e-8, wall time per solve,
      median of 3 after one warm-up, with its OpenMP products on the affinity's
      cores and on one thread."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    from oracle import oracle
    from sovlp.workloads import small_lp
    oracle.lib()
    affinity = len(os.sched_getaffinity(0))
    cap = 16
    cores = max(1, min(affinity, cap))

    def rate(threads, budget):
        done = []
        t0 = time.perf_counter()
        deadline = t0 + budget

        def work(i):
            if time.perf_counter() > deadline:
                return None
            r = oracle.solve(problems[i], eps_abs=tol, eps_rel=tol, feas_cap=tol, threads=1,
                             hist_cap=1)
            return r.status

        with ThreadPoolExecutor(threads) as ex:
            for st in ex.map(work, range(len(problems))):
                if st is not None:
                    done.append(st)
        return done, time.perf_counter() - t0

    done, el = rate(cores, budget_s)
    one, el1 = rate(1, budget_s / 4)
    ko = small_lp(0)

    def per_solve(threads):
        oracle.solve(ko.problem, threads=threads, hist_cap=1)
        ts, it = [], 0
        for _ in range(3):
            t = time.perf_counter()
            r = oracle.solve(ko.problem, threads=threads, hist_cap=1)
            ts.append(time.perf_counter() - t)
            it = r.iterations
        return float(np.median(ts)), it, abs(r.objective - ko.f_star) / max(1.0, abs(ko.f_star))

    c0_all, it0, err0 = per_solve(cores)
    c0_one, _, _ = per_solve(1)
    return dict(value=len(done) / el, unit="LPs/s", cores=cores, kind="port",
                sample=f"first {len(done)} LPs of the batch, FP64 C restatement of the "
                       f"reference PDLP, one LP per thread on {cores} threads, {el:.1f} s wall",
                optimal=int(sum(1 for s in done if s == 0)),
                affinity_cores=affinity, thread_cap=cap,
                one_thread={"value": len(one) / el1, "unit": "LPs/s", "lps": len(one),
                            "seconds": el1},
                configs0={"lp": "500 x 2000 (250 eq, 250 <=), 5 nnz per column, seed 0",
                          "tolerance": 1e-8, "iterations": it0, "rel_obj_err": err0,
                          "seconds_per_solve_median_of_3": c0_all, "threads": cores,
                          "seconds_per_solve_one_thread": c0_one})


def gpu_configs0(reps: int = 3) -> dict:
    """The 500 x 2000 LP on the GPU's single-LP engine (FP64, one workgroup, the
    reference's arithmetic) at rel KKT 1e-8: wall time per solve with the
    inputs resident, median of ``reps`` after a warm-up."""
    import numpy as np
    import torch
    from sovlp.lp.pdlp import DeviceLP, Engine, PDLPParams
    from sovlp.workloads import small_lp
    ko = small_lp(0)
    par = PDLPParams(device="gpu", precision=64, team="block")
    dlp = DeviceLP(ko.problem, 64)
    ts, it = [], 0
    for k in range(reps + 1):
        eng = Engine(dlp, par, history_cap=1)
        try:
            torch.cuda.synchronize()
            t = time.perf_counter()
            while eng.info().stage != 3:
                eng.advance(par.chunk_iters, par.time_limit)
            dt = time.perf_counter() - t
            inf = eng.info()
        finally:
            eng.close()
        if k:
            ts.append(dt)
        it = int(inf.iterations)
    return {"seconds_per_solve_median": float(np.median(ts)), "iterations": it,
            "rel_obj_err": abs(inf.objective - ko.f_star) / max(1.0, abs(ko.f_star))}


def large_lp_section(ko, given: bool = False) -> dict:
    """The 200k x 1M skewed LP solved from scratch: FP64 (the reference's
    arithmetic) and FP32 + iterative refinement, rel KKT 1e-6 (feas_cap 1e-6),
    with the engine's defaults for an LP of this size (columns in locality
    order, refinement rounds at 1e-3), and with ``given`` also in the caller's
    column order at the small-LP round tolerance; time on the device clock and
    the wall, iterations, objective error against the closed form, and for
    FP32 the refinement record."""
    from sovlp.lp.pdlp import PDLPParams, solve_pdlp
    f = ko.f_star
    rel = lambda v: abs(v - f) / max(1.0, abs(f))  # noqa: E731
    out = {"lp": f"{ko.problem.m} x {ko.problem.n}, {ko.problem.A.nnz} nonzeros, Zipf column "
                 f"degrees (s=1.3, 2..5000)", "tolerance": 1e-6}
    orders = ["auto"] + (["given"] if given else [])
    for order in orders:
        for prec in (64, 32):
            extra = {} if order == "auto" else {"refine_eps": 3e-4}
            par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6, device="gpu",
                             precision=prec, time_limit=240.0, column_order=order, **extra)
            t = time.perf_counter()
            sol = solve_pdlp(ko.problem, par)
            wall = time.perf_counter() - t
            nnz = int(ko.problem.A.nnz)
            r = {"status": sol.status.name, "iterations": int(sol.iterations),
                 "restarts": int(sol.info.restarts), "device_s": sol.info.device_seconds,
                 "wall_s": wall, "us_per_iteration": sol.info.device_seconds /
                 max(1, sol.iterations) * 1e6, "workgroups": sol.info.workgroups,
                 "column_order": par.resolved_column_order(nnz),
                 "rel_obj_err": rel(sol.objective)}
            if prec == 32:
                r["round_eps"] = par.resolved_refine(32, nnz)[1]
            if sol.info.refinement:
                r["refinement"] = [{"iterations": int(q[0]), "primal_res": q[1], "dual_res": q[2],
                                    "gap": q[3], "rel_obj_err": rel(q[4])}
                                   for q in sol.info.refinement]
            key = ("fp64" if prec == 64 else "fp32_refined") + \
                ("" if order == "auto" else "_given")
            out[key] = r
            log(f"large LP {key}: {r}")
    return out


def serve_section(a) -> dict:
    """A Poisson stream of blending LPs (K 20-60, P 5-20, T in the given range,
    30 % warm-started 1 % resolves) through the serving loop, admission by
    free HBM: p50/p99 latency, LPs/s, outcomes."""
    import numpy as np
    import torch
    from sovlp.lp.pdlp import PDLPParams
    from sovlp.lp.serve import (DeviceServer, HbmAdmission, ServeLoop, hbm_budget,
                                poisson_blending_requests, summarize)
    t0 = time.perf_counter()
    reqs = poisson_blending_requests(a.serve_rate, a.serve_duration, 0, 0.3, a.serve_tmin,
                                     a.serve_tmax)
    nnz = np.array([r.problem.A.nnz for r in reqs])
    log(f"serve: {len(reqs)} requests built in {time.perf_counter() - t0:.1f} s, nnz mean "
        f"{nnz.mean():.0f} max {nnz.max()}")
    tol = a.tol
    mk = lambda prec: PDLPParams(eps_abs=tol, eps_rel=tol, feas_cap=tol, device="gpu",  # noqa
                                 precision=prec, time_limit=60.0)
    srv = DeviceServer(a.serve_slots, mk(a.serve_precision), precision=a.serve_precision,
                       phases_per_launch=32, team=a.serve_team)
    fb = DeviceServer(32, mk(64), precision=64, phases_per_launch=32, team=a.serve_team) \
        if a.serve_precision == 32 else None
    adm = HbmAdmission(hbm_budget(0.8))
    try:
        ServeLoop(srv, HbmAdmission(adm.budget)).run(reqs[:1])     # warm-up
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out = ServeLoop(srv, adm, fallback=fb,
                        progress=lambda d: log(f"serve: {d}")).run(reqs)
        wall = time.perf_counter() - t1
    finally:
        srv.close()
        if fb is not None:
            fb.close()
    st = summarize(out)
    st.update(dict(workload=f"Poisson {a.serve_rate:g} req/s for {a.serve_duration:g} s, "
                            f"blending K 20-60, P 5-20, T {a.serve_tmin}-{a.serve_tmax}, "
                            f"30% warm-started resolves",
                   wall_s=wall, nnz_mean=float(nnz.mean()), nnz_max=int(nnz.max()),
                   hbm_budget_bytes=adm.budget, hbm_peak_bytes=adm.peak, slots=a.serve_slots,
                   team=a.serve_team, precision=a.serve_precision, tolerance=tol))
    return st


def copy_calibration(mib: int = 256, reps: int = 3) -> dict:
    """A plain device-to-device copy of a known size: rocprofv3's FETCH_SIZE and
    WRITE_SIZE read on it give the factor between what the counters report and
    the bytes moved, which then scales the solver kernels' readings."""
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"copy_bytes_read": n, "copy_bytes_written": n, "copies": reps}


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every LP of the batch lives on one device and "
              "sovlp has no multi-GPU path (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    from sovlp.lp.batch import BatchEngine
    from sovlp.lp.pdlp import PDLPParams
    from sovlp.lp.traffic import bytes_of_counts
    from sovlp.workloads import skewed_lp, uneven_batch

    big = {}

    def large():
        """The 200k x 1M skewed LP, built once."""
        if "ko" not in big:
            t = time.perf_counter()
            big["ko"] = skewed_lp(1)
            log(f"200k x 1M LP built in {time.perf_counter() - t:.1f} s")
        return big["ko"]

    def spmv_section():
        from sovlp.spmv_bench import skewed_roofline
        res, _ = skewed_roofline(A=large().problem.A, panels=a.panels)
        counted = spmv_traffic()
        for r in res:
            t = counted.get(tuple(r[k] for k in SPMV_KEYS))
            if t is not None:
                # the committed counter passes' bytes per launch (below L2:
                # Infinity Cache hits included) over this run's time
                r["traffic_bytes"] = t
                r["traffic_gbs"] = t / (r["ms"] * 1e-3) / 1e9
                r["traffic_frac"] = r["traffic_gbs"] / HBM_PEAK_GBS
            log(f"spmv {r['op']} {r['kernel']} FP{r['precision']}: {r['ms'] * 1e3:.1f} us, "
                f"{r['gbs']:.0f} GB/s ({r['frac']:.1%})"
                + (f", counted {r['traffic_gbs']:.0f} GB/s" if t is not None else ""))
        return {"lp": "200000 x 1000000 skewed LP", "peak_gbs": HBM_PEAK_GBS,
                "bytes_model": "nnz (s_val + 4) + (rows + 1) 4 + cols s_val + rows s_val",
                "results": res}

    if a.only:
        sec = {"spmv": spmv_section, "large_lp": lambda: large_lp_section(
            large(), a.large_lp_given), "bnr": bnr_rate, "certify": lambda: certify_rate(k=large()),
               "serve": lambda: serve_section(a), "copy": copy_calibration}[a.only]()
        print(json.dumps({a.only: sec}), flush=True)
        return

    t = time.perf_counter()
    kos = uneven_batch(a.lps, base_seed=a.seed)
    problems = [k.problem for k in kos]
    log(f"generated {a.lps} LPs in {time.perf_counter() - t:.1f} s "
        f"(sum nnz {sum(p.A.nnz for p in problems)})")
    par = PDLPParams(eps_abs=a.tol, eps_rel=a.tol, feas_cap=a.tol, device="gpu",
                     refine_stall=a.refine_stall, refine_stall_open=a.refine_stall_open,
                     lds_matrix=not a.no_lds_matrix)
    tuner = None
    if a.tuning != "none":
        from sovlp.lp.autotune import Autotuner, packaged_table
        path = packaged_table(a.precision) if a.tuning == "auto" else a.tuning
        if path and os.path.exists(path):
            tuner = Autotuner.load(path)
            if tuner.precision != a.precision:
                tuner = None
            else:
                log(f"kernel configurations from {os.path.relpath(path, ROOT)}: {tuner.table}")
    t = time.perf_counter()
    eng = BatchEngine(problems, par, precision=a.precision, block=a.block, tuner=tuner)
    log(f"batch laid out on the device in {time.perf_counter() - t:.1f} s")
    fb = None
    if a.precision == 32 and a.escalate:
        from sovlp.lp.serve import DeviceServer
        fb = DeviceServer(64, PDLPParams(eps_abs=a.tol, eps_rel=a.tol, feas_cap=a.tol,
                                         device="gpu", precision=64), precision=64)

    def escalate():
        """LPs the FP32 batch ended NUMERICAL or at the iteration limit: re-solved
        on the FP64 server, warm-started from their FP32 point."""
        if fb is None:
            return {}
        from sovlp.lp.serve import WarmStart, resolve_warm
        infos = eng.infos()
        bad = [i for i in range(eng.count) if int(infos[i].status) in (4, 8)]
        if not bad:
            return {}
        warm = [WarmStart(*eng.outputs_of(i, infos)) for i in bad]
        return dict(zip(bad, resolve_warm(fb, [problems[i] for i in bad], warm)))

    for w in range(a.warmup):
        g = eng.run()
        esc = escalate()
        log(f"warmup {w}: {g} graph replays, {len(esc)} LPs escalated to FP64")
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    graphs, batch_ms, esc_s = [], 0.0, 0.0
    for k in range(a.steps):
        ev0.record(stream)
        graphs.append(eng.run())
        ev1.record(stream)
        ev1.synchronize()
        t1 = time.perf_counter()
        esc = escalate()
        esc_s += time.perf_counter() - t1
        batch_ms += ev0.elapsed_time(ev1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = batch_ms
    log(f"timed {a.steps} steps: {wall:.3f} s wall, {dev_ms:.1f} ms of batch kernels on the "
        f"stream, graphs {graphs}, FP64 escalation {esc_s:.3f} s for {len(esc)} LPs per step")

    sols, lat = eng.solutions(keep_outputs=True)
    for i, s in esc.items():
        lat[i] += s.time
        s.info.refinement = sols[i].info.refinement
        sols[i] = s
    opt = sum(1 for s in sols if int(s.status) == 0)
    # objective error of every LP reported OPTIMAL, against its closed form
    rel = [abs(s.objective - k.f_star) / max(1.0, abs(k.f_star))
           for s, k in zip(sols, kos) if int(s.status) == 0]
    counts = eng.phase_counts()
    phase_prof = phase_profile(counts, eng.phase_ticks(), lat)
    slowest = slowest_lps(eng, sols, lat, problems)
    step_bytes = bytes_of_counts(counts, eng.sizes(), a.precision)
    launches = sum(graphs)            # one run-kernel launch per graph replay
    launch_ms = dev_ms / launches
    achieved = step_bytes / graphs[-1] / (launch_ms * 1e-3) / 1e9
    from sovlp.lp.batch import batch_geometry
    blocks = sorted(set(int(b) for b in eng.db.blocks))
    kernel = " + ".join(f"pdlp_run_kernel<{'double' if a.precision == 64 else 'float'}, {b}, "
                        f"{batch_geometry(a.precision, b)[1]}>" for b in blocks)
    modes = eng.lds_modes()
    n_resident, n_vec = int((modes == 1).sum()), int((modes == 2).sum())
    eng.close()
    if fb is not None:
        fb.close()

    out = {
        "metric": METRIC,
        "value": opt * a.steps / wall,     # LPs that met the criterion, per second
        "unit": "LPs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == 64 else "f32 iterates, f64 reductions and refinement",
        "data": "synthetic",
        "config": {"workload": f"{a.lps} uneven LPs with known optima, m log-uniform[20,800] "
                               f"(60% eq), n=U[2,4]m, 3-8 nnz/col, bounds [0,U(5,50)], "
                               f"seeds {a.seed}..{a.seed + a.lps - 1}",
                   "lps": a.lps, "tolerance": a.tol, "precision": a.precision,
                   "block": a.block, "parallelism": "one workgroup per LP",
                   "kernel_configs": "autotuned" if tuner is not None else f"block {a.block}",
                   "lds_matrix": not a.no_lds_matrix,
                   "refinement": refinement_config(par, a.precision)},
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "optimal": opt,
        "lds_resident_lps": n_resident,        # matrix and step vectors in LDS
        "lds_vectors_only_lps": n_vec,         # only the step vectors in LDS
        "lps_per_s_all_statuses": a.lps * a.steps / wall,
        "max_rel_obj_err_vs_closed_form": float(max(rel)) if rel else None,   # optimal LPs
        "rel_obj_err_p99": float(np.percentile(rel, 99)) if rel else None,
        "mean_iterations": float(np.mean([s.iterations for s in sols])),
        "mean_refinement_rounds": float(np.mean([len(s.info.refinement) for s in sols])),
        "escalated_to_fp64": len(esc),
        "phase_profile": phase_prof,
        "slowest_lps": slowest,
        "escalation_s_per_step": esc_s / a.steps,
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kernel, a.lps, a.seed, a.precision, a.block,
                                                 tuner is not None,
                                                 refinement_config(par, a.precision),
                                                 int(graphs[-1]), not a.no_lds_matrix),
                     "traffic_key": traffic_key(a.lps, a.seed, a.precision, a.block,
                                                tuner is not None,
                                                refinement_config(par, a.precision),
                                                not a.no_lds_matrix),
                     "algorithmic_bytes_per_launch": step_bytes / graphs[-1],
                     "avg_launch_us": launch_ms * 1e3, "launches": launches},
    }
    if a.precision == 32 and a.robust_sample > 0 and a.refine_stall_open != 0:
        rb = robust_rate(problems, par, tuner, a)
        log(f"default stall lengths: {rb}")
        out["engine_default_stalls"] = rb
    if a.plain_sample > 0:
        t = time.perf_counter()
        pl = plain_fp64_rate(problems, a.tol, min(a.plain_sample, a.lps))
        log(f"plain fp64 sample: {pl} ({time.perf_counter() - t:.1f} s)")
        out["plain_fp64"] = pl
        out["speedup_vs_plain_fp64"] = out["value"] / pl["value"]
    if a.precision == 32 and a.fp64_batch_steps > 0:
        t = time.perf_counter()
        fb64 = fp64_batch_rate(problems, a.tol, a.fp64_batch_steps)
        log(f"fp64 batch: {fb64} ({time.perf_counter() - t:.1f} s)")
        out["fp64_batch"] = fb64
        out["speedup_vs_fp64_batch"] = out["value"] / fb64["value"]
    if a.cpu_seconds > 0:
        cb = cpu_baseline(problems, a.tol, a.cpu_seconds)
        log(f"cpu baseline: {cb}")
        out["cpu_baseline"] = cb
    out["gpu_configs0"] = gpu_configs0()
    log(f"500 x 2000 LP on the GPU: {out['gpu_configs0']}")
    if not a.no_spmv:
        out["spmv"] = spmv_section()
    if not a.no_large_lp:
        out["large_lp"] = large_lp_section(large(), a.large_lp_given)
    if not a.no_serve:
        out["serve"] = serve_section(a)
        log(f"serve: {out['serve']}")
    if a.bnr:
        out["bnr"] = bnr_rate()
        log(f"bnr: {out['bnr']}")
    if a.certify:
        out["certify"] = certify_rate(k=large())
        log(f"certificate: {out['certify']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
