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
  roofline      algorithmic bytes of the step kernel (phase counts x the byte
                model of each phase) / its average launch time (HIP events on
                the launch stream), against 8 TB/s; `traffic` the HBM bytes per
                launch from the committed FETCH_SIZE / WRITE_SIZE passes when
                they were taken on this workload and launch structure
  cpu_baseline  the FP64 C restatement of the reference solver on a bounded
                sample of the batch, one LP per host thread
  spmv          the tile-pass SpMV of A and A^T of the 200k x 1M skewed LP, in
                the given and the locality column order (optional, --spmv)
  bnr           batched node relaxation, 128 nodes of a 500 x 2000 LP
                (optional, --bnr)
  certify       the certificate of a point of the 200k x 1M LP (optional,
                --certify)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--lps 4096]
                       [--precision 64|32] [--block 256|64]
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
    ap.add_argument("--plain-sample", type=int, default=48,
                    help="LPs solved one at a time for the FP64 single-LP rate (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="wall budget of the CPU baseline sample (0: skip)")
    ap.add_argument("--spmv", action="store_true", help="also time the SpMV of the 200k x 1M LP")
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


def certify_rate(reps: int = 20) -> dict:
    """One certificate of a point (x inside the box, y random) of the 200k x 1M
    skewed LP: device time per call (kernels, the 48-byte result copy and the
    synchronisation), its algorithmic bytes, and the C oracle on the same call
    (one host thread)."""
    import numpy as np
    import torch
    from sovlp.lp.certify import DeviceCertifier
    from sovlp.workloads import skewed_lp
    t0 = time.perf_counter()
    k = skewed_lp(seed=1)
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


def measured_traffic(kernel: str, lps: int, seed: int, precision: int, block: int, tuned: bool,
                     refine: dict | None, replays: int):
    """HBM bytes per launch of ``kernel`` on this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a plain copy's reading;
    see profiles/), or None when they were taken on another configuration or
    with another number of graph replays per step."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_pdlp_run_kernel_fp32_4096_lps.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    want = f"{lps} uneven LPs, seeds {seed}..{seed + lps - 1}, " \
           f"{'FP32 + refinement' if precision == 32 else 'FP64'}, block {block}" \
           f"{', autotuned' if tuned else ''}"
    if refine:
        want += f", {refine['rounds']} rounds at {refine['round_eps']:g}, stall " \
                f"{refine['stall_open']}/{refine['stall']}"
    if d.get("kernel") != kernel or d.get("workload") != want \
            or d.get("replays_per_step") != replays:
        return None
    return float(d["traffic_bytes_per_launch"])


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
    """The FP64 C restatement of the reference solver, one LP per host thread,
    over as many LPs of the batch as finish within about ``budget_s``."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle
    oracle.lib()
    cores = len(os.sched_getaffinity(0))
    cores = max(1, min(cores, 16))
    done = []
    t0 = time.perf_counter()
    deadline = t0 + budget_s

    def work(i):
        if time.perf_counter() > deadline:
            return None
        r = oracle.solve(problems[i], eps_abs=tol, eps_rel=tol, feas_cap=tol, threads=1,
                         hist_cap=1)
        return r.status

    with ThreadPoolExecutor(cores) as ex:
        for st in ex.map(work, range(len(problems))):
            if st is not None:
                done.append(st)
    el = time.perf_counter() - t0
    return dict(value=len(done) / el, unit="LPs/s", cores=cores, kind="port",
                sample=f"first {len(done)} LPs of the batch, FP64 C restatement of the "
                       f"reference PDLP, one LP per thread, {el:.1f} s wall",
                optimal=int(sum(1 for s in done if s == 0)))


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
    from sovlp.workloads import uneven_batch

    t = time.perf_counter()
    kos = uneven_batch(a.lps, base_seed=a.seed)
    problems = [k.problem for k in kos]
    log(f"generated {a.lps} LPs in {time.perf_counter() - t:.1f} s "
        f"(sum nnz {sum(p.A.nnz for p in problems)})")
    par = PDLPParams(eps_abs=a.tol, eps_rel=a.tol, feas_cap=a.tol, device="gpu",
                     refine_stall=a.refine_stall, refine_stall_open=a.refine_stall_open)
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
    step_bytes = bytes_of_counts(counts, eng.sizes(), a.precision)
    launches = sum(graphs)            # one run-kernel launch per graph replay
    launch_ms = dev_ms / launches
    achieved = step_bytes / graphs[-1] / (launch_ms * 1e-3) / 1e9
    from sovlp.lp.batch import batch_geometry
    blocks = sorted(set(int(b) for b in eng.db.blocks))
    kernel = " + ".join(f"pdlp_run_kernel<{'double' if a.precision == 64 else 'float'}, {b}, "
                        f"{batch_geometry(a.precision, b)[1]}>" for b in blocks)
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
                   "refinement": refinement_config(par, a.precision)},
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "optimal": opt,
        "lps_per_s_all_statuses": a.lps * a.steps / wall,
        "max_rel_obj_err_vs_closed_form": float(max(rel)) if rel else None,   # optimal LPs
        "rel_obj_err_p99": float(np.percentile(rel, 99)) if rel else None,
        "mean_iterations": float(np.mean([s.iterations for s in sols])),
        "mean_refinement_rounds": float(np.mean([len(s.info.refinement) for s in sols])),
        "escalated_to_fp64": len(esc),
        "escalation_s_per_step": esc_s / a.steps,
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kernel, a.lps, a.seed, a.precision, a.block,
                                                 tuner is not None,
                                                 refinement_config(par, a.precision),
                                                 int(graphs[-1])),
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
    if a.cpu_seconds > 0:
        cb = cpu_baseline(problems, a.tol, a.cpu_seconds)
        log(f"cpu baseline: {cb}")
        out["cpu_baseline"] = cb
    if a.spmv:
        from sovlp.spmv_bench import skewed_roofline
        res, gen = skewed_roofline()
        out["spmv"] = res
    if a.bnr:
        out["bnr"] = bnr_rate()
        log(f"bnr: {out['bnr']}")
    if a.certify:
        out["certify"] = certify_rate()
        log(f"certificate: {out['certify']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
