"""Serving measurement: a Poisson stream of multi-period blending LPs.

Requests arrive at ``rate`` per second for ``duration`` seconds (arrival times
drawn up front from a seeded exponential clock). A fraction ``resolve`` of them
re-solve an earlier model with ``b`` and ``c`` perturbed by 1 % and are
warm-started from that model's last solution; the rest are new models with
``K ~ U{20..60}`` components, ``P ~ U{5..20}`` products and ``T ~ U{tmin..tmax}``
hours (``sovlp.workloads.blending_lp``). Every model is built before the clock
starts, so the loop measures the device and the admission path only.

The primary device is an FP32 server (iterates and matrix FP32, FP64
reductions and refinement); an LP it cannot finish is re-solved on an FP64
server warm-started from its FP32 point. Admission is by free HBM.

    python -m sovlp.lp.serve_bench [--rate 10 --duration 10 --tmin 24 --tmax 36 --precision 32]
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np


def make_stream(rate: float, duration: float, seed: int, resolve: float, tmin: int, tmax: int,
                kmax: int = 60, pmax: int = 20):
    from ..workloads import blending_lp, perturbed
    from .serve import Request
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--rate", type=float, default=10.0)
    ap.add_argument("--duration", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--resolve", type=float, default=0.3)
    ap.add_argument("--tmin", type=int, default=24)
    ap.add_argument("--tmax", type=int, default=36)
    ap.add_argument("--slots", type=int, default=128)
    ap.add_argument("--team", type=int, default=8,
                    help="workgroups per served LP and phase (1: one workgroup runs all phases)")
    ap.add_argument("--fallback-slots", type=int, default=32)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--hbm-fraction", type=float, default=0.8)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64),
                    help="primary server's precision (FP64 then has no fallback)")
    ap.add_argument("--time-limit", type=float, default=60.0,
                    help="device seconds per LP on each server")
    a = ap.parse_args()
    import torch
    from .pdlp import PDLPParams
    from .serve import DeviceServer, HbmAdmission, ServeLoop, hbm_budget, summarize
    t0 = time.perf_counter()
    reqs = make_stream(a.rate, a.duration, a.seed, a.resolve, a.tmin, a.tmax)
    nnz = np.array([r.problem.A.nnz for r in reqs])
    gen_s = time.perf_counter() - t0
    print(f"[serve] {len(reqs)} requests built in {gen_s:.1f} s; nnz mean {nnz.mean():.0f} "
          f"max {nnz.max()}", flush=True)
    par32 = PDLPParams(eps_abs=a.tol, eps_rel=a.tol, feas_cap=a.tol, device="gpu", precision=32,
                       time_limit=a.time_limit)
    par64 = PDLPParams(eps_abs=a.tol, eps_rel=a.tol, feas_cap=a.tol, device="gpu", precision=64,
                       time_limit=a.time_limit)
    if a.precision == 32:
        srv = DeviceServer(a.slots, par32, precision=32, phases_per_launch=32, team=a.team)
        fb = DeviceServer(a.fallback_slots, par64, precision=64, phases_per_launch=32,
                          team=a.team)
    else:
        srv = DeviceServer(a.slots, par64, precision=64, phases_per_launch=32, team=a.team)
        fb = None
    adm = HbmAdmission(hbm_budget(a.hbm_fraction))
    try:
        # warm-up: one small request through both servers (module load, graphs)
        ServeLoop(srv, HbmAdmission(adm.budget)).run(reqs[:1])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out = ServeLoop(srv, adm, fallback=fb,
                        progress=lambda d: print(f"[serve] {d}", flush=True)).run(reqs)
        wall = time.perf_counter() - t1
    finally:
        srv.close()
        if fb is not None:
            fb.close()
    st = summarize(out)
    st.update(dict(workload=f"Poisson {a.rate:g} req/s for {a.duration:g} s, blending K 20-60, "
                            f"P 5-20, T {a.tmin}-{a.tmax}, {a.resolve:.0%} warm-started resolves",
                   wall_s=wall, nnz_mean=float(nnz.mean()), nnz_max=int(nnz.max()),
                   hbm_budget_bytes=adm.budget, hbm_peak_bytes=adm.peak, slots=a.slots,
                   fallback_slots=a.fallback_slots if fb is not None else 0, team=a.team,
                   tolerance=a.tol,
                   precision=a.precision, time_limit_s=a.time_limit))
    print(json.dumps(st), flush=True)


if __name__ == "__main__":
    main()
