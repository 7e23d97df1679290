"""One large LP end to end: the 200k x 1M skewed LP solved by the single-LP engine.

* FP64 iterates (the reference's arithmetic), relative KKT 1e-6;
* FP32 matrix and iterates with FP64 reductions and iterative refinement, with
  the primal residual, duality gap and objective error of every refinement round.

The LP has a closed-form optimum (``sovlp.workloads.known_optimum_lp``), so each
result is reported with its relative objective error. With more than 200k
nonzeros the engine runs every phase over W workgroups (``team="grid"``).

    python -m sovlp.large_run [--m 200000 --n 1000000 --nnz 8000000] [--precision 64 32]
                              [--column-order given locality]
"""

from __future__ import annotations

import argparse
import json
import time


def run(m: int, n: int, nnz: int, precision: int, tol: float, time_limit: float, seed: int = 1,
        column_order: str = "given"):
    from .lp.pdlp import PDLPParams, solve_pdlp
    from .workloads import skewed_lp
    t = time.perf_counter()
    ko = skewed_lp(seed, m, n, nnz)
    gen = time.perf_counter() - t
    p = ko.problem
    par = PDLPParams(eps_abs=tol, eps_rel=tol, feas_cap=tol, device="gpu", precision=precision,
                     time_limit=time_limit, column_order=column_order)
    s = solve_pdlp(p, par)
    f = ko.f_star
    out = dict(m=p.m, n=p.n, nnz=int(p.A.nnz), precision=precision, tol=tol,
               column_order=column_order,
               status=s.status.name, iterations=int(s.iterations), seconds=s.time,
               device_seconds=s.info.device_seconds, workgroups=s.info.workgroups,
               restarts=int(s.info.restarts), objective=s.objective, f_star=f,
               rel_obj_err=abs(s.objective - f) / max(1.0, abs(f)), generate_s=gen)
    if s.info.refinement:
        out["refinement"] = [dict(iterations=r[0], primal_res=r[1], dual_res=r[2], gap=r[3],
                                  objective=r[4], rel_obj_err=abs(r[4] - f) / max(1.0, abs(f)))
                             for r in s.info.refinement]
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--m", type=int, default=200_000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nnz", type=int, default=8_000_000)
    ap.add_argument("--precision", type=int, nargs="+", default=[64, 32])
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--time-limit", type=float, default=240.0)
    ap.add_argument("--column-order", nargs="+", default=["given"], choices=("given", "locality"))
    a = ap.parse_args()
    for order in a.column_order:
        for prec in a.precision:
            print(json.dumps(run(a.m, a.n, a.nnz, prec, a.tol, a.time_limit, column_order=order)),
                  flush=True)


if __name__ == "__main__":
    main()
