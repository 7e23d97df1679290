"""Autotuner of the batched engine's kernel configuration, per LP shape.

No reference counterpart: the reference launches one fixed shape, a 32-lane
warp per row in 128-thread blocks (``src/sovopt/core/backend.py:373-384``).

A configuration is ``(block, tile_rows, tile_nnz)``: the workgroup size that
runs the LP (64, 128, 256 or 512 threads; one compiled kernel each), and the caps of
the host tile plan — rows per tile (``block`` to ``4 block``: up to four short
rows per lane) and nonzeros per tile (up to the kernel's LDS tile). LPs are
bucketed by their nonzero count; for every bucket the tuner solves a sample of
the caller's LPs, replicated to fill the GPU, once under each candidate and
keeps the one with the highest bulk throughput: 95 % of the LPs over the device
time from the first start to the 95th-percentile finish (the last few LPs of a
batch are the slowest-converging ones whatever the configuration). The table is keyed by the
bucket only, so an LP's configuration — and with it its arithmetic — depends
only on its own shape, never on the other LPs of a batch.

``python -m sovlp.lp.autotune --out table.json`` tunes on the uneven-batch
workload and writes the table; ``Autotuner.load`` reads it back.
"""

from __future__ import annotations

import argparse
import json
from dataclasses import dataclass, field

import numpy as np

from ..core.problem import Problem
from .batch import batch_geometry

__all__ = ["Autotuner", "CANDIDATES", "NNZ_EDGES", "bucket_of", "packaged_table"]

# (block, rows per lane, fraction of the kernel's tile nonzeros)
CANDIDATES = [(512, 4, 1.0), (512, 4, 0.5), (256, 4, 1.0), (256, 1, 1.0), (256, 4, 0.5),
              (128, 4, 1.0), (128, 1, 1.0), (64, 4, 1.0)]
# nonzero-count bucket edges
NNZ_EDGES = [0, 1024, 4096, 16384, 1 << 62]


def packaged_table(precision: int) -> str:
    """Path of the table tuned on an MI355X that ships with the package."""
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        f"tuned_mi355x_fp{int(precision)}.json")


def bucket_of(p: Problem) -> int:
    z = int(p.A.nnz)
    for b in range(len(NNZ_EDGES) - 1):
        if NNZ_EDGES[b] <= z < NNZ_EDGES[b + 1]:
            return b
    return len(NNZ_EDGES) - 2


def bulk_rate(infos, hz: float, frac: float = 0.95) -> float:
    """LPs per second of the first ``frac`` of a batch to finish, from the
    engines' device clocks (start tick, ticks used)."""
    start = np.array([inf.start_tick for inf in infos], dtype=np.float64)
    end = start + np.array([inf.device_ticks for inf in infos], dtype=np.float64)
    k = max(1, int(frac * len(end)))
    t = (np.sort(end)[k - 1] - start.min()) / hz
    return k / t if t > 0 else float("inf")


def config_of(precision: int, block: int, rows_per_lane: int, nnz_frac: float):
    rmax, tmax = batch_geometry(precision, block)
    return (int(block), int(block * rows_per_lane), max(64, int(tmax * nnz_frac)))


@dataclass
class Autotuner:
    precision: int = 32
    table: dict = field(default_factory=dict)        # bucket -> (block, tile_rows, tile_nnz)
    measured: dict = field(default_factory=dict)     # bucket -> {config: LPs/s}
    device: str = ""

    def default(self):
        return config_of(self.precision, 256, 4, 1.0)

    def choose(self, p: Problem):
        """The configuration of LP ``p`` (a function of its shape only).  A
        bucket the tuning sample left empty takes the configuration of the
        nearest tuned bucket below it (LPs larger than any the sample held
        behave most like the largest it held), else above it."""
        b = bucket_of(p)
        if b in self.table:
            return tuple(self.table[b])
        below = [k for k in self.table if k < b]
        above = [k for k in self.table if k > b]
        if below:
            return tuple(self.table[max(below)])
        if above:
            return tuple(self.table[min(above)])
        return self.default()

    def tune(self, problems: list[Problem], params, fill: int = 1024, sample: int = 64,
             log=print) -> dict:
        """Measure every candidate on every bucket that ``problems`` populate."""
        import torch
        from .batch import BatchEngine
        self.device = torch.cuda.get_device_name(0)
        import ctypes as C
        from .. import _lib
        hz = C.c_double()
        _lib.call("sovlp_clock_rate", 0, C.byref(hz))
        hz = hz.value
        warmed = False
        by = {}
        for p in problems:
            by.setdefault(bucket_of(p), []).append(p)
        for b in sorted(by):
            base = by[b][:sample]
            reps = [base[i % len(base)] for i in range(max(fill, len(base)))]
            rates = {}
            for cand in CANDIDATES:
                cfg = config_of(self.precision, *cand)
                eng = BatchEngine(reps, params, precision=self.precision,
                                  configs=[cfg] * len(reps))
                try:
                    if not warmed:                  # module load and first launches
                        eng.run()
                        warmed = True
                    eng.run()
                    rates[cfg] = bulk_rate(eng.infos(), hz)
                finally:
                    eng.close()
                if log:
                    log(f"bucket {b} nnz [{NNZ_EDGES[b]}, {NNZ_EDGES[b + 1]}) {len(base)} LPs "
                        f"x{len(reps) // max(len(base), 1)}: block {cfg[0]} rows {cfg[1]} "
                        f"nnz {cfg[2]}: {rates[cfg]:.1f} LPs/s")
            best = max(rates, key=rates.get)
            self.table[b] = best
            self.measured[b] = rates
        return self.table

    def to_json(self) -> dict:
        return {"precision": self.precision, "device": self.device, "nnz_edges": NNZ_EDGES,
                "table": {str(b): list(c) for b, c in self.table.items()},
                "measured_lps_per_s": {str(b): {",".join(map(str, c)): r for c, r in m.items()}
                                       for b, m in self.measured.items()}}

    def save(self, path: str) -> None:
        with open(path, "w") as f:
            json.dump(self.to_json(), f, indent=1)

    @classmethod
    def from_json(cls, d: dict) -> "Autotuner":
        if d.get("nnz_edges") != NNZ_EDGES:
            raise ValueError("table was tuned with other buckets")
        t = cls(precision=int(d["precision"]), device=d.get("device", ""))
        t.table = {int(b): tuple(int(v) for v in c) for b, c in d["table"].items()}
        for b, c in t.table.items():
            rmax, tmax = batch_geometry(t.precision, c[0])
            if not (c[1] <= rmax and c[2] <= tmax):
                raise ValueError(f"bucket {b}: configuration {c} exceeds its kernel")
        t.measured = {int(b): {tuple(int(v) for v in k.split(",")): float(r)
                               for k, r in m.items()}
                      for b, m in d.get("measured_lps_per_s", {}).items()}
        return t

    @classmethod
    def load(cls, path: str) -> "Autotuner":
        with open(path) as f:
            return cls.from_json(json.load(f))


def main():
    ap = argparse.ArgumentParser(description="tune the batched engine on the uneven-batch LPs")
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--lps", type=int, default=1024, help="LPs of the workload to bucket")
    ap.add_argument("--seed", type=int, default=20000,
                    help="first seed (default: LPs disjoint from the benchmark's seeds 1000..5095)")
    ap.add_argument("--fill", type=int, default=1024)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    from ..workloads import uneven_batch
    from .pdlp import PDLPParams
    probs = [k.problem for k in uneven_batch(a.lps, a.seed)]
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6, device="gpu")
    t = Autotuner(a.precision)
    t.tune(probs, par, fill=a.fill, log=lambda s: print(s, flush=True))
    t.save(a.out)
    print(json.dumps(t.to_json()["table"]))


if __name__ == "__main__":
    main()
