"""Algorithmic HBM bytes of the device PDLP phases.

Each phase of the device state machine (one launch's work for one LP) reads
and writes a known set of arrays.  ``phase_bytes`` counts every compulsory
array once: matrix values and indices once per pass, the gathered vector once
(its length, not once per nonzero), pointer and tile arrays once, and each
row/column vector the phase touches once per read and once per write.  Cache
reuse of a gathered vector is therefore not charged, and neither are re-reads
that a perfect cache would avoid.  This is the yardstick that a kernel's
achieved bandwidth (bytes / time) is measured against; hardware counters give
the actual traffic.

An SpMV ``y = M v`` over R rows, C columns and Z nonzeros costs
``Z (s_val + 4) + (R + 1) 4 + C s_v + R s_y`` (4-byte indices and pointers).

The engines count how often each phase ran (``sovlp_*_phase_counts``); the
B phases that also formed the next primal step are counted a second time in
the ``*_NEXT`` slots, which carry only the extra bytes.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib

__all__ = ["phase_names", "phase_bytes", "bytes_of_counts", "traffic_record", "counter_per_kernel"]

_F64 = 8
_IDX = 4


def phase_names() -> list[str]:
    """Slot names in the library's order."""
    n = C.c_int()
    _lib.call("sovlp_phase_slots", C.byref(n))
    return [_lib.lib().sovlp_phase_name(i).decode() for i in range(n.value)]


def phase_bytes(name: str, m: int, n: int, nnz: int, ntiles_a: int, ntiles_t: int,
                precision: int) -> int:
    """Algorithmic bytes one execution of phase ``name`` moves for one LP."""
    s = 8 if precision == 64 else 4
    # structure of one pass over A (rows m) or A^T (rows n): indices, pointers, tiles
    pa = nnz * _IDX + (m + 1) * _IDX + (ntiles_a + 1) * _IDX
    pt = nnz * _IDX + (n + 1) * _IDX + (ntiles_t + 1) * _IDX
    b = {
        "INIT": _F64 * (m + n),
        # |a| r c products over A and A^T (FP64 original values); factors gathered
        "RUIZ_MAX": pa + pt + 2 * nnz * _F64 + 3 * _F64 * (m + n),
        "RUIZ_APPLY": 3 * _F64 * (m + n),
        "PC_SUM": pa + pt + 2 * nnz * _F64 + 3 * _F64 * (m + n),
        "PC_APPLY": 3 * _F64 * m + 4 * _F64 * n,
        # scaled copy of A and A^T, bounds, c, x0, y0
        "SCALE": pa + pt + 2 * nnz * (_F64 + s) + 2 * _F64 * (m + n)
                 + n * (5 * _F64 + 7 * s) + m * (3 * _F64 + 5 * s),
        "AX0": pa + nnz * s + n * s + m * s + n * (_F64 + s),
        "POW_A": pa + nnz * s + n * s + m * s,
        "POW_T": pt + nnz * s + m * s + n * s,
        "KKT0": 4 * s * m + 5 * s * n,
        "P_TRIAL": 7 * s * n,
        "A_TRIAL": pa + nnz * s + n * s + 7 * s * m,
        "B_ADAPT": pt + nnz * s + m * s + 6 * s * n + 7 * s * m,
        "A_FUSED": pa + nnz * s + n * s + 9 * s * m,
        "B_FUSED": pt + nnz * s + m * s + 5 * s * n,
        "RESTART": 8 * s * (m + n),
        "COPY_AVG": 2 * s * (m + n),
        "FINAL": n * (s + 5 * _F64) + m * (s + 2 * _F64),
        "FINAL_D": pt + nnz * _F64 + m * _F64 + 2 * _F64 * n,
        # refinement rounds (FP32 engines): X += x col / Delta
        "REF_X": n * (s + 5 * _F64),
        # FP64 KKT of X on the original: A X (kept) and A^T y
        "REF_KKT": (pa + nnz * _F64 + n * _F64 + m * (s + 5 * _F64))
                   + (pt + nnz * _F64 + m * (s + _F64) + n * 4 * _F64),
        # shifted, scaled bounds of the next round and its starting point
        "REF_SHIFT": m * (6 * _F64 + 5 * s) + n * (6 * _F64 + 5 * s),
        "DONE": 0,
        # extra bytes of a B phase that also formed the next primal step:
        # read c, lo, hi; write dx and x (or x')
        "B_ADAPT_NEXT": 5 * s * n,
        "B_FUSED_NEXT": 5 * s * n,
    }
    if name == "CHECK":
        # two-vector passes over A and A^T (current and average points)
        v = (pa + nnz * s + 2 * s * n + 7 * s * m) + (pt + nnz * s + 2 * s * m + 8 * s * n)
        if precision == 64:
            # unscaled KKT from the same products: row/col factors and original bounds
            v += 3 * _F64 * m + 4 * _F64 * n
        else:
            # unscaled KKT recomputed in FP64 from the original matrix, one
            # vector per pass: current point (sum kept in tmpm / tmpn), then average
            v += pa + nnz * _F64 + n * (s + _F64) + m * _F64
            v += pa + nnz * _F64 + n * (s + _F64) + m * (2 * s + 4 * _F64)
            v += pt + nnz * _F64 + m * (s + _F64) + n * _F64
            v += pt + nnz * _F64 + m * (s + _F64) + n * (2 * s + 5 * _F64)
        return int(v)
    return int(b.get(name, 0))


def bytes_of_counts(counts: np.ndarray, sizes, precision: int) -> float:
    """Total algorithmic bytes of a batch: ``counts`` is [LPs, slots] as the
    engine reports it, ``sizes`` a list of (m, n, nnz, ntiles_a, ntiles_t)."""
    names = phase_names()
    counts = np.asarray(counts, dtype=np.float64).reshape(len(sizes), len(names))
    used = [k for k in range(len(names)) if counts[:, k].any()]
    total = 0.0
    for k in used:
        per_lp = np.array([phase_bytes(names[k], *sz, precision) for sz in sizes], np.float64)
        total += float(counts[:, k] @ per_lp)
    return total


# ---------------------------------------------------------------------------
# The traffic record bench.py reports as a kernel's `traffic`: rocprofv3
# counter readings (FETCH_SIZE, WRITE_SIZE, in KiB per dispatch) of the
# profiled kernels, each scaled by what the same counter reads for a plain
# device copy of known size.
#
#   python -m sovlp.lp.traffic --bench-json B --fetch F --write W \
#       --copy-fetch CF --copy-write CW --out profiles/<record>.json
#
# B is the JSON line of the profiled bench run, F / W the output directories of
# its two counter passes (`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`,
# csv output), CF / CW the same two passes over `bench.py --only copy`.  When B
# is the line of `bench.py --only spmv`, the record holds each tiled product's
# traffic per launch instead (spmv_record).
# ---------------------------------------------------------------------------

def counter_rows(path: str) -> list[dict]:
    """Every row of the *counter_collection.csv files under ``path``."""
    import csv
    import glob
    import os
    rows = []
    files = [path] if os.path.isfile(path) else sorted(
        glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True))
    for f in files:
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    if not rows:
        raise FileNotFoundError(f"no counter_collection.csv rows under {path}")
    return rows


def counter_per_kernel(path: str, counter: str) -> dict[str, tuple[float, int]]:
    """(sum of the counter over dispatches, dispatches) per kernel name."""
    out: dict[str, list] = {}
    for r in counter_rows(path):
        if r.get("Counter_Name") != counter:
            continue
        k = r["Kernel_Name"]
        acc = out.setdefault(k, [0.0, set()])
        acc[0] += float(r["Counter_Value"])
        acc[1].add(r.get("Dispatch_Id") or r.get("Correlation_Id"))
    return {k: (v[0], len(v[1])) for k, v in out.items()}


def copy_scale(path: str, counter: str, nbytes: int, match: str = "copy") -> tuple[float, float]:
    """(KiB per copy dispatch, bytes moved / KiB read) of the copy kernels
    (kernel names containing ``match``) of a ``bench.py --only copy`` run."""
    per = {k: v for k, v in counter_per_kernel(path, counter).items() if match in k.lower()}
    if not per:
        raise ValueError(f"no kernel named like '{match}' with {counter} under {path}")
    total = sum(v[0] for v in per.values())
    disp = sum(v[1] for v in per.values())
    kib = total / disp
    return kib, nbytes / (kib * 1024.0)


def traffic_record(bench_line: dict, fetch: str, write: str, copy_fetch: str, copy_write: str,
                   copy_bytes: int = 256 << 20) -> dict:
    """The record for the bench line's dominant kernels (the roofline's
    `kernel`, several joined by ' + '): their counter sums over the profiled
    step, scaled by the copy calibration, per graph replay."""
    roof = bench_line["roofline"]
    names = [k.strip() for k in roof["kernel"].split(" + ")]
    replays = roof["launches"] // bench_line["steps"]

    def pick(path, counter):
        per = counter_per_kernel(path, counter)
        got = {}
        for n in names:
            hits = {k: v for k, v in per.items() if n in k}
            if not hits:
                raise ValueError(f"kernel {n} has no {counter} rows under {path}")
            got[n] = sum(v[0] for v in hits.values())
        return got

    f, w = pick(fetch, "FETCH_SIZE"), pick(write, "WRITE_SIZE")
    fk, fs = copy_scale(copy_fetch, "FETCH_SIZE", copy_bytes)
    wk, ws = copy_scale(copy_write, "WRITE_SIZE", copy_bytes)
    per_launch = (sum(f.values()) * fs + sum(w.values()) * ws) * 1024.0 / (
        replays * bench_line["steps"])
    return {
        "kernel": roof["kernel"],
        "workload": roof["traffic_key"],
        "replays_per_step": replays,
        "fetch_size_kib_per_step": {k: v / bench_line["steps"] for k, v in f.items()},
        "write_size_kib_per_step": {k: v / bench_line["steps"] for k, v in w.items()},
        "copy_calibration": {"bytes_read_and_written": copy_bytes, "fetch_size_kib": fk,
                             "write_size_kib": wk, "fetch_scale": fs, "write_scale": ws},
        "traffic_bytes_per_launch": per_launch,
        "algorithmic_bytes_per_launch": roof["algorithmic_bytes_per_launch"],
    }


def dispatch_series(path: str, counter: str, match: str) -> list[float]:
    """The counter's value per dispatch of the kernels named like ``match``, in
    dispatch order."""
    per: dict[int, float] = {}
    for r in counter_rows(path):
        if r.get("Counter_Name") != counter or match not in r["Kernel_Name"]:
            continue
        d = int(r.get("Dispatch_Id") or r.get("Correlation_Id"))
        per[d] = per.get(d, 0.0) + float(r["Counter_Value"])
    return [per[d] for d in sorted(per)]


_SPMV_KERNEL = {64: "spmv_tiled_kernel<double", 32: "spmv_tiled_kernel<float"}
SPMV_KEYS = ("op", "kernel", "precision", "rows", "cols", "nnz", "grid", "tiles")


def spmv_record(section: dict, fetch: str, write: str, copy_fetch: str, copy_write: str,
                copy_bytes: int = 256 << 20) -> dict:
    """Counted traffic per launch of each tiled product of a ``bench.py --only
    spmv`` run: the dispatches of each precision's kernel, in order, are the
    products of that precision in the section's order, ``warmup + reps`` each;
    the mean over each product's last ``reps`` dispatches, scaled by the copy
    calibration."""
    _, fs = copy_scale(copy_fetch, "FETCH_SIZE", copy_bytes)
    _, ws = copy_scale(copy_write, "WRITE_SIZE", copy_bytes)
    entries = []
    for prec, match in _SPMV_KERNEL.items():
        prods = [r for r in section["results"]
                 if r["precision"] == prec and r["kernel"].startswith("tiled")]
        f, w = dispatch_series(fetch, "FETCH_SIZE", match), dispatch_series(write, "WRITE_SIZE", match)
        need = sum(r["warmup"] + r["reps"] for r in prods)
        if len(f) != need or len(w) != need:
            raise ValueError(f"{match}: {len(f)} / {len(w)} dispatches counted, {need} expected")
        at = 0
        for r in prods:
            lo, hi = at + r["warmup"], at + r["warmup"] + r["reps"]
            fk, wk = float(np.mean(f[lo:hi])), float(np.mean(w[lo:hi]))
            e = {k: r[k] for k in SPMV_KEYS}
            e.update(fetch_kib_per_launch=fk, write_kib_per_launch=wk,
                     traffic_bytes_per_launch=(fk * fs + wk * ws) * 1024.0,
                     algorithmic_bytes=r["bytes"])
            entries.append(e)
            at = hi
    return {"lp": section["lp"], "copy_scale": {"fetch": fs, "write": ws}, "products": entries}


def _main(argv=None) -> int:
    import argparse
    import json
    ap = argparse.ArgumentParser(description="assemble the traffic record of a bench run")
    ap.add_argument("--bench-json", required=True, help="file whose last line is the bench JSON")
    for k in ("fetch", "write", "copy-fetch", "copy-write"):
        ap.add_argument(f"--{k}", required=True)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    line = json.loads(open(a.bench_json).read().strip().splitlines()[-1])
    if "spmv" in line and "roofline" not in line:
        rec = spmv_record(line["spmv"], a.fetch, a.write, a.copy_fetch, a.copy_write)
        summary = [(e["op"], e["kernel"], e["precision"], e["traffic_bytes_per_launch"])
                   for e in rec["products"]]
    else:
        rec = traffic_record(line, a.fetch, a.write, a.copy_fetch, a.copy_write)
        summary = {k: rec[k] for k in ("traffic_bytes_per_launch", "algorithmic_bytes_per_launch")}
    with open(a.out, "w") as fh:
        json.dump(rec, fh, indent=1)
        fh.write("\n")
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
