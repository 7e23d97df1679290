"""Roofline measurement of the engines' sparse product (the tile pass).

Times ``y = A x`` and ``y = Aᵀ x`` on the large skewed LP with HIP events on
the stream the kernel runs on, and reports algorithmic bandwidth

    B = Z·(s_val + s_idx) + (R+1)·s_ptr + C·s_x + R·s_y

per launch against MI355X's 8 TB/s HBM peak. A rotating set of vector copies is
used so the gathered vector is not trivially cache-resident between launches.
"""

from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from .lp.pdlp import plan_tiles, tile_geometry

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(rows: int, cols: int, nnz: int, s_val: int, s_idx: int = 4,
                      s_ptr: int = 4) -> int:
    return nnz * (s_val + s_idx) + (rows + 1) * s_ptr + cols * s_val + rows * s_val


def measure(ptr: np.ndarray, idx: np.ndarray, val: np.ndarray, ncols: int, precision: int = 64,
            reps: int = 50, warmup: int = 5, grid: int | None = None):
    """Returns dict(ms, gbs, bytes, ...) for the tiled SpMV of the CSR (ptr, idx, val)."""
    import torch
    rows = ptr.size - 1
    nnz = int(ptr[-1])
    block, trows, tnnz = tile_geometry(precision)
    tiles = plan_tiles(ptr, tnnz, trows, block)
    nt = tiles.size - 1
    if grid is None:
        props = torch.cuda.get_device_properties(0)
        grid = max(1, min(nt, props.multi_processor_count * 8))
    part = np.zeros(grid + 1, dtype=np.int32)
    _lib.call("sovlp_partition_tiles", nt, tiles.ctypes.data,
              np.ascontiguousarray(ptr, np.int64).ctypes.data, grid, part.ctypes.data)
    td = torch.float64 if precision == 64 else torch.float32
    d_ptr = torch.from_numpy(ptr.astype(np.int32)).cuda()
    d_idx = torch.from_numpy(idx.astype(np.int32)).cuda()
    d_val = torch.from_numpy(val).to(td).cuda()
    d_tiles = torch.from_numpy(tiles).cuda()
    d_part = torch.from_numpy(part).cuda()
    nrot = 4
    xs = [torch.randn(ncols, dtype=td, device="cuda") for _ in range(nrot)]
    y = torch.empty(rows, dtype=td, device="cuda")
    fn = getattr(_lib.lib(), "sovlp_spmv_tiled_f64" if precision == 64 else "sovlp_spmv_tiled_f32")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def launch(i):
        _lib.check(fn(rows, d_ptr.data_ptr(), d_idx.data_ptr(), d_val.data_ptr(), d_tiles.data_ptr(),
                      nt, d_part.data_ptr(), grid, xs[i % nrot].data_ptr(), y.data_ptr(), sh),
                   "spmv_tiled")

    for i in range(warmup):
        launch(i)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for i in range(reps):
        launch(i)
    stop.record(stream)
    stop.synchronize()
    ms = start.elapsed_time(stop) / reps
    s_val = 8 if precision == 64 else 4
    b = algorithmic_bytes(rows, ncols, nnz, s_val)
    return dict(rows=rows, cols=ncols, nnz=nnz, precision=precision, grid=grid, tiles=nt,
                ms=ms, bytes=b, gbs=b / (ms * 1e-3) / 1e9, frac=b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)


def skewed_roofline(reps: int = 50, seed: int = 1):
    """SpMV roofline of A and Aᵀ of the 200k x 1M skewed LP, FP64 and FP32."""
    from .workloads import skewed_lp
    t = time.perf_counter()
    A = skewed_lp(seed).problem.A
    gen_s = time.perf_counter() - t
    out = []
    for prec in (64, 32):
        out.append(dict(op="A*x", **measure(A.rp, A.ri, A.rx, A.n, prec, reps)))
        out.append(dict(op="At*y", **measure(A.cp, A.ci, A.cx, A.m, prec, reps)))
    return out, gen_s


if __name__ == "__main__":
    import json
    res, gen = skewed_roofline()
    print(json.dumps(dict(gen_s=gen, results=res), indent=1))
