"""Roofline measurement of the engines' sparse product (the tile pass).

Times ``y = A x`` and ``y = Aᵀ x`` on the large skewed LP with HIP events on
the stream the kernel runs on (and ``A x`` also as eight column panels, one per
XCD: ``PanelSpmv``; and both with the columns relabelled by
``locality_order``, which leaves every row sum of A x bit for bit the same),
and reports algorithmic bandwidth

    B = Z·(s_val + s_idx) + (R+1)·s_ptr + C·s_x + R·s_y

per launch against MI355X's 8 TB/s HBM peak. A rotating set of vector copies is
used so the gathered vector is not trivially cache-resident between launches.
"""

from __future__ import annotations

import time

import numpy as np

from . import _lib
from .core.sparse import locality_order, relabel_columns
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
    ptr64 = np.ascontiguousarray(ptr, np.int64)      # held: C reads it by address
    _lib.call("sovlp_partition_tiles", nt, tiles.ctypes.data, ptr64.ctypes.data, grid,
              part.ctypes.data)
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
                ms=ms, bytes=b, gbs=b / (ms * 1e-3) / 1e9, frac=b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                warmup=warmup, reps=reps)


def split_panels(ptr: np.ndarray, idx: np.ndarray, val: np.ndarray, ncols: int, panels: int):
    """Column panels of a CSR matrix: panel p holds columns [p w, (p+1) w), w =
    ceil(ncols / panels), as its own CSR. Returns (ptrs [panels, rows+1] int32
    offsets into the concatenated arrays, idx, val), panel-major, each row's
    entries in their original (column) order."""
    rows = ptr.size - 1
    w = -(-ncols // panels)
    row_of = np.repeat(np.arange(rows, dtype=np.int64), np.diff(ptr))
    key = (idx.astype(np.int64) // w) * rows + row_of
    order = np.argsort(key, kind="stable")
    counts = np.bincount(key, minlength=panels * rows)
    ptrs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ptrs2 = np.empty((panels, rows + 1), dtype=np.int32)
    for p in range(panels):
        ptrs2[p] = ptrs[p * rows: (p + 1) * rows + 1]
    return ptrs2, idx[order].astype(np.int32), val[order]


class PanelSpmv:
    """The column-panel SpMV (``sovlp_spmv_panels_*``) of one CSR matrix, laid out
    on the device: per-panel CSR, tile plans and tile-range splits."""

    def __init__(self, ptr, idx, val, ncols: int, precision: int = 64, panels: int = 8,
                 groups: int | None = None):
        import torch
        self.rows, self.panels, self.precision = ptr.size - 1, int(panels), int(precision)
        block, trows, tnnz = tile_geometry(precision)
        ptrs, pidx, pval = split_panels(np.asarray(ptr, np.int64), np.asarray(idx), np.asarray(val),
                                        ncols, panels)
        tl, offs, nts = [], [0], []
        for p in range(panels):
            t = plan_tiles(ptrs[p].astype(np.int64), tnnz, trows, block)
            tl.append(t)
            nts.append(t.size - 1)
            offs.append(offs[-1] + t.size)
        self.groups = int(groups) if groups else max(max(nts), 1)
        parts = np.zeros((panels, self.groups + 1), dtype=np.int32)
        for p in range(panels):
            ptr64 = np.ascontiguousarray(ptrs[p], np.int64)   # held: C reads it by address
            _lib.call("sovlp_partition_tiles", nts[p], tl[p].ctypes.data, ptr64.ctypes.data,
                      self.groups, parts[p].ctypes.data)
        td = torch.float64 if precision == 64 else torch.float32
        self.d = dict(ptrs=torch.from_numpy(ptrs.ravel()).cuda(),
                      idx=torch.from_numpy(pidx).cuda(),
                      val=torch.from_numpy(pval).to(td).cuda(),
                      tiles=torch.from_numpy(np.concatenate(tl)).cuda(),
                      toff=torch.from_numpy(np.array(offs[:-1], np.int32)).cuda(),
                      parts=torch.from_numpy(parts.ravel()).cuda())
        self.partial = torch.empty(panels * self.rows, dtype=td, device="cuda")
        self.fn = getattr(_lib.lib(), "sovlp_spmv_panels_f64" if precision == 64
                          else "sovlp_spmv_panels_f32")
        self.tiles = int(sum(nts))

    def __call__(self, x, y, stream) -> None:
        d = self.d
        _lib.check(self.fn(self.panels, self.rows, d["ptrs"].data_ptr(), d["idx"].data_ptr(),
                           d["val"].data_ptr(), d["tiles"].data_ptr(), d["toff"].data_ptr(),
                           d["parts"].data_ptr(), self.groups, x.data_ptr(),
                           self.partial.data_ptr(), y.data_ptr(), stream), "spmv_panels")


def measure_panels(ptr, idx, val, ncols: int, precision: int = 64, panels: int = 8,
                   groups: int | None = None, reps: int = 50, warmup: int = 5):
    """Time the column-panel SpMV (both launches) like ``measure``."""
    import torch
    sp = PanelSpmv(ptr, idx, val, ncols, precision, panels, groups)
    td = torch.float64 if precision == 64 else torch.float32
    xs = [torch.randn(ncols, dtype=td, device="cuda") for _ in range(4)]
    y = torch.empty(sp.rows, dtype=td, device="cuda")
    stream = torch.cuda.current_stream()
    for i in range(warmup):
        sp(xs[i % 4], y, stream.cuda_stream)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for i in range(reps):
        sp(xs[i % 4], y, stream.cuda_stream)
    stop.record(stream)
    stop.synchronize()
    ms = start.elapsed_time(stop) / reps
    s_val = 8 if precision == 64 else 4
    nnz = int(ptr[-1])
    b = algorithmic_bytes(sp.rows, ncols, nnz, s_val)
    return dict(rows=sp.rows, cols=ncols, nnz=nnz, precision=precision, panels=panels,
                groups=sp.groups, tiles=sp.tiles, ms=ms, bytes=b, gbs=b / (ms * 1e-3) / 1e9,
                frac=b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)


def skewed_roofline(reps: int = 50, seed: int = 1, A=None, panels: bool = False):
    """SpMV roofline of A and Aᵀ of the 200k x 1M skewed LP (``A``: its matrix,
    built from ``seed`` when not given), FP64 and FP32, in the given column
    order and in locality order; ``panels``: also A x as eight column panels."""
    from .workloads import skewed_lp
    t = time.perf_counter()
    if A is None:
        A = skewed_lp(seed).problem.A
    gen_s = time.perf_counter() - t
    B = relabel_columns(A, locality_order(A))
    out = []
    for prec in (64, 32):
        out.append(dict(op="A*x", kernel="tiled", **measure(A.rp, A.ri, A.rx, A.n, prec, reps)))
        out.append(dict(op="A*x", kernel="tiled, columns in locality order",
                        **measure(B.rp, B.ri, B.rx, B.n, prec, reps)))
        if panels:
            out.append(dict(op="A*x", kernel="8 column panels",
                            **measure_panels(A.rp, A.ri, A.rx, A.n, prec, 8, reps=reps)))
        out.append(dict(op="At*y", kernel="tiled", **measure(A.cp, A.ci, A.cx, A.m, prec, reps)))
        out.append(dict(op="At*y", kernel="tiled, columns in locality order",
                        **measure(B.cp, B.ci, B.cx, B.m, prec, reps)))
    return out, gen_s
