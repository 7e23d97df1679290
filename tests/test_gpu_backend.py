"""GPU kernels against the CPU oracle: the reference's tests/test_gpu.py restated
(seeded ``_random_csr``, rtol 1e-12 / 1e-13 / 1e-4, edge row lengths with a
poisoned output, bitwise repeatability), plus the engines' tiled SpMV."""

import numpy as np
import pytest

from oracle import oracle
from sovlp.core.backend import get_backend, gpu_selftest, Backend
from sovlp.core.sparse import SparseMatrix

pytestmark = pytest.mark.gpu


def _random_csr(m, n, per_col, seed):
    rng = np.random.default_rng(seed)
    cols = np.repeat(np.arange(n), per_col)
    rows = rng.integers(0, m, cols.size)
    vals = rng.standard_normal(cols.size)
    return SparseMatrix.from_triplets(rows, cols, vals, m, n), rng


def _to(bk, A):
    return bk.pointer(A.rp), bk.index(A.ri), bk.to_device(A.rx)


def test_selftest_and_backend():
    ok, why = gpu_selftest()
    assert ok, why
    assert get_backend("gpu").device == "gpu"
    assert get_backend("auto").device == "gpu"


@pytest.mark.parametrize("lanes", [1, 4, 16, 64])
def test_spmv_matches_oracle(lanes):
    A, rng = _random_csr(400, 600, 5, seed=1)
    x = rng.standard_normal(A.n)
    gpu = get_backend("gpu")
    out = gpu.zeros(A.m)
    gpu.spmv(*_to(gpu, A), gpu.to_device(x), out, A.m, lanes_per_row=lanes)
    gpu.sync()
    np.testing.assert_allclose(gpu.to_host(out), oracle.spmv(A.rp, A.ri, A.rx, x, A.m),
                               rtol=1e-12, atol=1e-12)


def test_spmv_handles_empty_and_wave_boundary_rows():
    lens = [0, 1, 31, 32, 33, 63, 64, 65, 127, 129]
    m, n = len(lens), 256
    rng = np.random.default_rng(2)
    rows, cols, vals = [], [], []
    for i, length in enumerate(lens):
        c = rng.choice(n, size=length, replace=False)
        rows += [i] * length
        cols += list(c)
        vals += list(rng.standard_normal(length))
    A = SparseMatrix.from_triplets(np.array(rows), np.array(cols), np.array(vals), m, n)
    x = rng.standard_normal(n)
    gpu = get_backend("gpu")
    for lanes in (1, 8, 64):
        out = gpu.zeros(m) + 12345.0
        gpu.spmv(*_to(gpu, A), gpu.to_device(x), out, m, lanes_per_row=lanes)
        gpu.sync()
        np.testing.assert_allclose(gpu.to_host(out), oracle.spmv(A.rp, A.ri, A.rx, x, m),
                                   rtol=1e-12, atol=1e-12)


def test_spmv_is_bitwise_reproducible():
    A, rng = _random_csr(500, 700, 6, seed=3)
    x = rng.standard_normal(A.n)
    gpu = get_backend("gpu")
    dev, xd = _to(gpu, A), gpu.to_device(x)
    first = None
    for _ in range(4):
        out = gpu.zeros(A.m)
        gpu.spmv(*dev, xd, out, A.m)
        got = gpu.to_host(out)
        if first is None:
            first = got
        else:
            assert np.array_equal(got, first)


def test_float32_kernels_run():
    A, rng = _random_csr(200, 300, 4, seed=4)
    x = rng.standard_normal(A.n)
    gpu = Backend("gpu", dtype=np.float32)
    out = gpu.zeros(A.m)
    gpu.spmv(*_to(gpu, A), gpu.to_device(x), out, A.m)
    gpu.sync()
    assert gpu.to_host(out).dtype == np.float32
    np.testing.assert_allclose(gpu.to_host(out), A.matvec(x), rtol=1e-4, atol=1e-4)


def _tiled(A_ptr, A_idx, A_val, x, rows, dtype, grid=7):
    import ctypes as C
    import torch
    from sovlp import _lib
    from sovlp.lp.pdlp import plan_tiles, tile_geometry
    prec = 64 if dtype == np.float64 else 32
    block, trows, tnnz = tile_geometry(prec)
    tiles = plan_tiles(A_ptr, tnnz, trows, block)
    nt = tiles.size - 1
    part = np.zeros(grid + 1, dtype=np.int32)
    ptr64 = np.ascontiguousarray(A_ptr, np.int64)    # held: C reads it by address
    _lib.call("sovlp_partition_tiles", nt, tiles.ctypes.data, ptr64.ctypes.data, grid,
              part.ctypes.data)
    td = torch.float64 if prec == 64 else torch.float32
    dp = torch.from_numpy(A_ptr.astype(np.int32)).cuda()
    di = torch.from_numpy(A_idx.astype(np.int32)).cuda()
    dv = torch.from_numpy(A_val).to(td).cuda()
    dt = torch.from_numpy(tiles).cuda()
    dpart = torch.from_numpy(part).cuda()
    dx = torch.from_numpy(x).to(td).cuda()
    y = torch.full((rows,), 12345.0, dtype=td, device="cuda")
    fn = "sovlp_spmv_tiled_f64" if prec == 64 else "sovlp_spmv_tiled_f32"
    _lib.call(fn, rows, dp.data_ptr(), di.data_ptr(), dv.data_ptr(), dt.data_ptr(), nt,
              dpart.data_ptr(), grid, dx.data_ptr(), y.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    return y.cpu().numpy()


def test_tiled_spmv_on_skewed_rows_matches_oracle():
    # rows of length 0..5000 interleaved: long-row tiles, stream tiles, empty rows
    rng = np.random.default_rng(9)
    n = 6000
    lens = np.concatenate([rng.integers(0, 4, 3000), [5000, 2049, 2048, 4097], rng.integers(0, 300, 50)])
    rng.shuffle(lens)
    m = lens.size
    rows = np.repeat(np.arange(m), lens)
    cols = np.concatenate([rng.choice(n, size=l, replace=False) for l in lens])
    A = SparseMatrix.from_triplets(rows, cols, rng.standard_normal(rows.size), m, n)
    x = rng.standard_normal(n)
    ref = oracle.spmv(A.rp, A.ri, A.rx, x, m)
    got = _tiled(A.rp, A.ri, A.rx, x, m, np.float64)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    got2 = _tiled(A.rp, A.ri, A.rx, x, m, np.float64)
    assert np.array_equal(got, got2)
    # the transpose through the CSC arrays
    ref_t = oracle.spmv(A.cp, A.ci, A.cx, ref, n)
    got_t = _tiled(A.cp, A.ci, A.cx, ref, n, np.float64)
    np.testing.assert_allclose(got_t, ref_t, rtol=1e-11, atol=1e-11)
    got32 = _tiled(A.rp, A.ri, A.rx, x, m, np.float32)
    np.testing.assert_allclose(got32, ref, rtol=1e-4, atol=2e-4)


def test_primal_and_dual_steps_match_oracle_formulas():
    rng = np.random.default_rng(5)
    n, m = 1000, 700
    x, c, aty = (rng.standard_normal(n) for _ in range(3))
    lo, hi = -np.abs(rng.standard_normal(n)), np.abs(rng.standard_normal(n))
    y, ax = rng.standard_normal(m), rng.standard_normal(m)
    rl, ru = -np.abs(rng.standard_normal(m)), np.abs(rng.standard_normal(m))
    tau, sigma = 0.37, 1.9
    exp_x = np.clip(x - tau * (c + aty), lo, hi)
    w = y + sigma * ax
    exp_y = w - sigma * np.clip(w / sigma, rl, ru)
    gpu = get_backend("gpu")
    got_x = gpu.primal_step(gpu.zeros(n), *(gpu.to_device(v) for v in (x, c, aty, lo, hi)), tau, n)
    got_y = gpu.dual_step(gpu.zeros(m), *(gpu.to_device(v) for v in (y, ax, rl, ru)), sigma, m)
    np.testing.assert_allclose(gpu.to_host(got_x), exp_x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(gpu.to_host(got_y), exp_y, rtol=1e-13, atol=1e-13)


def test_fused_pdhg_steps_match_unfused():
    rng = np.random.default_rng(6)
    n, m = 800, 600
    x0, c, aty = (rng.standard_normal(n) for _ in range(3))
    lo, hi = -np.ones(n), np.ones(n)
    y0, ax0, adx = (rng.standard_normal(m) for _ in range(3))
    rl, ru = -np.ones(m), np.ones(m)
    tau, sigma, eta = 0.25, 1.4, 0.6
    # unfused formulas (backend.py:441-479 CPU branch)
    xn = np.clip(x0 - tau * (c + aty), lo, hi)
    dx = xn - x0
    ax1 = ax0 + adx
    w = y0 + sigma * (ax1 + adx)
    yn = w - sigma * np.clip(w / sigma, rl, ru)
    dy = yn - y0
    acc = aty + dx
    run = eta * xn
    gpu = get_backend("gpu")
    gx, gdx = gpu.to_device(x0.copy()), gpu.zeros(n)
    gpu.pdhg_primal(gx, gdx, *(gpu.to_device(v) for v in (c, aty, lo, hi)), tau, n)
    gy, gdy, gax = gpu.to_device(y0.copy()), gpu.zeros(m), gpu.to_device(ax0.copy())
    gpu.pdhg_dual(gy, gdy, gax, *(gpu.to_device(v) for v in (adx, rl, ru)), sigma, m)
    gacc, grun = gpu.to_device(aty.copy()), gpu.zeros(n)
    gpu.pdhg_accum(gacc, gdx, grun, gx, eta, n)
    for exp, got in zip((xn, dx, yn, dy, ax1, acc, run), (gx, gdx, gy, gdy, gax, gacc, grun)):
        np.testing.assert_allclose(gpu.to_host(got), exp, rtol=1e-13, atol=1e-13)


def test_reductions_match():
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal(4096), rng.standard_normal(4096)
    gpu = get_backend("gpu")
    ag, bg = gpu.to_device(a), gpu.to_device(b)
    assert gpu.dot(ag, bg) == pytest.approx(float(np.dot(a, b)), rel=1e-12)
    assert gpu.nrm2(ag) == pytest.approx(float(np.linalg.norm(a)), rel=1e-12)
    assert gpu.nrm_inf(ag) == pytest.approx(float(np.abs(a).max()), rel=1e-12)
    assert gpu.nrm_inf(gpu.zeros(0)) == 0.0
    # fixed order: repeats are bitwise equal
    assert gpu.dot(ag, bg) == gpu.dot(ag, bg)


@pytest.mark.parametrize("precision,panels", [(64, 8), (64, 3), (32, 8)])
def test_column_panel_spmv_matches_oracle(precision, panels):
    import torch
    from sovlp.spmv_bench import PanelSpmv
    rng = np.random.default_rng(12)
    m, n = 700, 5000
    lens = np.concatenate([rng.integers(0, 6, m - 4), [0, 3000, 64, 4100]])
    rows = np.repeat(np.arange(m), lens)
    cols = rng.integers(0, n, rows.size)
    A = SparseMatrix.from_triplets(rows, cols, rng.standard_normal(rows.size), m, n)
    x = rng.standard_normal(n)
    sp = PanelSpmv(A.rp, A.ri, A.rx, n, precision, panels)
    td = torch.float64 if precision == 64 else torch.float32
    y = torch.full((m,), 12345.0, dtype=td, device="cuda")
    sp(torch.from_numpy(x).to(td).cuda(), y, torch.cuda.current_stream().cuda_stream)
    ref = oracle.spmv(A.rp, A.ri, A.rx, x, m)
    tol = 1e-12 if precision == 64 else 1e-4
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=tol, atol=tol)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tiled_spmv_with_columns_in_locality_order_is_bitwise_the_same(dtype):
    """Relabelling the columns (``locality_order``) changes where x is gathered
    from, not one product or one addition of A x: y = A x is bit for bit that of
    the given order. A^T y's rows move, so its tiles (and each row's reduction
    tree) change: it agrees to rounding."""
    from sovlp.core.sparse import locality_order, relabel_columns
    from sovlp.workloads import skewed_lp_degrees
    rng = np.random.default_rng(17)
    m, n = 3000, 40_000
    deg = skewed_lp_degrees(n, 300_000, dmax=2000)[rng.permutation(n)]
    cols = np.repeat(np.arange(n), deg)
    A = SparseMatrix.from_triplets(rng.integers(0, m, cols.size), cols,
                                   rng.standard_normal(cols.size), m, n)
    new = locality_order(A)
    B = relabel_columns(A, new)
    x = rng.standard_normal(n).astype(dtype)
    xn = np.empty_like(x)
    xn[new] = x
    y = _tiled(A.rp, A.ri, A.rx, x, m, dtype)
    yb = _tiled(B.rp, B.ri, B.rx, xn, m, dtype)
    assert np.array_equal(y.view(np.uint8), yb.view(np.uint8))
    v = rng.standard_normal(m).astype(dtype)
    t = _tiled(A.cp, A.ci, A.cx, v, n, dtype)
    tb = _tiled(B.cp, B.ci, B.cx, v, n, dtype)
    f64 = dtype == np.float64
    np.testing.assert_allclose(tb[new], t, rtol=1e-12 if f64 else 1e-4, atol=1e-12 if f64 else 2e-3)
    np.testing.assert_allclose(y, A.matvec(x.astype(np.float64)),
                               rtol=1e-10 if f64 else 1e-3, atol=1e-10 if f64 else 1e-3)


@pytest.mark.parametrize("k", [1, 3, 64, 300])
def test_spmm_is_the_reference_cpu_product_bit_for_bit(k):
    """``Backend.spmm`` (backend.py:386-400): out = A X, batch index fastest,
    with empty rows and rows of 1..9 nonzeros; FP64 against the oracle's CSR-order
    product (the reference's _csr_spmm), FP32 against the same in float32."""
    from oracle import bnr_oracle
    bk = get_backend("gpu")
    A, rng = _random_csr(70, 50, 3, 41 + k)
    X = rng.standard_normal((A.n, k))
    out = bk.zeros((A.m, k))
    bk.spmm(*_to(bk, A), bk.to_device(X), out, A.m, k)
    ref = bnr_oracle.spmm(A.rp, A.ri, A.rx, X, A.m)
    assert np.array_equal(bk.to_host(out), ref)
    b32 = get_backend("gpu", np.float32)
    X32 = X.astype(np.float32)
    rx32 = A.rx.astype(np.float32)
    out32 = b32.zeros((A.m, k))
    b32.spmm(bk.pointer(A.rp), bk.index(A.ri), b32.to_device(rx32), b32.to_device(X32), out32,
             A.m, k)
    ref32 = np.zeros((A.m, k), np.float32)
    for i in range(A.m):
        acc = np.zeros(k, np.float32)
        for p in range(A.rp[i], A.rp[i + 1]):
            acc = acc + rx32[p] * X32[A.ri[p]]
        ref32[i] = acc
    assert np.array_equal(b32.to_host(out32), ref32)
