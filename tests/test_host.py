"""Host-side logic: matrix builder conventions, generators, boundary types, and the
C ABI library's load and exports (no compute calls: this suite runs without a GPU)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from sovlp import _lib
from sovlp.core.problem import Solution, Status
from sovlp.core.sparse import SparseMatrix
from sovlp.workloads import (skewed_lp_degrees, small_lp, uneven_batch_member)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_triplets_sum_duplicates_and_drop_zeros():
    # sparse.py:198-236: (0,0) listed twice is summed; an explicit zero disappears
    A = SparseMatrix.from_triplets([0, 0, 1, 1], [0, 0, 1, 2], [1.0, 2.5, 0.0, 4.0], 2, 3)
    assert A.nnz == 2
    np.testing.assert_array_equal(A.to_dense(), [[3.5, 0, 0], [0, 0, 4.0]])
    np.testing.assert_array_equal(A.cp, [0, 1, 1, 2])
    np.testing.assert_array_equal(A.rp, [0, 1, 2])


def test_csr_and_csc_hold_the_same_matrix():
    rng = np.random.default_rng(0)
    r, c = rng.integers(0, 30, 200), rng.integers(0, 40, 200)
    A = SparseMatrix.from_triplets(r, c, rng.standard_normal(200), 30, 40)
    D = np.zeros((30, 40))
    for i in range(30):
        D[i, A.ri[A.rp[i]:A.rp[i + 1]]] = A.rx[A.rp[i]:A.rp[i + 1]]
    np.testing.assert_array_equal(D, A.to_dense())
    # columns ascending inside every row
    for i in range(30):
        assert np.all(np.diff(A.ri[A.rp[i]:A.rp[i + 1]]) > 0)


def test_status_codes_and_objective_drop():
    assert [Status.OPTIMAL, Status.ITERATION_LIMIT, Status.TIME_LIMIT, Status.NUMERICAL] == [0, 4, 5, 8]
    s = Solution(Status.NUMERICAL, objective=3.0, dual_bound=3.0).drop_objective_if_unsolved()
    assert np.isnan(s.objective) and np.isnan(s.dual_bound)
    s = Solution(Status.TIME_LIMIT, objective=3.0).drop_objective_if_unsolved()
    assert s.objective == 3.0


def test_small_lp_shape_and_optimum_certificate():
    ko = small_lp(0)
    p = ko.problem
    assert (p.m, p.n) == (500, 2000) and p.nnz <= 10_000
    eq, _, le, _, _ = p.row_types()
    assert (eq, le) == (250, 250)
    # KKT of (x*, y*) exactly: reduced costs have the right signs at the bounds
    d = p.c - p.A.rmatvec(ko.y_star)
    lo = ko.x_star == 0.0
    hi = ko.x_star == 10.0
    assert np.all(d[lo] > 0) and np.all(d[hi] < 0)
    assert np.allclose(d[~lo & ~hi], 0.0, atol=1e-12)
    assert p.violation(ko.x_star)[0] <= 1e-12


def test_skewed_degrees_follow_the_rank_law():
    deg = skewed_lp_degrees()
    assert abs(int(deg.sum()) - 8_000_000) <= 5000
    assert deg.max() == 5000 and deg.min() == 2
    assert int((deg == 5000).sum()) > 100 and int((deg == 2).sum()) > 800_000


def test_batch_member_ranges():
    for s in range(20):
        ko = uneven_batch_member(5000 + s)
        p = ko.problem
        assert 20 <= p.m <= 800 and 2 * p.m <= p.n <= 4 * p.m
        ub = p.col_ub[0]
        assert 5.0 <= ub <= 50.0 and np.all(p.col_ub == ub)


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "sovlp.h")).read()
    declared = set(re.findall(r"\b(sovlp_[a-z0-9_]+)\s*\(", hdr))
    L = C.CDLL(_lib.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared == set(_lib.exported_symbols())
    assert _lib.lib().sovlp_version() >= 100


def test_tile_plan_host_function():
    rng = np.random.default_rng(1)
    lens = np.concatenate([rng.integers(0, 5, 3000), [5000, 3, 4096, 2049, 0, 7]])
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    out = np.zeros(lens.size + 1, dtype=np.int32)
    nt = C.c_int32()
    _lib.call("sovlp_plan_tiles", lens.size, ptr.ctypes.data, 2048, 1024, 256, out.ctypes.data,
              C.byref(nt))
    t = out[: nt.value + 1]
    assert t[0] == 0 and t[-1] == lens.size and np.all(np.diff(t) > 0)
    for a, b in zip(t[:-1], t[1:]):
        nz = ptr[b] - ptr[a]
        if nz > 2048:
            assert b - a == 1               # a long row is a tile of its own
        else:
            assert b - a <= 1024
            # one lane's serial share of the row sums: at most 64 additions
            # (a multi-row tile never strands a long row on one lane)
            rows, longest = b - a, int((ptr[a + 1:b + 1] - ptr[a:b]).max())
            if rows > 1:
                tpr = 64
                while tpr > 1 and tpr * rows > 256:
                    tpr //= 2
                work = -(-longest // tpr) if rows <= 256 else -(-rows // 256) * longest
                assert work <= 64, (a, b, rows, longest)
    part = np.zeros(9, dtype=np.int32)
    _lib.call("sovlp_partition_tiles", nt.value, t.ctypes.data, ptr.ctypes.data, 8, part.ctypes.data)
    assert part[0] == 0 and part[-1] == nt.value and np.all(np.diff(part) >= 0)


def test_backend_refuses_cpu_and_missing_gpu():
    from sovlp.core import backend
    with pytest.raises(ValueError):
        backend.Backend("cpu")
    if not backend.gpu_available():
        with pytest.raises(RuntimeError):
            backend.get_backend("gpu")


def test_column_panel_split_reproduces_the_matrix():
    from sovlp.spmv_bench import split_panels
    rng = np.random.default_rng(4)
    A = SparseMatrix.from_triplets(rng.integers(0, 40, 600), rng.integers(0, 97, 600),
                                   rng.standard_normal(600), 40, 97)
    for panels in (1, 3, 8):
        ptrs, idx, val = split_panels(A.rp, A.ri, A.rx, A.n, panels)
        w = -(-A.n // panels)
        D = np.zeros((A.m, A.n))
        for p in range(panels):
            for r in range(A.m):
                cols = idx[ptrs[p, r]:ptrs[p, r + 1]]
                assert np.all((cols // w) == p) and np.all(np.diff(cols) > 0)
                D[r, cols] += val[ptrs[p, r]:ptrs[p, r + 1]]
        np.testing.assert_array_equal(D, A.to_dense())
        assert ptrs[-1, -1] == A.nnz


def test_locality_order_relabels_without_changing_any_sum():
    from sovlp.core.sparse import locality_order, relabel_columns
    from sovlp.workloads import skewed_lp_degrees
    rng = np.random.default_rng(3)
    m, n = 300, 2000
    deg = skewed_lp_degrees(n, 12_000, dmax=200)[rng.permutation(n)]
    cols = np.repeat(np.arange(n), deg)
    A = SparseMatrix.from_triplets(rng.integers(0, m, cols.size), cols,
                                   rng.standard_normal(cols.size), m, n)
    new = locality_order(A, hot_degree=8)
    assert np.array_equal(np.sort(new), np.arange(n))
    cnt = np.diff(A.cp)
    by_new = cnt[np.argsort(new)]
    hot = by_new >= 8
    assert hot[: hot.sum()].all() and np.all(np.diff(by_new[: hot.sum()]) <= 0)
    first = A.ci[A.cp[:-1]][np.argsort(new)][hot.sum():]     # cold: by first row
    assert np.all(np.diff(first[by_new[hot.sum():] > 0]) >= 0)
    B = relabel_columns(A, new)
    D = A.to_dense()
    Dn = np.zeros_like(D)
    Dn[:, new] = D
    np.testing.assert_array_equal(B.to_dense(), Dn)
    np.testing.assert_array_equal(B.rp, A.rp)
    np.testing.assert_array_equal(B.rx, A.rx)              # rows keep their entry order
    x = rng.standard_normal(n)
    xn = np.empty(n)
    xn[new] = x
    for i in range(0, m, 7):                                 # row sums in storage order
        a = 0.0
        b = 0.0
        for k in range(A.rp[i], A.rp[i + 1]):
            a += A.rx[k] * x[A.ri[k]]
            b += B.rx[k] * xn[B.ri[k]]
        assert a == b
    for j in range(0, n, 37):                                # columns move whole
        q = new[j]
        np.testing.assert_array_equal(B.ci[B.cp[q]:B.cp[q + 1]], A.ci[A.cp[j]:A.cp[j + 1]])
        np.testing.assert_array_equal(B.cx[B.cp[q]:B.cp[q + 1]], A.cx[A.cp[j]:A.cp[j + 1]])
    with pytest.raises(ValueError, match="permutation"):
        relabel_columns(A, np.zeros(n, dtype=np.int64))


def test_large_lp_defaults():
    """LPs of LARGE_NNZ nonzeros or more default to the locality column order
    and refinement rounds at 1e-3; smaller ones to the given order and 3e-4;
    explicit settings win."""
    from sovlp.lp.pdlp import LARGE_NNZ, PDLPParams
    p = PDLPParams(eps_abs=1e-6, eps_rel=1e-6)
    assert p.resolved_column_order(LARGE_NNZ - 1) == "given"
    assert p.resolved_column_order(LARGE_NNZ) == "locality"
    assert p.resolved_refine(32) == (6, 3e-4)
    assert p.resolved_refine(32, LARGE_NNZ - 1) == (6, 3e-4)
    assert p.resolved_refine(32, LARGE_NNZ) == (6, 1e-3)
    assert p.resolved_refine(64, LARGE_NNZ) == (0, 1e-3)
    q = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, column_order="given", refine_eps=2e-4)
    assert q.resolved_column_order(LARGE_NNZ) == "given"
    assert q.resolved_refine(32, LARGE_NNZ) == (6, 2e-4)
    assert PDLPParams(eps_rel=5e-3).resolved_refine(32, LARGE_NNZ)[1] == 5e-3
    assert q.to_c(32, LARGE_NNZ).refine_eps == 2e-4
    assert p.to_c(32, LARGE_NNZ).refine_eps == 1e-3


def test_column_order_is_checked():
    from sovlp.lp.pdlp import PDLPParams, solve_pdlp
    from sovlp.workloads import small_lp
    with pytest.raises(ValueError, match="column_order"):
        solve_pdlp(small_lp(0).problem, PDLPParams(column_order="sorted"))


def test_bench_reports_traffic_only_for_the_measured_configuration():
    """bench.py reads the committed PMC traffic only for the workload, kernel and
    replays per step it was measured on."""
    import json
    import bench
    from sovlp.lp.pdlp import PDLPParams
    d = json.load(open(os.path.join(bench.ROOT, "profiles",
                                    "pmc_traffic_pdlp_run_kernel_fp32_4096_lps.json")))
    par = PDLPParams(refine_stall_open=1024)
    refine = bench.refinement_config(par, 32)
    args = (d["kernel"], 4096, 1000, 32, 256, True, refine)
    assert bench.measured_traffic(*args, d["replays_per_step"]) == d["traffic_bytes_per_launch"]
    assert bench.measured_traffic(*args, d["replays_per_step"] + 25) is None
    assert bench.measured_traffic(d["kernel"], 4096, 1000, 32, 256, False, refine,
                                  d["replays_per_step"]) is None
    assert bench.measured_traffic(*args[:6], bench.refinement_config(PDLPParams(), 32),
                                  d["replays_per_step"]) is None
    assert bench.measured_traffic(*args, d["replays_per_step"], False) is None
    assert d["workload"] == bench.traffic_key(4096, 1000, 32, 256, True, refine, True)
    assert 0.5 < d["traffic_bytes_per_launch"] / d["algorithmic_bytes_per_launch"] < 1.5


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(os.path.dirname(__file__), "..", "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bench_spmv_traffic_record_matches_the_section_keys():
    """The committed SpMV counter record holds the eight tiled products of the
    bench's spmv section, keyed as bench.py looks them up, with counted traffic
    of at least the bytes of the matrix itself."""
    import json
    import bench
    d = json.load(open(os.path.join(bench.ROOT, "profiles",
                                    "pmc_traffic_spmv_tiled_200k_x_1M.json")))
    rec = bench.spmv_traffic()
    assert len(rec) == len(d["products"]) == 8
    for e in d["products"]:
        key = tuple(e[k] for k in bench.SPMV_KEYS)
        assert rec[key] == e["traffic_bytes_per_launch"]
        s_val = 8 if e["precision"] == 64 else 4
        assert e["traffic_bytes_per_launch"] >= e["nnz"] * (s_val + 4)
